// gpu_png_assemble.h -- one IDAT chunk from a device slice of DEFLATE output.
// Drop-in for include/gpu_png_assemble.h:38-84 of the reference.  The chunk
// CRC-32 is computed on the device, so the copy to host is a plain D2H.
#pragma once
#include <cstddef>
#include <cstdint>

struct GpuPngAssembleContext;

GpuPngAssembleContext* gpu_png_assemble_create(size_t max_chunk_data_bytes);
void                   gpu_png_assemble_destroy(GpuPngAssembleContext* ctx);

// [len BE]["IDAT"][78 FLG if is_first][slice][adler BE if is_last][crc BE].
// Returns the chunk's total size; asynchronous until copy_to_host.
size_t gpu_png_assemble_idat_chunk(GpuPngAssembleContext* ctx, const uint8_t* d_compressed_slice,
                                   size_t slice_bytes, bool is_first, bool is_last,
                                   uint32_t running_adler, int zlib_level);

// Blocking copy of the most recent chunk (num_bytes = idat_chunk's result).
void gpu_png_assemble_copy_to_host(const GpuPngAssembleContext* ctx, uint8_t* h_dst,
                                   size_t num_bytes);

struct GpuPngAssembleStats {
    long long chunks        = 0;
    long long header_h2d_us = 0;
    long long data_d2d_us   = 0;
    long long presync_us    = 0;
    long long crc_us        = 0;
    long long crc_h2d_us    = 0;
    long long copy_d2h_us   = 0;
};
GpuPngAssembleStats gpu_png_assemble_get_stats(const GpuPngAssembleContext* ctx);
void                gpu_png_assemble_reset_stats(GpuPngAssembleContext* ctx);
