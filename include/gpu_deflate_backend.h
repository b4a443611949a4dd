// gpu_deflate_backend.h -- strip-incremental DEFLATE on the GPU.
// Drop-in for include/gpu_deflate_backend.h:40-127 of the reference.
//
// Semantics kept: compress_strip is asynchronous, stream_sync must precede
// reading the output, raw DEFLATE only (zlib header/trailer by the caller).
// Semantics upgraded: every non-final strip ends byte-aligned (sync flush), so
// gpu_deflate_flushable_byte_length() == gpu_deflate_output_byte_length()
// after each strip; strips are cut into 64 KiB segments compressed in
// parallel, each primed with the preceding 32 KiB of input.
#pragma once
#include <cstddef>
#include <cstdint>

struct GpuDeflateContext;

struct LzStats {
    uint64_t matches_found;
    uint64_t matched_bytes;
    uint32_t max_match_len;
    double   avg_match_len;
    double   avg_match_dist;
    uint64_t total_positions;
    double   coverage;
};

// max_rows x max_row_bytes bounds one strip; max_total_bits bounds the whole
// image's output (total_filtered*9 + strips*10 suffices, as in the reference).
GpuDeflateContext* gpu_deflate_create(int max_rows, int max_row_bytes, size_t max_total_bits,
                                      bool use_lz77 = false);
void gpu_deflate_destroy(GpuDeflateContext* ctx);
void gpu_deflate_reset(GpuDeflateContext* ctx);

// d_filtered: actual_rows x row_bytes device bytes (filter byte + row).
void gpu_deflate_compress_strip(GpuDeflateContext* ctx, const uint8_t* d_filtered, int actual_rows,
                                int row_bytes, bool is_last);
void gpu_deflate_stream_sync(GpuDeflateContext* ctx);

const uint8_t* gpu_deflate_output(const GpuDeflateContext* ctx);
size_t gpu_deflate_output_byte_length(const GpuDeflateContext* ctx);
size_t gpu_deflate_flushable_byte_length(const GpuDeflateContext* ctx);
void gpu_deflate_copy_to_host(const GpuDeflateContext* ctx, uint8_t* h_dst, size_t num_bytes);

LzStats gpu_deflate_lz_stats(const GpuDeflateContext* ctx);

struct GpuDeflatePhaseStats {
    long long lz77_us      = 0;  // fused into the segment kernel: always 0
    long long bitlen_us    = 0;  // fused: always 0
    long long scan_us      = 0;  // segment offset scan + concatenation
    long long encode_us    = 0;  // segment kernel (LZ77 + Huffman + packing)
    long long sync_us      = 0;  // host wait in gpu_deflate_stream_sync
    long long stats_d2h_us = 0;
};
GpuDeflatePhaseStats gpu_deflate_get_phase_stats(const GpuDeflateContext* ctx);
void                 gpu_deflate_reset_phase_stats(GpuDeflateContext* ctx);
