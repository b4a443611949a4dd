// gpu_crc32.h -- CRC-32 (PNG/zlib polynomial) of device-resident bytes.
// Drop-in for include/gpu_crc32.h:19-38 of the reference.
#pragma once
#include <cstddef>
#include <cstdint>

struct GpuCrc32Context;

GpuCrc32Context* gpu_crc32_create(size_t max_input_bytes, size_t chunk_bytes = 65536);
void             gpu_crc32_destroy(GpuCrc32Context* ctx);

// CRC-32 of d_data[0..len), blocking.
uint32_t gpu_crc32_compute(GpuCrc32Context* ctx, const uint8_t* d_data, size_t len);

// crc(A || B) from crc(A), crc(B) and len(B) (zlib crc32_combine semantics).
uint32_t gpu_crc32_combine(uint32_t crc1, uint32_t crc2, size_t len2);
