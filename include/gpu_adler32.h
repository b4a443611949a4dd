// gpu_adler32.h -- Adler-32 of device-resident bytes.
// Drop-in for include/gpu_adler32.h:24-44 of the reference.  Partials per
// 64 KiB block are combined on the device (no host combine loop).
#pragma once
#include <cstddef>
#include <cstdint>

struct GpuAdler32Context;

GpuAdler32Context* gpu_adler32_create(size_t max_input_bytes, size_t chunk_bytes = 65536);
void               gpu_adler32_destroy(GpuAdler32Context* ctx);

// launch + collect
uint32_t gpu_adler32_compute(GpuAdler32Context* ctx, const uint8_t* d_data, size_t len);

// Asynchronous form: launch queues the kernels; collect waits and returns the
// value.  Exactly one collect per launch.
void     gpu_adler32_launch(GpuAdler32Context* ctx, const uint8_t* d_data, size_t len);
uint32_t gpu_adler32_collect(GpuAdler32Context* ctx);
