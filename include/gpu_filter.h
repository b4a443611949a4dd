// gpu_filter.h -- window + adaptive PNG filter on the GPU, strip by strip.
// Drop-in for include/gpu_filter.h:8-90 of the reference (same functions and
// argument meaning); implemented over r2p_filter_rows (r2p.h).
#pragma once
#include <cstddef>
#include <cstdint>

#include "dicom_params.h"

struct GpuFilterContext;

// Device timing of one call, milliseconds (gpu_filter.h:12-16).
struct GpuTimings {
    float h2d_ms    = 0.f;
    float kernel_ms = 0.f;
    float d2h_ms    = 0.f;
};

// width pixels of bpp bytes, strips of up to strip_height rows.  With
// device_output_only no pinned host mirror is allocated (gpu_filter.h:22-23).
GpuFilterContext* gpu_filter_create(int width, int bpp, int strip_height,
                                    bool device_output_only = false);
void gpu_filter_destroy(GpuFilterContext* ctx);

// Pinned host buffers (gpu_filter.h:29-30).
uint8_t* gpu_filter_alloc_pinned(size_t bytes);
void     gpu_filter_free_pinned(uint8_t* ptr);

// Zero the carried prior row before the first strip of a new image (:35).
void gpu_filter_reset(GpuFilterContext* ctx);

// Strip already in device memory (:41-47).  d_prior_row (device, optional)
// overrides the carried row above the strip.  dicom != nullptr windows the raw
// little-endian samples first (dicom_preprocess_kernel semantics).
const uint8_t* gpu_filter_process_from_device(GpuFilterContext* ctx, const uint8_t* d_input,
                                              const uint8_t* d_prior_row, int actual_rows,
                                              GpuTimings* timings = nullptr,
                                              const DicomPixelParams* dicom = nullptr);

// Strip in host memory (:56-63).  Returns the pinned host copy of the filtered
// strip, or nullptr when device_output_only (then read gpu_filter_device_output).
const uint8_t* gpu_filter_process_from_host(GpuFilterContext* ctx, const uint8_t* h_input,
                                            const uint8_t* h_prior_row, int actual_rows,
                                            GpuTimings* timings = nullptr,
                                            const DicomPixelParams* dicom = nullptr,
                                            bool device_output_only = false);

// actual_rows * (width * bpp + 1) (:67).
size_t gpu_filter_output_size(const GpuFilterContext* ctx, int actual_rows);

// Device pointer to the last filtered strip (:73).  Stream-ordered: work the
// caller enqueues after gpu_filter_stream() (or after a sync) sees it complete.
const uint8_t* gpu_filter_device_output(const GpuFilterContext* ctx);

// Copies the carried (windowed) last row to host memory, blocking (:90).
void gpu_filter_copy_prior_row_to_host(const GpuFilterContext* ctx, uint8_t* h_dst);

// The context's HIP stream (hipStream_t as void*), so consumers can order
// their work after the filter without a host sync.  No reference counterpart:
// it replaces the reference's implicit cross-stream hand-off
// (pipeline.cpp:618-623, 689-698), which races on the last strip.
void* gpu_filter_stream(const GpuFilterContext* ctx);
