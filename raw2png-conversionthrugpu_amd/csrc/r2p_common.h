// r2p_common.h -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "r2p_internal.h"

namespace r2p {

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define R2P_HIP(call)                                              \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return ::r2p::hip_fail(_e, #call);   \
    } while (0)

struct LaunchStatus;

// Launch wrappers (r2p_kernels.hip).
struct FilterLaunch;
struct DeflateLaunch;
struct AssembleLaunch;
hipError_t launch_build_lut(const LutBuild& lb, void* out, hipStream_t s);
hipError_t launch_filter(const FilterLaunch& L, int grid, size_t lds, hipStream_t s);
// Rows wider than kMaxRowBytesLds: one row per workgroup, in tiles.
hipError_t launch_filter_tiled(const FilterLaunch& L, int grid, size_t lds, hipStream_t s);
hipError_t launch_deflate(const DeflateLaunch& DL, int lane_bytes, int wg_threads, int grid, hipStream_t s);
hipError_t launch_scan(const AssembleLaunch& AL, int njobs, hipStream_t s);
// Dynamic-Huffman tables per segment, then the slices' bits.
hipError_t launch_huffman(const HuffLaunch& HL, int grid_tables, int grid_emit, hipStream_t s);
hipError_t launch_assemble(const AssembleLaunch& AL, int grid, hipStream_t s);
// A single job small enough for k_assemble to scan it itself (no k_scan
// launch; the scan arrays stay in LDS): launch_assemble_self_scan.
bool assemble_scans_itself(const AssembleJob& A);
hipError_t launch_assemble_self_scan(const AssembleLaunch& AL, int grid, hipStream_t s);
hipError_t launch_adler(const uint8_t* p, int64_t n, uint32_t* part, uint32_t* result, hipStream_t s);
hipError_t launch_crc(const uint8_t* p, int64_t n, uint32_t* part, uint32_t* result, hipStream_t s);
hipError_t launch_put_bytes(uint8_t* dst, const uint8_t* bytes16, int n, const uint32_t* crc_be_src,
                            hipStream_t s);
hipError_t upload_tables(const uint32_t* lit, const uint32_t* len, const uint32_t* dist_rev,
                         const uint32_t* crc4, const uint32_t* x8pow, const uint32_t* crc_shift,
                         const uint32_t* crc_shift32);

// Image geometry resolved from an r2p_image.
struct Geometry {
    int out_bits = 0;
    int bpp = 0;            // PNG bytes per complete pixel
    size_t row_bytes = 0;   // PNG-ready bytes per row
    size_t filtered = 0;    // height * (row_bytes + 1)
    size_t in_row = 0;      // raw bytes per row (packed)
    int64_t pitch = 0;      // raw bytes between rows
    int color_type = 0;
};
int resolve(const r2p_image* img, Geometry* g);

// Window-stage setup for one image: mode, table requirements.
struct WindowPlan {
    int32_t wmode = WM_COPY;
    int32_t shift = 0;
    uint32_t mask = 0xFFFFu;
    int32_t q15_a = 0, q15_b = 0, pixel_rep = 0;
    int lut_kind = -1;      // -1 none, 0 a2 float window, 1 a4 8-bit LUT
    int lut_entries = 0;
    int lut_bytes = 0;      // bytes of the device table
    int lds_bytes = 0;      // bytes of it the filter stages in LDS (FilterJob::lut_bytes)
};
int plan_window(const r2p_image* img, WindowPlan* wp);

// Host restatement of try_apply_linear_window's coefficients
// (batch_processor.cpp:202-248).  Returns true when the Q15 path applies.
bool q15_coeffs(const r2p_dicom_params& p, int16_t* a, int32_t* b);

size_t png_bound_bytes(size_t filtered, int segment_bytes, int chunk_bytes);
// CRC-32 of a PNG's IHDR chunk (its type and 13 data bytes).
uint32_t ihdr_crc(uint32_t width, uint32_t height, uint8_t bit_depth, uint8_t color_type);
// Filter bands: LDS bytes of a band (the band's rows and the row above it,
// beside the table), and the band height to use.
size_t filter_lds_bytes(int band_rows, int row_cap, int lut_cap);
size_t filter_tiled_lds_bytes(int lut_cap);
int filter_band_rows(int requested, int64_t total_rows, int row_cap, int lut_cap);
// Compute units of the current device (after r2p_init; 256 when unknown).
int cu_count();
// Workgroups of a filter launch (k_filter or k_filter_split by the row width).
int filter_grid(const FilterLaunch& L);
void fill_defaults(r2p_options* o);
// DEFLATE slice size for a stream of `filtered` bytes: the options' value,
// or (0, auto) 16, 8 or 4 KiB -- the largest that still gives about 512
// slices, two per CU: small images get more, shorter workgroups.
int resolve_slice_bytes(const r2p_options& o, size_t filtered);
// Slice dwords of one slice of `slice_bytes` (the dynamic-Huffman bound with huffman).
int slice_cap_for(int slice_bytes, int huffman);
// Slice buffers any stream of up to `filtered` bytes needs with these
// options: the most slices, and the most dwords of slice output, over the
// sizes resolve_slice_bytes can pick.
void slice_arena_bound(const r2p_options& o, size_t filtered, size_t* max_slices, size_t* max_dwords);

// Cached device tables keyed by DICOM parameter set.
struct LutCache {
    void* d_lut = nullptr;
    size_t cap = 0;
    bool valid = false;
    int kind = -1, entries = 0;
    r2p_dicom_params p{};
    ~LutCache();
    // Ensures the table for `wp`/`p` is on the device (built on `s`).
    int ensure(const WindowPlan& wp, const r2p_dicom_params& p, hipStream_t s);
};

}  // namespace r2p
