// r2p_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the PNG encoder.
//
//   k_build_lut     DICOM pixel tables (a2 float window, a4 8-bit LUT)
//   k_filter        fused window + adaptive PNG filter, rows staged in LDS
//   k_deflate       per-segment LZ77 + fixed-Huffman DEFLATE, 32 KiB primed
//                   window, sync-flush segment ends, per-segment Adler-32
//   k_scan          segment byte offsets, zlib payload size, Adler combine
//   k_assemble      concatenation / IDAT chunking with on-device CRC-32
//   k_adler_*, k_crc_*   standalone checksums (partials + combine)
//
// Everything is integer byte work bound by HBM bandwidth; nothing is a dense
// contraction, so no MFMA.  Wave64 throughout: reductions and scans use
// cross-lane shuffles, bit packing uses per-lane bit offsets from a block scan.
// Float code (the DICOM window) is compiled with -ffp-contract=off and IEEE
// division so it reproduces the oracle bit for bit.
#include <hip/hip_runtime.h>

#include "r2p_internal.h"

namespace r2p {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef u32x4 u32x4_ua __attribute__((aligned(1)));
typedef uint32_t u32_ua __attribute__((aligned(1)));

// ---------------------------------------------------------------------------
// Constant tables (filled by r2p_upload_tables)
// ---------------------------------------------------------------------------
__constant__ uint32_t c_lit[256];      // reversed fixed-Huffman literal code | nbits << 16
__constant__ uint32_t c_len[259];      // length L: (revcode | extra << cbits) | total bits << 16
__constant__ uint32_t c_dist_rev[32];  // 5-bit reversed distance codes
__constant__ uint32_t c_crc[4][256];   // slice-by-4 CRC-32 tables
__constant__ uint32_t c_x8pow[64];     // x^(8 * 2^k) mod P, reflected
__device__ uint32_t c_crcshift[kCrcShiftLevels][4][256];     // c -> c * x^(8*64*2^k) mod P, by input byte
__device__ uint32_t c_crcshift32[4][256];                   // c -> c * x^(8*32) mod P, by input byte

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ld32u(const uint8_t* p) { return *(const u32_ua*)p; }

// Loads from pointers known to address global memory.  The job descriptors
// carry generic pointers; casting to the global address space lets the
// compiler emit global_load (vmcnt only) instead of flat_load, whose
// completion is also tracked by lgkmcnt and would serialise it against the
// LDS traffic it is interleaved with.
#define R2P_GLOBAL __attribute__((address_space(1)))
typedef R2P_GLOBAL const u32_ua gu32_ua;
typedef R2P_GLOBAL const u32x4_ua gu32x4_ua;
typedef R2P_GLOBAL const uint8_t gu8;
__device__ __forceinline__ uint32_t g32u(const uint8_t* p) { return *(gu32_ua*)p; }
__device__ __forceinline__ u32x4 g128u(const uint8_t* p) { return *(gu32x4_ua*)p; }
__device__ __forceinline__ uint32_t g8(const uint8_t* p) { return *(gu8*)p; }
typedef uint64_t u64_ua __attribute__((aligned(1)));
typedef R2P_GLOBAL const u64_ua gu64_ua;
__device__ __forceinline__ uint64_t g64u(const uint8_t* p) { return *(gu64_ua*)p; }
__device__ __forceinline__ int64_t min64(int64_t a, int64_t b) { return a < b ? a : b; }
__device__ __forceinline__ int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

// Byte-wise a - b (mod 256 in every byte), no borrow between bytes.
__device__ __forceinline__ uint32_t sub8(uint32_t x, uint32_t y)
{
    return ((x | 0x80808080u) - (y & 0x7F7F7F7Fu)) ^ ((x ^ ~y) & 0x80808080u);
}
// Byte-wise floor((a + b) / 2).
__device__ __forceinline__ uint32_t avg8(uint32_t a, uint32_t b)
{
    return (a & b) + (((a ^ b) >> 1) & 0x7F7F7F7Fu);
}
// Sum over the 4 bytes of |int8(v)| = min(v, 256 - v), added to acc.
__device__ __forceinline__ uint32_t cost8(uint32_t v, uint32_t acc)
{
    return __builtin_amdgcn_sad_u8(v ^ 0x80808080u, 0x80808080u, acc);
}
typedef unsigned short us2 __attribute__((ext_vector_type(2)));
typedef short ss2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
// Inclusive prefix sum across the 64 lanes of a wave (DPP: row shifts by
// 1, 2, 4 and 8 scan each row of 16 lanes, then the row broadcasts of lanes
// 15 and 31 carry the rows' totals forward).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);   // row_bcast:31
    return v;
}

// clamp((x*A + B) >> 15, 0, 255) with int32 wrap-around (batch_processor.cpp:250-263).
// The clamp is applied to the Q15 accumulator before the shift (identical
// result: acc < 0 -> 0, acc >= 256<<15 -> 255), which keeps the compiler from
// fusing shift+saturate+pack into v_ashr_pk_u8_i32 -- that fused form gave
// results different from the scalar definition on gfx950.
__device__ __forceinline__ uint32_t q15_clamp(int32_t x, int32_t a, int32_t b)
{
    const int32_t acc = (int32_t)((uint32_t)x * (uint32_t)a + (uint32_t)b);
    return (uint32_t)min(max(acc, 0), (256 << 15) - 1) >> 15;
}

// float -> int32 conversion with x86 cvttss2si semantics (out of range or
// NaN gives INT32_MIN), as the reference's host-side LUT builder executes it.
__device__ __forceinline__ int32_t x86_f2i(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : (int32_t)0x80000000;
}

// ---------------------------------------------------------------------------
// DICOM pixel arithmetic (a2: gpu_filter.cu:71-162; a4: batch_processor.cpp:100-147)
// ---------------------------------------------------------------------------
// Value of a post-shift/mask stored code (bits_allocated 16) or raw byte (8).
__device__ float dicom_code_value(uint32_t code, const r2p_dicom_params& p)
{
    float val;
    if (p.bits_allocated == 16) {
        if (p.pixel_rep == 1) {
            int32_t sv;
            if (p.bits_stored == 16) {
                sv = (int32_t)(int16_t)(uint16_t)code;
            } else {
                const int half = 1 << (p.bits_stored - 1);
                sv = (int32_t)code;
                if (sv >= half) sv -= (1 << p.bits_stored);
            }
            if (p.apply_rescale) {
                const float m = (float)sv * p.rescale_slope;
                val = m + p.rescale_intercept;
            } else {
                val = (float)sv;
            }
        } else {
            val = (float)code;
            if (p.apply_rescale) {
                const float m = val * p.rescale_slope;
                val = m + p.rescale_intercept;
            }
        }
    } else {
        val = p.pixel_rep == 1 ? (float)(int8_t)(uint8_t)code : (float)(uint8_t)code;
        if (p.apply_rescale) {
            const float m = val * p.rescale_slope;
            val = m + p.rescale_intercept;
        }
    }
    return val;
}

__device__ uint32_t dicom_window(float val, const r2p_dicom_params& p)
{
    const int out_max = (p.bits_allocated == 16) ? 65535 : 255;
    if (p.apply_window) {
        const float wm1 = p.window_width - 1.0f;
        const float cen = p.window_center;
        if (wm1 <= 0.0f) {
            val = (val >= cen) ? (float)out_max : 0.0f;
        } else {
            const float half = wm1 * 0.5f;
            const float lo = (cen - half) - 0.5f;
            const float hi = (cen + half) - 0.5f;
            if (val <= lo) {
                val = 0.0f;
            } else if (val > hi) {
                val = (float)out_max;
            } else {
                float t = val - cen;
                t = t + 0.5f;
                t = t / wm1;
                t = t + 0.5f;
                val = t * (float)out_max;
            }
        }
    }
    val = fmaxf(0.f, fminf((float)out_max, val));
    return (uint32_t)val;
}

__device__ uint32_t lut8_entry(int i, const r2p_dicom_params& p, int lut_size)
{
    const int mask = lut_size - 1;
    float norm_scale = 1.f, norm_offset = 0.f;
    if (!p.apply_window) {
        if (p.pixel_rep == 1) {
            const float half = (float)(lut_size >> 1);
            const float range = (float)lut_size;
            norm_scale = 1.f / range;
            norm_offset = half / range;
        } else {
            norm_scale = 1.f / (float)(lut_size - 1);
        }
    }
    float lo = 0.f, inv_range = 1.f;
    if (p.apply_window) {
        lo = p.window_center - p.window_width * 0.5f;
        const float hi = p.window_center + p.window_width * 0.5f;
        inv_range = (hi > lo) ? 1.f / (hi - lo) : 0.f;
    }
    float val;
    if (p.pixel_rep == 1 && p.bits_stored < 16) {
        const int sign_bit = lut_size >> 1;
        int32_t s = i;
        if (s & sign_bit) s |= ~mask;
        val = (float)s;
    } else if (p.pixel_rep == 1) {
        val = (float)(int16_t)(uint16_t)(unsigned)i;
    } else {
        val = (float)(unsigned)i;
    }
    if (p.apply_rescale) {
        const float m = val * p.rescale_slope;
        val = m + p.rescale_intercept;
    }
    float norm;
    if (p.apply_window) {
        norm = (val - lo) * inv_range;
    } else {
        const float m = val * norm_scale;
        norm = m + norm_offset;
    }
    const float s255 = norm * 255.f;
    const int v = x86_f2i(s255 + 0.5f);
    return (uint32_t)(v < 0 ? 0 : v > 255 ? 255 : v);
}

__global__ void k_build_lut(LutBuild lb, void* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= lb.entries) return;
    if (lb.kind == 0) {
        const uint32_t v = dicom_window(dicom_code_value((uint32_t)i, lb.p), lb.p);
        if (lb.p.bits_allocated == 16) ((uint16_t*)out)[i] = (uint16_t)v;
        else                           ((uint8_t*)out)[i] = (uint8_t)v;
    } else {
        ((uint8_t*)out)[i] = (uint8_t)lut8_entry(i, lb.p, lb.entries);
    }
}

// The even bytes (0, 2) and the odd bytes (1, 3) of a dword as two 16-bit lanes.
__device__ __forceinline__ us2 lanes_even(uint32_t v) { return __builtin_bit_cast(us2, v & 0x00FF00FFu); }
__device__ __forceinline__ us2 lanes_odd(uint32_t v) { return __builtin_bit_cast(us2, __builtin_amdgcn_perm(0u, v, 0x0C030C01u)); }
// The low bytes of the lanes of ev and od, back in byte order (ev0 od0 ev1 od1).
__device__ __forceinline__ uint32_t join_lanes(us2 ev, us2 od)
{
    return __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, od), __builtin_bit_cast(uint32_t, ev), 0x06020400u);
}
// Paeth predictor (PNG spec 9.4) on two 16-bit lanes of bytes:
//   p - a = b - c, p - b = a - c, p - c = (b - c) + (a - c)
//   a if |p-a| <= min(|p-b|, |p-c|), else b if |p-b| <= |p-c|, else c.
__device__ __forceinline__ us2 paeth_pred(us2 a, us2 b, us2 c)
{
    const ss2 da = __builtin_bit_cast(ss2, (us2)(b - c));
    const ss2 db = __builtin_bit_cast(ss2, (us2)(a - c));
    const ss2 dc = da + db;
    const ss2 pa = __builtin_elementwise_max(da, (ss2)(-da));
    const ss2 pb = __builtin_elementwise_max(db, (ss2)(-db));
    const ss2 pc = __builtin_elementwise_max(dc, (ss2)(-dc));
    uint32_t na = __builtin_bit_cast(uint32_t, (ss2)((__builtin_elementwise_min(pb, pc) - pa) >> 15));
    uint32_t tc = __builtin_bit_cast(uint32_t, (ss2)((pc - pb) >> 15));
    // the lane masks stay masks (v_bfi_b32): as selects the compiler splits
    // them into per-lane compares
    asm("" : "+v"(na), "+v"(tc));
    const uint32_t a32 = __builtin_bit_cast(uint32_t, a), b32 = __builtin_bit_cast(uint32_t, b);
    const uint32_t bc = (__builtin_bit_cast(uint32_t, c) & tc) | (b32 & ~tc);
    return __builtin_bit_cast(us2, (bc & na) | (a32 & ~na));
}
// Costs of the five filters on one dword (x; a, b, c: its left, up and
// upper-left neighbours) added to s[0..4], the bytes outside m not counted;
// returns the Paeth-filtered dword.
__device__ __forceinline__ uint32_t cost_dword(uint32_t x, uint32_t a, uint32_t b, uint32_t c, uint32_t m, uint32_t* s)
{
    const us2 xe = lanes_even(x), xo = lanes_odd(x), ae = lanes_even(a), ao = lanes_odd(a);
    const us2 be = lanes_even(b), bo = lanes_odd(b), ce = lanes_even(c), co = lanes_odd(c);
    const uint32_t fp = join_lanes(xe - paeth_pred(ae, be, ce), xo - paeth_pred(ao, bo, co));
    s[0] = cost8(x & m, s[0]);
    s[1] = cost8(join_lanes(xe - ae, xo - ao) & m, s[1]);
    s[2] = cost8(join_lanes(xe - be, xo - bo) & m, s[2]);
    s[3] = cost8(join_lanes(xe - ((ae + be) >> 1), xo - ((ao + bo) >> 1)) & m, s[3]);
    s[4] = cost8(fp & m, s[4]);
    return fp;
}

// ---------------------------------------------------------------------------
// k_filter: fused window + adaptive filter (dicom_preprocess_kernel and
// filter_select_kernel, gpu_filter.cu:71-309), for rows of up to 6 KiB.
//
// A workgroup of 4 waves owns a band of consecutive rows (one per wave by
// default).  The band and the row above it are windowed from the raw input
// into PNG-ready rows in LDS (raw loads two or four items in flight), then
// each wave filters whole rows alone: lane l scores items l, l + 64, ... (K
// of them) for the five PNG filters (the even and odd bytes of each dword as
// packed 16-bit lanes, v_sad_u8 for the |int8| cost), five DPP wave
// reductions pick the filter (ties to the lowest index, as
// gpu_filter.cu:267-273), and the wave writes [filter byte | chosen row],
// the Paeth results kept in registers from the scoring pass.
// ---------------------------------------------------------------------------


// Raw input of 16-byte output item i of a row: 16 bytes, or 32 for the modes
// that narrow 16-bit samples to 8 bits.
struct RawItem { u32x4 v0, v1; };
__device__ __forceinline__ bool narrows(int wmode) { return wmode == WM_Q15 || wmode == WM_LUT8_FROM16; }
__device__ __forceinline__ RawItem load_raw(const FilterJob& J, const uint8_t* src_row, int i)
{
    RawItem r;
    if (narrows(J.wmode)) { r.v0 = g128u(src_row + 32 * i); r.v1 = g128u(src_row + 32 * i + 16); }
    else { r.v0 = g128u(src_row + 16 * i); r.v1 = r.v0; }
    return r;
}

// PNG-ready bytes of a full 16-byte item from its raw input.
__device__ __forceinline__ u32x4 window_raw(const FilterJob& J, const RawItem& R, const uint8_t* lut_lds)
{
    u32x4 o;
    switch (J.wmode) {
    case WM_COPY: {
        o = R.v0;
        break;
    }
    case WM_SWAP16: {
        const u32x4 v = R.v0;
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = ((v[k] & 0x00FF00FFu) << 8) | ((v[k] >> 8) & 0x00FF00FFu);
        break;
    }
    case WM_LUT16: {
        const u32x4 v = R.v0;
        const uint16_t* t16 = (const uint16_t*)lut_lds;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t s0 = t16[((v[k] & 0xFFFFu) >> J.shift) & J.mask];
            const uint32_t s1 = t16[((v[k] >> 16) >> J.shift) & J.mask];
            const uint32_t w = s0 | (s1 << 16);
            o[k] = ((w & 0x00FF00FFu) << 8) | ((w >> 8) & 0x00FF00FFu);
        }
        break;
    }
    case WM_LUT16G: {
        // the 128 KiB table of 16-bit stored samples, gathered through the caches
        const u32x4 v = R.v0;
        const R2P_GLOBAL uint16_t* t16 = (const R2P_GLOBAL uint16_t*)J.lut;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t s0 = t16[((v[k] & 0xFFFFu) >> J.shift) & J.mask];
            const uint32_t s1 = t16[((v[k] >> 16) >> J.shift) & J.mask];
            const uint32_t w = s0 | (s1 << 16);
            o[k] = ((w & 0x00FF00FFu) << 8) | ((w >> 8) & 0x00FF00FFu);
        }
        break;
    }
    case WM_LUT8_FROM8: {
        const u32x4 v = R.v0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            o[k] = (uint32_t)lut_lds[v[k] & 0xFF] | ((uint32_t)lut_lds[(v[k] >> 8) & 0xFF] << 8) |
                   ((uint32_t)lut_lds[(v[k] >> 16) & 0xFF] << 16) | ((uint32_t)lut_lds[v[k] >> 24] << 24);
        }
        break;
    }
    case WM_Q15: {
        const u32x4 v0 = R.v0, v1 = R.v1;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t w0 = k < 2 ? v0[2 * k] : v1[2 * k - 4];
            const uint32_t w1 = k < 2 ? v0[2 * k + 1] : v1[2 * k - 3];
            uint32_t r = 0;
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const uint32_t raw = ((h < 2 ? w0 : w1) >> (16 * (h & 1))) & 0xFFFFu;
                const int32_t x = J.pixel_rep == 1 ? (int32_t)(int16_t)raw : (int32_t)raw;
                r |= q15_clamp(x, J.q15_a, J.q15_b) << (8 * h);
            }
            o[k] = r;
        }
        break;
    }
    default: {  // WM_LUT8_FROM16
        const u32x4 v0 = R.v0, v1 = R.v1;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t w0 = k < 2 ? v0[2 * k] : v1[2 * k - 4];
            const uint32_t w1 = k < 2 ? v0[2 * k + 1] : v1[2 * k - 3];
            uint32_t r = 0;
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const uint32_t raw = ((h < 2 ? w0 : w1) >> (16 * (h & 1))) & 0xFFFFu;
                r |= (uint32_t)lut_lds[(raw >> J.shift) & J.mask] << (8 * h);
            }
            o[k] = r;
        }
        break;
    }
    }
    return o;
}

// PNG-ready byte j of a raw row (the tail of a row whose length is not a
// multiple of 16).
__device__ __forceinline__ uint32_t window_byte(const FilterJob& J, const uint8_t* src_row, int j, const uint8_t* lut_lds)
{
    switch (J.wmode) {
    case WM_COPY: return src_row[j];
    case WM_SWAP16: return src_row[j ^ 1];
    case WM_LUT16:
    case WM_LUT16G: {
        const int s = j >> 1;
        const uint32_t raw = (uint32_t)src_row[2 * s] | ((uint32_t)src_row[2 * s + 1] << 8);
        const uint16_t* t16 = J.wmode == WM_LUT16 ? (const uint16_t*)lut_lds : (const uint16_t*)J.lut;
        const uint32_t w = t16[(raw >> J.shift) & J.mask];
        return (j & 1) ? (w & 0xFF) : (w >> 8);
    }
    case WM_LUT8_FROM8: return lut_lds[src_row[j]];
    case WM_Q15: {
        const uint32_t raw = (uint32_t)src_row[2 * j] | ((uint32_t)src_row[2 * j + 1] << 8);
        const int32_t x = J.pixel_rep == 1 ? (int32_t)(int16_t)raw : (int32_t)raw;
        return q15_clamp(x, J.q15_a, J.q15_b);
    }
    default: {
        const uint32_t raw = (uint32_t)src_row[2 * j] | ((uint32_t)src_row[2 * j + 1] << 8);
        return lut_lds[(raw >> J.shift) & J.mask];
    }
    }
}

// Work item (DEFLATE slice, filter band) of workgroup b: workgroups are dealt
// to the 8 XCDs in turn (b on XCD b % 8), so consecutive b run on different
// L2s.  Mapping b to items in runs of kXcdRun consecutive items per XCD puts
// most of what an item reads of its neighbours (a slice's 32 KiB window, a
// band's row above) in the L2 of its own XCD, where the neighbouring
// workgroups have read or written it, while the runs stay interleaved across
// the XCDs so that their loads remain even.  A bijection on [0, total): the
// items past the last whole 8 * kXcdRun keep their order.
__device__ __forceinline__ uint32_t xcd_slice(uint32_t b, uint32_t total)
{
    constexpr uint32_t X = 8, R = kXcdRun;
    if (R <= 1 || b >= total / (X * R) * (X * R)) return b;
    const uint32_t x = b % X, i = b / X;
    return ((i / R) * X + x) * R + i % R;
}

// Sum of v over the 64 lanes of the wave (DPP: quad swaps, half-row and row
// mirrors, then the row broadcasts), returned uniform.
__device__ __forceinline__ uint32_t wave_total(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);   // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);   // row_mirror
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);   // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// The 6 dwords around item i: d[0..1] = 8 bytes before, d[2..5] = the item.
__device__ __forceinline__ void load_item(const uint32_t* row32, int i, uint32_t d[6])
{
    const u32x2 lo = *(const u32x2*)(row32 + 4 * i - 2);
    const u32x4 hi = *(const u32x4*)(row32 + 4 * i);
    d[0] = lo[0]; d[1] = lo[1]; d[2] = hi[0]; d[3] = hi[1]; d[4] = hi[2]; d[5] = hi[3];
}

// Left neighbours (distance bpp) of the 4 dwords of an item.
__device__ __forceinline__ void left4(const uint32_t d[6], int near, uint32_t sh, uint32_t a[4])
{
#pragma unroll
    for (int k = 0; k < 4; k++)
        a[k] = near ? __builtin_amdgcn_alignbyte(d[2 + k], d[1 + k], sh)
                    : __builtin_amdgcn_alignbyte(d[1 + k], d[k], sh);
}

__device__ __forceinline__ uint32_t byte_mask(int nvalid_in_dword)
{
    return nvalid_in_dword >= 4 ? 0xFFFFFFFFu : nvalid_in_dword <= 0 ? 0u : ((1u << (8 * nvalid_in_dword)) - 1u);
}

template <int K>   // items per lane: ceil(16-byte items of the widest row / 64)
__global__ __launch_bounds__(kFilterThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_filter(FilterLaunch L)
{
    constexpr bool KEEP = K <= 8;                         // Paeth results stay in registers
    extern __shared__ u32x4 smem_filter[];
    uint8_t* lut_lds = (uint8_t*)smem_filter;
    uint8_t* bufs = lut_lds + L.lut_cap;
    const int rbuf = L.row_cap + 32;                      // 16 B front pad + row + 16 B slack
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    constexpr int NW = kFilterThreads / 64;

    __shared__ uint32_t next_band;
    int cur_job = -1;
    const uint32_t total = L.counts ? L.counts[1] : L.total_bands;
    const uint32_t first = L.counts ? L.counts[2] : 0u;
    const int njobs = L.counts ? (int)L.counts[0] : L.njobs;
    // bands after a workgroup's first: every grid-th, or from the queue; a
    // single job's launch of one band per workgroup deals them to the XCDs in
    // runs (xcd_slice), so that a band's row above is mostly in its XCD's L2
    const bool runs = !L.jobs && !L.queue && first == 0 && gridDim.x >= total;
    for (uint32_t g = runs ? xcd_slice(blockIdx.x, total) : first + blockIdx.x; g < total;) {
        int j = 0;
        if (L.jobs) {
            int lo = 0, hi = njobs - 1;
            while (lo < hi) {                              // last job with band_start <= g
                const int mid = (lo + hi + 1) >> 1;
                if (L.band_start[mid] <= g) lo = mid; else hi = mid - 1;
            }
            j = lo;
        }
        const FilterJob J = L.jobs ? L.jobs[j] : L.single;
        const int band = (int)(g - (L.jobs ? L.band_start[j] : 0u));
        const int r0 = band * J.band_rows;
        const int nr = min(J.rows, r0 + J.band_rows) - r0;
        const int rb = J.row_bytes;
        const int nfull = rb >> 4, nitems = (rb + 15) >> 4, tail = rb & 15;

        __syncthreads();                                   // the previous band is done with LDS
        if (j != cur_job) {
            if (J.lut_bytes > 0) {
                const int n16 = (J.lut_bytes + 15) >> 4;
                // (at 6 items per lane the first index is made opaque here,
                // so that the loop's 64-bit start address is not kept live
                // across the kernel, where it took a spill slot)
                int i0 = tid;
                if (K == 6) asm volatile("" : "+v"(i0));
                for (int i = i0; i < n16; i += blockDim.x)
                    ((u32x4*)lut_lds)[i] = ((const u32x4*)J.lut)[i];
            }
            cur_job = j;
        }
        // zero the front pads (left neighbours of every row's first item)
        for (int b = tid; b <= nr; b += blockDim.x) *(u32x4*)(bufs + b * rbuf) = (u32x4){0, 0, 0, 0};
        __syncthreads();

        // ---- window the band and the row above it: buffer b holds row r0 - 1 + b
        {
            // row above: windowed from the source, or the caller's prior row, or zeros
            const bool above_src = r0 > 0;
            const int b_first = above_src ? 0 : 1;
            if (!above_src) {
                uint8_t* prev = bufs + 16;
                if (J.prior) for (int i = tid; i < rb; i += blockDim.x) prev[i] = J.prior[i];
                else for (int i = tid; i < ((rb + 3) >> 2); i += blockDim.x) ((uint32_t*)prev)[i] = 0;
            }
            // full items of buffers b_first..nr, flattened over the rows; the
            // raw loads of WU items are issued together (global latency)
            const int nbuf = nr + 1 - b_first;
            if (nfull > 0 && narrows(J.wmode)) {
                constexpr int WU = 2;                  // 32 raw bytes per item
                const int tot = nbuf * nfull;
                for (int t0 = tid; t0 < tot; t0 += WU * kFilterThreads) {
                    RawItem raw[WU];
                    int bi[WU], ii[WU];
#pragma unroll
                    for (int u = 0; u < WU; u++) {
                        const int t = min(t0 + u * kFilterThreads, tot - 1);
                        bi[u] = t / nfull; ii[u] = t - bi[u] * nfull;
                        raw[u] = load_raw(J, J.src + (int64_t)(r0 - 1 + b_first + bi[u]) * J.src_pitch, ii[u]);
                    }
#pragma unroll
                    for (int u = 0; u < WU; u++)
                        if (t0 + u * kFilterThreads < tot)
                            *(u32x4*)(bufs + (b_first + bi[u]) * rbuf + 16 + 16 * ii[u]) = window_raw(J, raw[u], lut_lds);
                }
            } else if (nfull > 0) {
                // 16 raw bytes per item: twice as many in flight (but for K = 6,
                // whose registers are at the 4-waves-per-SIMD limit)
                constexpr int WU = K == 6 ? 2 : 4;
                const int tot = nbuf * nfull;
                for (int t0 = tid; t0 < tot; t0 += WU * kFilterThreads) {
                    u32x4 raw[WU];
                    int bi[WU], ii[WU];
#pragma unroll
                    for (int u = 0; u < WU; u++) {
                        const int t = min(t0 + u * kFilterThreads, tot - 1);
                        bi[u] = t / nfull; ii[u] = t - bi[u] * nfull;
                        raw[u] = g128u(J.src + (int64_t)(r0 - 1 + b_first + bi[u]) * J.src_pitch + 16 * ii[u]);
                    }
#pragma unroll
                    for (int u = 0; u < WU; u++)
                        if (t0 + u * kFilterThreads < tot)
                            *(u32x4*)(bufs + (b_first + bi[u]) * rbuf + 16 + 16 * ii[u]) =
                                window_raw(J, RawItem{raw[u], raw[u]}, lut_lds);
                }
            }
            for (int t = tid; t < nbuf * tail; t += blockDim.x) {
                const int b = t / tail, jj = (nfull << 4) + (t - b * tail);
                const int r = r0 - 1 + b_first + b;
                bufs[(b_first + b) * rbuf + 16 + jj] = (uint8_t)window_byte(J, J.src + (int64_t)r * J.src_pitch, jj, lut_lds);
            }
        }
        __syncthreads();

        // ---- each wave filters whole rows of the band
        const int near = J.bpp <= 4;                      // neighbour within 1 dword back
        const uint32_t sh = (uint32_t)((4 - (J.bpp & 3)) & 3);
        for (int jr = wave; jr < nr; jr += NW) {
            const uint8_t* cur8 = bufs + (jr + 1) * rbuf + 16;
            const uint32_t* cur = (const uint32_t*)cur8;
            const uint32_t* prv = (const uint32_t*)(bufs + jr * rbuf + 16);
            uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
            uint32_t f4c[KEEP ? K : 1][4], f4t[4];
            // scores of the item at i; the first nv bytes count (a row's
            // last item may be partial: only its lane takes the masked path)
            auto score = [&](int i, int nv, uint32_t* f4o) {
                uint32_t dc[6], dp[6], a[4], c[4];
                load_item(cur, i, dc);
                load_item(prv, i, dp);
                left4(dc, near, sh, a);
                left4(dp, near, sh, c);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t m = nv >= 16 ? ~0u : byte_mask(nv - 4 * q);
                    uint32_t sv[5] = {s0, s1, s2, s3, s4};
                    f4o[q] = cost_dword(dc[2 + q], a[q], dp[2 + q], c[q], m, sv);
                    s0 = sv[0]; s1 = sv[1]; s2 = sv[2]; s3 = sv[3]; s4 = sv[4];
                }
            };
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = lane + 64 * k;
                // one item in flight at a time bounds the registers; the
                // other waves of the SIMD hide the LDS latency
                __builtin_amdgcn_sched_barrier(0);
                if (i < nfull) score(i, 16, f4c[KEEP ? k : 0]);
            }
            const bool ragged = tail != 0;
            const int tl = nfull & 63;                    // lane of the partial item
            if (ragged && lane == tl) score(nfull, tail, f4t);
            const uint32_t sc[5] = {wave_total(s0), wave_total(s1), wave_total(s2), wave_total(s3), wave_total(s4)};
            int best = 0; uint32_t bsc = sc[0];
#pragma unroll
            for (int f = 1; f < 5; f++) if (sc[f] < bsc) { bsc = sc[f]; best = f; }
            // the write pass reads the rows from LDS again rather than keeping
            // every item of the cost pass in registers
            asm volatile("" ::: "memory");

            const int r = r0 + jr;
            uint8_t* dst = J.dst + (int64_t)r * (rb + 1);
            if (lane == 0) dst[0] = (uint8_t)best;
            // filtered bytes of the item at i (f4: its Paeth result, if kept)
            auto filtered = [&](int i, const uint32_t* f4, bool kept, uint32_t* o) {
                if (kept && best == 4) {
#pragma unroll
                    for (int q = 0; q < 4; q++) o[q] = f4[q];
                    return;
                }
                uint32_t dc[6], dp[6], a[4], c[4];
                load_item(cur, i, dc);
                load_item(prv, i, dp);
                left4(dc, near, sh, a);
                if (best == 4) left4(dp, near, sh, c);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t x = dc[2 + q], b = dp[2 + q];
                    o[q] = best == 0 ? x : best == 1 ? sub8(x, a[q]) : best == 2 ? sub8(x, b)
                         : best == 3 ? sub8(x, avg8(a[q], b))
                         : join_lanes(lanes_even(x) - paeth_pred(lanes_even(a[q]), lanes_even(b), lanes_even(c[q])),
                                      lanes_odd(x) - paeth_pred(lanes_odd(a[q]), lanes_odd(b), lanes_odd(c[q])));
                }
            };
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = lane + 64 * k;
                __builtin_amdgcn_sched_barrier(0);
                if (i < nfull) {
                    uint32_t o[4];
                    filtered(i, f4c[KEEP ? k : 0], KEEP, o);
                    *(u32x4_ua*)(dst + 1 + 16 * i) = (u32x4){o[0], o[1], o[2], o[3]};
                }
            }
            if (ragged && lane == tl) {
                uint32_t o[4];
                filtered(nfull, f4t, true, o);
                uint8_t* p = dst + 1 + 16 * nfull;
                for (int t = 0; t < tail; t++) p[t] = (uint8_t)(o[t >> 2] >> (8 * (t & 3)));
            }
            if (J.last_row && r == J.rows - 1)
                for (int i = lane; i < rb; i += 64) J.last_row[i] = cur8[i];
        }
        if (L.queue) {
            // (the loop's first barrier keeps this write after every read of
            // the previous value)
            if (tid == 0) next_band = first + gridDim.x + atomicAdd(L.queue, 1u);
            __syncthreads();
            g = next_band;
        } else {
            g += gridDim.x;
        }
    }
}

// ---------------------------------------------------------------------------
// k_filter_split: the same window + filter for rows wider than k_filter
// takes (6 KiB to 32 KiB), where one row per wave would hold too many items
// per lane (and one workgroup per CU, by its registers).
//
// Persistent workgroups of 8 waves, kFilterSplitWgPerCu per CU.  A band is BR = band_rows
// rows (1, 2, 4 or 8, chosen by the row width); each row is filtered by
// WPR = 8 / BR waves, thread p = (wave % WPR) * 64 + lane of the row taking
// items p, p + 64 WPR, ... (at most K).  LDS holds the window table and
// BR + 1 PNG-ready rows: row r of a job in slot (r + 1) mod (BR + 1), each
// slot a 16-byte zero front pad (the first item's left neighbours), the row
// and 16 bytes of slack.  A workgroup takes a contiguous run of a single
// job's bands (a band's row above is then the previous band's last row,
// already in its slot) or every gridDim.x-th band of a job table.  The raw
// input of the next band is loaded into registers before the current band
// is scored, so its latency hides behind the scoring and the write pass; it
// is windowed into the slots the band no longer needs once the write pass
// is done (three barriers per band).
// ---------------------------------------------------------------------------
template <int K>   // items per lane of a row: ceil(16-byte items / (64 * waves per row))
__global__ __launch_bounds__(kFilterRowThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_filter_split(FilterLaunch L)
{
    constexpr int NT = kFilterRowThreads, NW = NT / 64, PF = kFilterPrefetch, PH = PF / 2;
    static_assert(NW == 8, "band rows times waves per row is 8");
    static_assert(PF % 2 == 0, "the narrowing modes prefetch two halves per item");
    extern __shared__ u32x4 smem_filter[];
    uint8_t* lut_lds = (uint8_t*)smem_filter;
    uint8_t* bufs = lut_lds + L.lut_cap;
    const int rbuf = L.row_cap + 32;                      // 16 B front pad + row + 16 B slack
    __shared__ uint32_t red[NW][5];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

    const uint32_t total = L.counts ? L.counts[1] : L.total_bands;
    const uint32_t first = L.counts ? L.counts[2] : 0u;
    const int njobs = L.counts ? (int)L.counts[0] : L.njobs;
    if (total <= first) return;
    // this workgroup's bands: a contiguous run of a single job's bands, or
    // every gridDim.x-th band of a job table
    uint32_t g, g_hi, g_step;
    if (!L.jobs) {
        const uint64_t nb = total - first;
        g = first + (uint32_t)(nb * blockIdx.x / gridDim.x);
        g_hi = first + (uint32_t)(nb * (blockIdx.x + 1) / gridDim.x);
        g_step = 1;
    } else {
        g = first + blockIdx.x; g_hi = total; g_step = gridDim.x;
    }
    if (g >= g_hi) return;                                 // uniform per workgroup
    auto job_of = [&](uint32_t gb) {
        int lo = 0;
        if (L.jobs) {
            int hi = njobs - 1;
            while (lo < hi) {                              // last job with band_start <= gb
                const int mid = (lo + hi + 1) >> 1;
                if (L.band_start[mid] <= gb) lo = mid; else hi = mid - 1;
            }
        }
        return lo;
    };
    // the slots' front pads stay zero: row data never reaches them
    for (int b = tid; b <= L.band_rows; b += NT) *(u32x4*)(bufs + b * rbuf) = (u32x4){0, 0, 0, 0};

    // Raw input of rows [ra, re) of job JB, items flattened over the rows:
    // thread tid holds items tid + NT u in pre[u] (u < PF), or for the modes
    // that narrow 16-bit samples both 16-byte halves in pre[u] and pre[u + PF
    // / 2] (u < PF / 2); items past those are loaded when windowed.
    u32x4 pre[PF];
    auto prefetch = [&](const FilterJob& JB, int ra, int re) {
        const int nf = JB.row_bytes >> 4;
        const int tot = (re - ra) * nf;
        const bool nw = narrows(JB.wmode);
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int t = tid + NT * (nw ? u % PH : u);
            if (t < tot) {
                const int rr = t / nf, i = t - rr * nf;
                const uint8_t* srow = JB.src + (int64_t)(ra + rr) * JB.src_pitch;
                pre[u] = g128u(nw ? srow + 32 * i + 16 * (u / PH) : srow + 16 * i);
            }
        }
    };
    // PNG-ready bytes of rows [ra, re) of JB into their slots, row r in slot
    // (r + 1) mod (band rows + 1)
    auto window_rows = [&](const FilterJob& JB, int ra, int re) {
        const int nf = JB.row_bytes >> 4, tl = JB.row_bytes & 15, S = JB.band_rows + 1;
        const int tot = (re - ra) * nf;
        const bool nw = narrows(JB.wmode);
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int t = tid + NT * u;
            if (t < tot && (!nw || u < PH)) {
                const int rr = t / nf, i = t - rr * nf;
                const RawItem R = nw ? RawItem{pre[u], pre[(u + PH) % PF]} : RawItem{pre[u], pre[u]};
                *(u32x4*)(bufs + ((ra + rr + 1) % S) * rbuf + 16 + 16 * i) = window_raw(JB, R, lut_lds);
            }
        }
        for (int t = tid + NT * (nw ? PH : PF); t < tot; t += NT) {
            const int rr = t / nf, i = t - rr * nf;
            const RawItem R = load_raw(JB, JB.src + (int64_t)(ra + rr) * JB.src_pitch, i);
            *(u32x4*)(bufs + ((ra + rr + 1) % S) * rbuf + 16 + 16 * i) = window_raw(JB, R, lut_lds);
        }
        for (int t = tid; t < (re - ra) * tl; t += NT) {
            const int rr = t / tl, jj = (nf << 4) + (t - rr * tl);
            bufs[((ra + rr + 1) % S) * rbuf + 16 + jj] =
                (uint8_t)window_byte(JB, JB.src + (int64_t)(ra + rr) * JB.src_pitch, jj, lut_lds);
        }
    };
    // the row above a job's first row (slot 0): the caller's prior row or zeros
    auto above_first = [&](const FilterJob& JB) {
        uint8_t* prev = bufs + 16;
        if (JB.prior) for (int i = tid; i < JB.row_bytes; i += NT) prev[i] = JB.prior[i];
        else for (int i = tid; i < ((JB.row_bytes + 3) >> 2); i += NT) ((uint32_t*)prev)[i] = 0;
    };
    auto load_lut = [&](const FilterJob& JB) {
        const int n16 = (JB.lut_bytes + 15) >> 4;
        for (int i = tid; i < n16; i += NT) ((u32x4*)lut_lds)[i] = ((const u32x4*)JB.lut)[i];
    };

    // the first band: its rows and the row above, loaded and windowed
    int j = job_of(g);
    FilterJob J = L.jobs ? L.jobs[j] : L.single;
    int r0 = (int)(g - (L.jobs ? L.band_start[j] : 0u)) * J.band_rows;
    int nr = min(J.rows, r0 + J.band_rows) - r0;
    int lut_job = -1;
    {
        const int ra = r0 > 0 ? r0 - 1 : r0;
        prefetch(J, ra, r0 + nr);                          // in flight while the table is copied
        if (J.lut_bytes > 0) { load_lut(J); lut_job = j; }
        __syncthreads();
        if (r0 == 0) above_first(J);
        window_rows(J, ra, r0 + nr);
        __syncthreads();
    }

    for (;;) {
        // ---- the next band's raw input, in flight during this band's work
        const uint32_t gn = g + g_step;
        const bool has_next = gn < g_hi;
        int jn = j, r0n = 0, nrn = 0, ran = 0;
        FilterJob JN = J;
        if (has_next) {
            jn = job_of(gn);
            if (jn != j) JN = L.jobs[jn];
            r0n = (int)(gn - (L.jobs ? L.band_start[jn] : 0u)) * JN.band_rows;
            nrn = min(JN.rows, r0n + JN.band_rows) - r0n;
            // the row above is this band's last row, in its slot, when the
            // next band follows it in the same job
            const bool reuse = jn == j && r0n == r0 + nr;
            ran = reuse || r0n == 0 ? r0n : r0n - 1;
            prefetch(JN, ran, r0n + nrn);
        }

        // ---- score this band: row jr of the band by waves jr * WPR .. + WPR - 1
        const int rb = J.row_bytes, nfull = rb >> 4, tail = rb & 15, S = J.band_rows + 1;
        const int WPR = NW / J.band_rows;
        const int jr = wave / WPR, part = wave - jr * WPR;
        const bool active = jr < nr;
        const int near = J.bpp <= 4;                      // neighbour within 1 dword back
        const uint32_t sh = (uint32_t)((4 - (J.bpp & 3)) & 3);
        const uint8_t* cur8 = bufs + ((r0 + jr + 1) % S) * rbuf + 16;
        const uint32_t* cur = (const uint32_t*)cur8;
        const uint32_t* prv = (const uint32_t*)(bufs + ((r0 + jr) % S) * rbuf + 16);
        uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
        uint32_t f4c[K][4];
        // scores of the item at i; its first nv bytes count
        auto score = [&](int i, int nv, uint32_t* f4o) {
            uint32_t dc[6], dp[6], a[4], c[4];
            load_item(cur, i, dc);
            load_item(prv, i, dp);
            left4(dc, near, sh, a);
            left4(dp, near, sh, c);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t m = nv >= 16 ? ~0u : byte_mask(nv - 4 * q);
                uint32_t sv[5] = {s0, s1, s2, s3, s4};
                f4o[q] = cost_dword(dc[2 + q], a[q], dp[2 + q], c[q], m, sv);
                s0 = sv[0]; s1 = sv[1]; s2 = sv[2]; s3 = sv[3]; s4 = sv[4];
            }
        };
        const int p0 = part * 64 + lane, pstep = 64 * WPR;
        if (active) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = p0 + pstep * k;
                if (i < nfull) score(i, 16, f4c[k]);
                else if (i == nfull && tail) score(i, tail, f4c[k]);
            }
        }
        {
            const uint32_t sc[5] = {wave_total(s0), wave_total(s1), wave_total(s2), wave_total(s3), wave_total(s4)};
            if (lane == 0)
#pragma unroll
                for (int f = 0; f < 5; f++) red[wave][f] = sc[f];
        }
        __syncthreads();

        // ---- the row's filter (lowest index on ties) and its bytes
        if (active) {
            uint32_t sc[5] = {0, 0, 0, 0, 0};
            for (int w = jr * WPR; w < jr * WPR + WPR; w++)
#pragma unroll
                for (int f = 0; f < 5; f++) sc[f] += red[w][f];
            int best = 0; uint32_t bsc = sc[0];
#pragma unroll
            for (int f = 1; f < 5; f++) if (sc[f] < bsc) { bsc = sc[f]; best = f; }
            const int r = r0 + jr;
            uint8_t* dst = J.dst + (int64_t)r * (rb + 1);
            if (part == 0 && lane == 0) dst[0] = (uint8_t)best;
            // filtered bytes of the item at i (f4: its Paeth result)
            auto filtered = [&](int i, const uint32_t* f4, uint32_t* o) {
                if (best == 4) {
#pragma unroll
                    for (int q = 0; q < 4; q++) o[q] = f4[q];
                    return;
                }
                uint32_t dc[6], dp[6], a[4];
                load_item(cur, i, dc);
                load_item(prv, i, dp);
                left4(dc, near, sh, a);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t x = dc[2 + q], b = dp[2 + q];
                    o[q] = best == 0 ? x : best == 1 ? sub8(x, a[q]) : best == 2 ? sub8(x, b) : sub8(x, avg8(a[q], b));
                }
            };
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int i = p0 + pstep * k;
                if (i < nfull) {
                    uint32_t o[4];
                    filtered(i, f4c[k], o);
                    *(u32x4_ua*)(dst + 1 + 16 * i) = (u32x4){o[0], o[1], o[2], o[3]};
                } else if (i == nfull && tail) {
                    uint32_t o[4];
                    filtered(i, f4c[k], o);
                    uint8_t* p = dst + 1 + 16 * nfull;
                    for (int t = 0; t < tail; t++) p[t] = (uint8_t)(o[t >> 2] >> (8 * (t & 3)));
                }
            }
            if (J.last_row && r == J.rows - 1)
                for (int i = p0; i < rb; i += pstep) J.last_row[i] = cur8[i];
        }
        if (!has_next) break;
        __syncthreads();                                   // every read of this band's slots done

        // ---- the next band's rows into their slots
        if (jn != j) {
            J = JN;
            if (J.lut_bytes > 0 && lut_job != jn) {
                load_lut(J);
                lut_job = jn;
                __syncthreads();
            }
            j = jn;
        }
        if (r0n == 0) above_first(J);
        window_rows(J, ran, r0n + nrn);
        g = gn; r0 = r0n; nr = nrn;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_filter_tiled: rows wider than kMaxRowBytesLds (filter_select_kernel
// filters any width, gpu_filter.cu:206-233, 285-308; the reference sizes its
// buffers for 17043-pixel RGB16 rows, gpu_deflate_backend.h:7-9).  One row
// per workgroup at a time, in tiles of kTileBytes PNG-ready bytes: pass 1
// windows each tile of the row and of the row above into LDS, with the 16
// bytes before the tile as left neighbours, and accumulates the five scores;
// the workgroup then picks the filter (lowest index on ties) and pass 2
// windows the tiles again and writes the chosen filter's bytes.  The jobs of
// this launch have band_rows == 1, so a band is a row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFilterThreads) void k_filter_tiled(FilterLaunch L)
{
    extern __shared__ u32x4 smem_tiled[];
    uint8_t* lut_lds = (uint8_t*)smem_tiled;
    uint8_t* tcur = lut_lds + L.lut_cap;                  // 16 B front pad + tile + 16 B slack
    uint8_t* tprv = tcur + kTileBytes + 32;
    constexpr int NW = kFilterThreads / 64;
    __shared__ uint32_t red[5][NW];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

    int cur_job = -1;
    const uint32_t total = L.counts ? L.counts[1] : L.total_bands;
    const uint32_t first = L.counts ? L.counts[2] : 0u;
    const int njobs = L.counts ? (int)L.counts[0] : L.njobs;
    for (uint32_t g = first + blockIdx.x; g < total; g += gridDim.x) {
        int j = 0;
        if (L.jobs) {
            int lo = 0, hi = njobs - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (L.band_start[mid] <= g) lo = mid; else hi = mid - 1;
            }
            j = lo;
        }
        const FilterJob J = L.jobs ? L.jobs[j] : L.single;
        const int r = (int)(g - (L.jobs ? L.band_start[j] : 0u));
        const int rb = J.row_bytes;
        const int ntiles = (rb + kTileBytes - 1) / kTileBytes;
        const uint8_t* src_cur = J.src + (int64_t)r * J.src_pitch;
        const uint8_t* src_prv = r > 0 ? J.src + (int64_t)(r - 1) * J.src_pitch : nullptr;

        __syncthreads();                                   // the previous row is done with LDS
        if (j != cur_job) {
            if (J.lut_bytes > 0) {
                const int n16 = (J.lut_bytes + 15) >> 4;
                for (int i = tid; i < n16; i += kFilterThreads) ((u32x4*)lut_lds)[i] = ((const u32x4*)J.lut)[i];
            }
            cur_job = j;
            __syncthreads();
        }
        // tile t of the row (and, with_prev, of the row above: windowed from
        // the source, or the caller's prior row, or zeros), each with the
        // item before it as its front pad (zeros before the row)
        auto stage = [&](int t, bool with_prev) {
            const int base = t * kTileBytes;
            const int n = min(kTileBytes, rb - base);
            const int nfull = n >> 4, tail = n & 15;
            const int per = nfull + 1;
            const int nrow = with_prev ? 2 : 1;
            for (int q = tid; q < nrow * per; q += kFilterThreads) {
                const int which = q >= per ? 1 : 0;
                const int i = q - which * per - 1;
                const int item = (base >> 4) + i;
                u32x4 v = (u32x4){0, 0, 0, 0};
                if (item >= 0) {
                    if (which == 0) v = window_raw(J, load_raw(J, src_cur, item), lut_lds);
                    else if (src_prv) v = window_raw(J, load_raw(J, src_prv, item), lut_lds);
                    else if (J.prior) v = *(const u32x4_ua*)(J.prior + 16 * (int64_t)item);
                }
                *(u32x4*)((which ? tprv : tcur) + 16 + 16 * i) = v;
            }
            for (int q = tid; q < nrow * tail; q += kFilterThreads) {
                const int which = q >= tail ? 1 : 0;
                const int jj = (nfull << 4) + (q - which * tail);
                uint32_t b = 0;
                if (which == 0) b = window_byte(J, src_cur, base + jj, lut_lds);
                else if (src_prv) b = window_byte(J, src_prv, base + jj, lut_lds);
                else if (J.prior) b = J.prior[base + jj];
                (which ? tprv : tcur)[16 + jj] = (uint8_t)b;
            }
        };
        const int near = J.bpp <= 4;
        const uint32_t sh = (uint32_t)((4 - (J.bpp & 3)) & 3);
        const uint32_t* cur = (const uint32_t*)(tcur + 16);
        const uint32_t* prv = (const uint32_t*)(tprv + 16);

        // ---- pass 1: the five scores over all tiles
        uint32_t s[5] = {0, 0, 0, 0, 0};
        for (int t = 0; t < ntiles; t++) {
            if (t) __syncthreads();                        // the previous tile is scored
            stage(t, true);
            __syncthreads();
            const int n = min(kTileBytes, rb - t * kTileBytes);
            const int nitems = (n + 15) >> 4;
            for (int i = tid; i < nitems; i += kFilterThreads) {
                const int nv = min(16, n - 16 * i);
                uint32_t dc[6], dp[6], a[4], c[4];
                load_item(cur, i, dc);
                load_item(prv, i, dp);
                left4(dc, near, sh, a);
                left4(dp, near, sh, c);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t m = nv >= 16 ? ~0u : byte_mask(nv - 4 * q);
                    (void)cost_dword(dc[2 + q], a[q], dp[2 + q], c[q], m, s);
                }
            }
        }
#pragma unroll
        for (int f = 0; f < 5; f++) {
            const uint32_t v = wave_total(s[f]);
            if (lane == 0) red[f][wave] = v;
        }
        __syncthreads();
        uint32_t sc[5];
#pragma unroll
        for (int f = 0; f < 5; f++) {
            sc[f] = 0;
            for (int w = 0; w < NW; w++) sc[f] += red[f][w];
        }
        int best = 0;
        uint32_t bsc = sc[0];
#pragma unroll
        for (int f = 1; f < 5; f++) if (sc[f] < bsc) { bsc = sc[f]; best = f; }

        // ---- pass 2: the chosen filter's bytes
        uint8_t* dst = J.dst + (int64_t)r * (rb + 1);
        if (tid == 0) dst[0] = (uint8_t)best;
        for (int t = 0; t < ntiles; t++) {
            __syncthreads();                               // red is read / the previous tile is written
            stage(t, best >= 2);
            __syncthreads();
            const int base = t * kTileBytes;
            const int n = min(kTileBytes, rb - base);
            const int nitems = (n + 15) >> 4;
            for (int i = tid; i < nitems; i += kFilterThreads) {
                const int nv = min(16, n - 16 * i);
                uint32_t dc[6], dp[6] = {0, 0, 0, 0, 0, 0}, a[4], c[4] = {0, 0, 0, 0};
                load_item(cur, i, dc);
                if (best >= 2) load_item(prv, i, dp);
                left4(dc, near, sh, a);
                if (best == 4) left4(dp, near, sh, c);
                uint32_t o[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t x = dc[2 + q], b = dp[2 + q];
                    o[q] = best == 0 ? x : best == 1 ? sub8(x, a[q]) : best == 2 ? sub8(x, b)
                         : best == 3 ? sub8(x, avg8(a[q], b))
                         : join_lanes(lanes_even(x) - paeth_pred(lanes_even(a[q]), lanes_even(b), lanes_even(c[q])),
                                      lanes_odd(x) - paeth_pred(lanes_odd(a[q]), lanes_odd(b), lanes_odd(c[q])));
                }
                uint8_t* p = dst + 1 + base + 16 * i;
                if (nv == 16) *(u32x4_ua*)p = (u32x4){o[0], o[1], o[2], o[3]};
                else for (int k = 0; k < nv; k++) p[k] = (uint8_t)(o[k >> 2] >> (8 * (k & 3)));
            }
            if (J.last_row && r == J.rows - 1)
                for (int i = tid; i < n; i += kFilterThreads) J.last_row[base + i] = tcur[16 + i];
        }
    }
}

// ---------------------------------------------------------------------------
// Adler-32 of a byte range, computed by a whole workgroup.
// Per 16-byte item at offset o: S += sum(x), T += o*sum(x) + sum(j*x_j);
// then A = 1 + S, B = n + n*S - T  (mod 65521), B = n + sum((n-i) x_i).
// Returns the adler value on every thread.  `red` needs 2*nwaves u32.
// ---------------------------------------------------------------------------
__device__ uint32_t adler_finish(uint32_t S, uint32_t t32, int64_t n, uint32_t* red);

__device__ uint32_t block_adler(const uint8_t* p, int64_t n, uint32_t* red)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    uint32_t S = 0;
    uint64_t T = 0;
    const int64_t nfull = n >> 4;
    for (int64_t i0 = tid; i0 < nfull; i0 += 4 * (int64_t)blockDim.x) {
        u32x4 v[4];                                   // four loads in flight
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t i = i0 + u * (int64_t)blockDim.x;
            v[u] = i < nfull ? g128u(p + 16 * i) : (u32x4){0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t i = i0 + u * (int64_t)blockDim.x;
            uint32_t s16 = 0, w16 = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t sk = __builtin_amdgcn_udot4(v[u][k], 0x01010101u, 0u, false);
                w16 = __builtin_amdgcn_udot4(v[u][k], 0x03020100u, w16, false) + (uint32_t)(4 * k) * sk;
                s16 += sk;
            }
            S += s16;
            T += (uint64_t)(16 * (i < nfull ? i : 0)) * s16 + w16;
        }
    }
    for (int64_t j = (nfull << 4) + tid; j < n; j += blockDim.x) {
        const uint32_t b = g8(p + j);
        S += b;
        T += (uint64_t)j * b;
    }
    // S < 2^32 (n <= 64 KiB); T mod M from its 32-bit halves (2^32 = 225 mod M)
    constexpr uint32_t M = 65521;
    const uint32_t t32 = ((((uint32_t)(T >> 32)) % M) * 225u + ((uint32_t)T) % M) % M;
    return adler_finish(S, t32, n, red);
}

// The Adler-32 of n bytes from every thread's terms: S (its bytes' sum,
// < 2^26) and T (the sum of offset * byte, mod 65521).  Returns the value
// on every thread.  `red` needs 2*nwaves u32.
__device__ uint32_t adler_finish(uint32_t S, uint32_t t32, int64_t n, uint32_t* red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    constexpr uint32_t M = 65521;
    const uint32_t Sw = wave_total(S) % M;         // 64 lanes: no overflow
    const uint32_t Tw = wave_total(t32) % M;
    __syncthreads();
    if (lane == 0) { red[wave * 2] = Sw; red[wave * 2 + 1] = Tw; }
    __syncthreads();
    uint32_t St = 0, Tt = 0;
    for (int w = 0; w < nw; w++) { St += red[w * 2]; Tt += red[w * 2 + 1]; }
    St %= M; Tt %= M;
    const uint32_t A = (1u + St) % M;
    const uint32_t nm = (uint32_t)((uint64_t)n % M);
    const uint32_t B = (nm * A + M - Tt) % M;      // n (1 + S) - T
    return (B << 16) | A;
}

__device__ __forceinline__ uint32_t adler_combine_dev(uint32_t a1, uint32_t a2, uint64_t len2)
{
    const uint64_t M = 65521;
    const uint64_t A1 = a1 & 0xFFFF, B1 = a1 >> 16, A2 = a2 & 0xFFFF, B2 = a2 >> 16;
    const uint64_t n = len2 % M;
    const uint64_t A = (A1 + A2 + M - 1) % M;
    const uint64_t B = (B1 + B2 + n * ((A1 + M - 1) % M)) % M;
    return (uint32_t)((B << 16) | A);
}

// ---------------------------------------------------------------------------
// CRC-32 helpers (reflected polynomial 0xEDB88320)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t gf2_mulmod(uint32_t a, uint32_t b)
{
    uint32_t prod = 0;
    for (int k = 0; k < 32; k++) {
        if (a & (0x80000000u >> k)) prod ^= b;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
    }
    return prod;
}
__device__ __forceinline__ uint32_t gf2_x8n(uint64_t n)
{
    uint32_t r = 0x80000000u;
    for (int k = 0; n; k++, n >>= 1)
        if (n & 1) r = gf2_mulmod(r, c_x8pow[k]);
    return r;
}
__device__ __forceinline__ uint32_t crc_combine_dev(uint32_t c1, uint32_t c2, uint64_t len2)
{
    return len2 ? gf2_mulmod(gf2_x8n(len2), c1) ^ c2 : c1;
}
// crc32 update over n bytes from generic memory, slice-by-4 tables in LDS.
__device__ uint32_t crc_bytes(uint32_t crc, const uint8_t* p, int64_t n, const uint32_t (*T)[256])
{
    uint32_t c = ~crc;
    int64_t i = 0;
    for (; i + 4 <= n; i += 4) {
        c ^= ld32u(p + i);
        c = T[3][c & 0xFF] ^ T[2][(c >> 8) & 0xFF] ^ T[1][(c >> 16) & 0xFF] ^ T[0][c >> 24];
    }
    for (; i < n; i++) c = T[0][(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return ~c;
}

// CRC of [p, p+n) by a whole workgroup: contiguous per-thread ranges, then a
// tree combine in LDS.  `tab` = slice-by-4 tables in LDS; `red` needs
// 2*blockDim u32.  Result on every thread.
__device__ uint32_t block_crc(const uint8_t* p, int64_t n, const uint32_t (*tab)[256], uint32_t* red)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int64_t per = (((n + nt - 1) / nt) + 3) & ~(int64_t)3;   // dword-aligned ranges
    const int64_t b = min64(n, per * tid), e = min64(n, b + per);
    uint32_t c = crc_bytes(0, p + b, e - b, tab);
    uint32_t len = (uint32_t)(e - b);
    __syncthreads();
    red[tid] = c; red[nt + tid] = len;
    __syncthreads();
    for (int s = 1; s < nt; s <<= 1) {
        if ((tid & (2 * s - 1)) == 0 && tid + s < nt) {
            const uint32_t l2 = red[nt + tid + s];
            red[tid] = crc_combine_dev(red[tid], red[tid + s], l2);
            red[nt + tid] += l2;
        }
        __syncthreads();
    }
    const uint32_t r = red[0];
    __syncthreads();
    return r;
}

// One slice-by-4 step: 4 more bytes (w, little-endian) into register c.
__device__ __forceinline__ uint32_t crc_step4(uint32_t c, uint32_t w, const uint32_t (*T)[256])
{
    c ^= w;
    return T[3][c & 0xFF] ^ T[2][(c >> 8) & 0xFF] ^ T[1][(c >> 16) & 0xFF] ^ T[0][c >> 24];
}

// c times the fixed polynomial a shift table S holds, by input byte.
__device__ __forceinline__ uint32_t crc_shift(uint32_t c, const uint32_t (*S)[256])
{
    return S[0][c & 0xFF] ^ S[1][(c >> 8) & 0xFF] ^ S[2][(c >> 16) & 0xFF] ^ S[3][c >> 24];
}

__device__ void load_crc_tables(uint32_t (*tab)[256])
{
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) (&tab[0][0])[i] = (&c_crc[0][0])[i];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Standalone checksums: a partial per kChecksumBlock-byte block, then one
// combine workgroup.
// ---------------------------------------------------------------------------
// Per block its two terms of the sum form of the Adler-32 combine (see
// k_checksum_final), A_i - 1 and B_i + (A_i - 1) R_i with R_i the bytes
// after the block, both mod 65521, packed as low | high << 16.
__global__ __launch_bounds__(256) void k_adler_partial(const uint8_t* p, int64_t n, uint32_t* part)
{
    __shared__ uint32_t red[16];
    constexpr uint32_t M = 65521;
    const int64_t b = (int64_t)blockIdx.x * kChecksumBlock;
    const int64_t len = min64(kChecksumBlock, n - b);
    const uint32_t a = block_adler(p + b, len, red);
    if (threadIdx.x == 0) {
        const uint32_t R = (uint32_t)((uint64_t)(n - b - len) % M);
        const uint32_t am = (a & 0xFFFFu) ? (a & 0xFFFFu) - 1u : M - 1u;
        part[blockIdx.x] = am | ((((a >> 16) + (am * R) % M) % M) << 16);
    }
}

// XOR over the 64 lanes of the wave (the DPP pattern of wave_total), uniform.
__device__ __forceinline__ uint32_t wave_xor(uint32_t v)
{
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

constexpr int kFinalThreads = 1024;     // the combining workgroup
constexpr int kCrcPartialGrid = 2048;   // 8 workgroups on each of 256 CUs
constexpr int kCrcRow = 256 * 32;       // a row: 32 bytes for each of 256 threads
constexpr int kCrcRows = kCrcBlock / kCrcRow;
static_assert(kCrcRows % 4 == 0, "rows in groups of four");

// CRC partials: per block, its register run from 0 with no final inversion.
// Blocks are counted back from the end: block j covers [n - (j + 1) B,
// n - j B) (the data's first block may be short: leading zeros leave a
// register that starts at 0 unchanged, so it is read as if padded at the
// front).  A block is kCrcRows rows of 8 KiB; thread t takes the 32 bytes at
// 32 t of every row, so each load of a wave covers 1 KiB of whole cache lines.
// A thread runs four rows' pieces as four chains and folds them Horner-wise
// (h = h x^(8*8 KiB) + c_q), so its register covers its pieces of every row;
// the threads then combine in a tree from the 32-byte shift up, the tree and
// the workgroup's one barrier once per block.  One table lookup per byte and
// one more per 32.
__global__ __launch_bounds__(256) void k_crc_partial(const uint8_t* p, int64_t n, uint32_t* part)
{
    __shared__ uint32_t tab[4][256];
    __shared__ uint32_t srow[4][256];     // x^(8 kCrcRow): one row on
    __shared__ uint32_t red[2][4];
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) (&srow[0][0])[i] = (&c_crcshift[7][0][0])[i];
    static_assert((64 << 7) == kCrcRow, "row shift table");
    load_crc_tables(tab);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nb = (int)((n + kCrcBlock - 1) / kCrcBlock);
    // persistent: the tables are loaded once, blocks are dealt round-robin
    int it = 0;
    for (int blk = blockIdx.x; blk < nb; blk += gridDim.x, it ^= 1) {
        const int64_t pb = n - (int64_t)(blk + 1) * kCrcBlock;    // negative for a short first block
        const uint8_t* s0 = p + pb + tid * 32;                    // row q's piece at s0 + kCrcRow q
        uint32_t h = 0;
        if (pb >= 0) {
            for (int q0 = 0; q0 < kCrcRows; q0 += 4) {
                u32x4 v[4][2];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    v[q][0] = g128u(s0 + (int64_t)kCrcRow * (q0 + q));
                    v[q][1] = g128u(s0 + (int64_t)kCrcRow * (q0 + q) + 16);
                }
                uint32_t c[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int k = 0; k < 8; k++)
#pragma unroll
                    for (int q = 0; q < 4; q++) c[q] = crc_step4(c[q], v[q][k >> 2][k & 3], tab);
#pragma unroll
                for (int q = 0; q < 4; q++) h = crc_shift(h, srow) ^ c[q];
            }
        } else {
            for (int q = 0; q < kCrcRows; q++) {
                const int64_t e = pb + tid * 32 + (int64_t)kCrcRow * q + 32;
                uint32_t c = 0;
                for (int64_t i = max64(0, e - 32); i < e; i++) c = tab[0][(c ^ p[i]) & 0xFF] ^ (c >> 8);
                h = crc_shift(h, srow) ^ c;
            }
        }
        // thread t's register stands (255 - t) 32 bytes before the block's end
        {
            const uint32_t right = __shfl_down(h, 1, 64);
            if ((lane & 1) == 0) h = crc_shift(h, c_crcshift32) ^ right;
        }
#pragma unroll
        for (int k = 1; k < 6; k++) {
            const uint32_t right = __shfl_down(h, 1 << k, 64);
            if ((lane & ((2 << k) - 1)) == 0) h = crc_shift(h, c_crcshift[k - 1]) ^ right;
        }
        // red is double-buffered by iteration, so one barrier per block
        if (lane == 0) red[it][wave] = h;
        __syncthreads();
        if (tid == 0) {
            const uint32_t v01 = crc_shift(red[it][0], c_crcshift[5]) ^ red[it][1];
            const uint32_t v23 = crc_shift(red[it][2], c_crcshift[5]) ^ red[it][3];
            part[blk] = crc_shift(v01, c_crcshift[6]) ^ v23;
        }
    }
}

// Total register XOR_j p_j * x^(8 j B) of the CRC partials: thread t folds
// the blocks [t m, (t + 1) m) Horner-wise from the top (h = h * x^(8B) + p_j),
// then the threads combine in a tree whose level-k shift is x^(8 m B 2^k);
// m is a power of two, so every shift is one of the tables.  The result is
// ~(init ^ total) with init = ~0 * x^(8n), the initial register's term.
constexpr int kBlockLevel = 11;                        // x^(8 B): 64 * 2^11 = B
static_assert((64 << kBlockLevel) == kCrcBlock, "block shift table");
__global__ __launch_bounds__(kFinalThreads) void k_crc_final(const uint32_t* part, int nb, int lm, uint32_t init,
                                                             uint32_t* result)
{
    constexpr int NW = kFinalThreads / 64;
    __shared__ uint32_t shb[4][256];
    __shared__ uint32_t lvl[10][4][256];                   // the tree's ten shifts, from x^(8 m B) up
    __shared__ uint32_t red[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < 1024; i += kFinalThreads) (&shb[0][0])[i] = (&c_crcshift[kBlockLevel][0][0])[i];
    for (int i = tid; i < 10 * 1024; i += kFinalThreads) (&lvl[0][0][0])[i] = (&c_crcshift[kBlockLevel + lm][0][0])[i];
    __syncthreads();
    const int m = 1 << lm;
    const int lo = min(nb, tid << lm), hi = min(nb, lo + m);
    uint32_t h = 0;
    for (int j0 = hi - 1; j0 >= lo; j0 -= 8) {
        uint32_t pv[8];                                     // eight loads in flight
#pragma unroll
        for (int u = 0; u < 8; u++) pv[u] = j0 - u >= lo ? part[j0 - u] : 0u;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (j0 - u >= lo) h = crc_shift(h, shb) ^ pv[u];
    }
    // thread t's register is worth x^(8 t m B): level k shifts by m B 2^k
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const uint32_t right = __shfl_down(h, 1 << k, 64);
        if ((lane & ((2 << k) - 1)) == 0) h ^= crc_shift(right, lvl[k]);
    }
    if (lane == 0) red[wave] = h;
    __syncthreads();
    if (wave == 0) {
        h = lane < NW ? red[lane] : 0u;
#pragma unroll
        for (int k = 6; k < 10; k++) {
            const uint32_t right = __shfl_down(h, 1 << (k - 6), 64);
            if ((lane & ((2 << (k - 6)) - 1)) == 0) h ^= crc_shift(right, lvl[k]);
        }
        if (lane == 0) *result = ~(init ^ h);
    }
}

// Adler-32 of the nb blocks (every block kChecksumBlock bytes but the last):
// for blocks with sums A_i - 1 and R_i bytes after them, A = 1 + sum(A_i - 1)
// and B = sum(B_i + (A_i - 1) R_i) (mod 65521); the partials hold the two
// terms of each block, so this is a plain sum.
__global__ __launch_bounds__(kFinalThreads) void k_adler_final(const uint32_t* part, int nb, uint32_t* result)
{
    constexpr int NW = kFinalThreads / 64;
    __shared__ uint32_t red[2 * NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        // 16 terms < 2^16 per round, reduced mod M after each round
        constexpr uint32_t M = 65521;
        uint32_t sa = 0, sb = 0;
        for (int i0 = tid; i0 < nb; i0 += 16 * kFinalThreads) {
            uint32_t a[16];                               // sixteen loads in flight
#pragma unroll
            for (int u = 0; u < 16; u++) a[u] = i0 + u * kFinalThreads < nb ? part[i0 + u * kFinalThreads] : 0u;
#pragma unroll
            for (int u = 0; u < 16; u++) { sa += a[u] & 0xFFFFu; sb += a[u] >> 16; }
            sa %= M; sb %= M;
        }
        sa = wave_total(sa) % M;
        sb = wave_total(sb) % M;
        if (lane == 0) { red[wave] = sa; red[NW + wave] = sb; }
        __syncthreads();
        if (tid == 0) {
            uint32_t A = 1, B = 0;
            for (int w = 0; w < NW; w++) { A += red[w]; B += red[NW + w]; }
            *result = ((B % M) << 16) | (A % M);
        }
        return;
    }
}

// ---------------------------------------------------------------------------
// k_deflate: one workgroup per slice.
//
// Segment layout on the wire: [BFINAL|BTYPE=01 fixed-Huffman block][EOB]
// then, unless it is the stream's final segment, an empty stored block
// (sync flush: 3 header bits, pad to a byte, 00 00 FF FF).  Every segment is
// therefore a byte-aligned, independently decodable piece whose history
// (the 32 KiB before it) is already in memory, so all segments of all
// images run in parallel and are joined by a prefix sum of their lengths.
// A segment (64 KiB by default) is cut into slices (16 KiB by default), the
// unit of work: each slice is one workgroup, emits a bit string, and
// k_scan / k_assemble join a segment's slices at bit granularity.
//
// A slice first enters its 32 KiB window into T1 (every kPrimeStride-th
// position, atomicMax), then runs sub-blocks of NT * 16 positions; thread t
// owns the 16 consecutive positions of its parse lane and keeps them in
// registers:
//   A. the positions' 4-byte hashes (kept, two per register); every
//      SS-th position enters T2 (first position of the sub-block
//      with that hash, atomicMin);
//   B. candidates per position: distance 1, distance bpp and one row up
//      (equality bitmasks from the lane's own bytes, no extra loads), and
//      one hashed earlier position -- T2's entry when it precedes the
//      position, else T1's (most recent position before the sub-block)
//      within the window.  The longest match wins, ties to the shorter
//      distance (a max over length << 16 | (0xFFFF - distance) keys);
//      matches stop at the lane end, so each lane parses greedily on its own,
//      after which lanes that are one 16-byte match at the distance of the
//      lane before them join that lane's token (runs of up to 256 bytes);
//   C. per-lane bit counts, block exclusive scan, then each lane packs its
//      tokens into 32-bit words at its bit offset (the two words shared
//      with neighbouring lanes by LDS atomicOr) and full words stream out;
//      every SS-th position of the sub-block then enters T1.
// Table entries carry a tag (the slice for T1, the sub-block for T2), so
// stale entries never match and the tables are not cleared per slice.
// The result depends only on the input bytes: every table update is an
// order-independent min/max, fenced by barriers.
// ---------------------------------------------------------------------------


#ifndef R2P_MATCH_GROUP
#define R2P_MATCH_GROUP 4
#endif
constexpr int kHashBits = 12;      // T1: 4 Ki entries
constexpr int kHash2Bits = 12;     // T2: 4 Ki entries
constexpr int kLane = 16;          // positions per parse lane
// Table entries are tag << kPosBits | position (relative to the window base,
// < 32 KiB + 64 KiB).  The tag names the slice (T1) or sub-block (T2) that
// wrote the entry, so entries of earlier ones never match and the tables
// are cleared only when the workgroup starts or the tags run out.
constexpr int kPosBits = 17;
constexpr uint32_t kPosMask = (1u << kPosBits) - 1;
constexpr uint32_t kTagMax = (1u << (32 - kPosBits)) - 1;
constexpr int kPrimeStride = 8;    // window positions entered into T1 at this stride
// The slice's own positions enter T1 and T2, and are searched for a hashed
// candidate, at a stride SS of 4 (r2p_options.lz77 = 1) or 2 (lz77 = 2): a
// template parameter of k_deflate.
constexpr int kHashCmpDwords = 2;  // a hashed candidate is compared over at most 8 bytes

__device__ __forceinline__ uint32_t hash4(uint32_t w) { return (w * 2654435761u) >> (32 - kHashBits); }

// Bit 4k + i flags a zero byte i of x[k] (k, i < 4).  Each byte's flag is
// its bit 7 (0x80 or 0); full-rate dot products against the byte weights
// 2^i and 2^(4 + i) gather two dwords' flags into one byte, times 128.
__device__ __forceinline__ uint32_t zero_bytes16(const uint32_t* x)
{
    uint32_t z[4];
#pragma unroll
    for (int k = 0; k < 4; k++) z[k] = ~(((x[k] & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x[k]) & 0x80808080u;
    const uint32_t lo = __builtin_amdgcn_udot4(z[1], 0x80402010u, __builtin_amdgcn_udot4(z[0], 0x08040201u, 0u, false), false);
    const uint32_t hi = __builtin_amdgcn_udot4(z[3], 0x80402010u, __builtin_amdgcn_udot4(z[2], 0x08040201u, 0u, false), false);
    return (lo + (hi << 8)) >> 7;
}

// The 4 bytes starting at byte i (< 16) of the dword array o; i is a
// constant after unrolling, so o stays in registers.
__device__ __forceinline__ uint32_t bytes_at(const uint32_t* o, int i)
{
    return __builtin_amdgcn_alignbyte(o[(i >> 2) + 1], o[i >> 2], i & 3);
}

// Positions j < 16 where byte j equals byte j - d (1 <= d <= 8).  X[0..1] are
// the 8 bytes before the lane, X[2..] the lane's own bytes.
__device__ __forceinline__ uint32_t eq_mask(const uint32_t* X, int d)
{
    static_assert(kLane == 16, "four dwords per lane");
    uint32_t x[kLane / 4];
#pragma unroll
    for (int k = 0; k < kLane / 4; k++) {
        const uint32_t prev = d <= 4 ? __builtin_amdgcn_alignbyte(X[2 + k], X[1 + k], 4 - d)
                                     : __builtin_amdgcn_alignbyte(X[1 + k], X[k], 8 - d);
        x[k] = X[2 + k] ^ prev;
    }
    return zero_bytes16(x);
}

// Fixed-Huffman codes (RFC 1951 3.2.6), bit-reversed for the LSB-first
// stream, as code | nbits << 26.  Literal b: 8-bit codes 0x30+b (b < 144),
// 9-bit 0x190+b-144.
__device__ __forceinline__ uint32_t lit_tok(uint32_t b)
{
    const bool hi = b >= 144;
    return (__builtin_bitreverse32(b + (hi ? 0x100u : 0x30u)) >> (hi ? 23 : 24)) | ((hi ? 9u : 8u) << 26);
}

// Length symbol (257..285) of a match length 3..258 and its extra bits
// (RFC 1951 3.2.5): 257 + m for m = L - 3 < 8, then four symbols per extra
// bit count e with base (4 + i) << e.
__device__ __forceinline__ uint32_t huff_len_sym(uint32_t L)
{
    if (L == 258) return 285;
    const uint32_t m = L - 3;
    if (m < 8) return 257 + m;
    const uint32_t e = 29u - __builtin_clz(m);          // floor(log2 m) - 2
    return 257 + 4 * e + (m >> e);
}
__device__ __forceinline__ uint32_t huff_len_extra_bits(uint32_t L)
{
    const uint32_t m = L - 3;
    return (L == 258 || m < 8) ? 0u : 29u - __builtin_clz(m);
}
// Distance code (0..29) of a distance 1..32768: with dm = d - 1 and
// e = max(0, floor(log2 dm) - 1) extra bits, code = 2e + (dm >> e).
__device__ __forceinline__ uint32_t huff_dist_sym(uint32_t d)
{
    const uint32_t dm = d - 1;
    const uint32_t e = 30u - __builtin_clz(dm | 2u);
    return (e << 1) + (dm >> e);
}
constexpr uint32_t kNoToken = 0x7FFFFFFFu;   // (a parse position where no token starts)
// In the dynamic-Huffman token slots, a position where no token starts: an
// index of the emit table whose entry is zero (no bits).
constexpr uint32_t kTokNone = (uint32_t)kHuffTab;

// A match as a key: len << 16 | (0xFFFF - dist).  The larger key is the
// longer match, and of equal lengths the shorter distance.
__device__ __forceinline__ uint32_t mkey(uint32_t len, uint32_t dist) { return (len << 16) | (0xFFFFu - dist); }

// Smallest key of a match that is cheaper than its literals: lengths of 4 and
// more always are (at most 7 + 1 + 5 + 13 = 26 bits against 32), length 3
// only up to distance 16384 (12 + 12 bits; beyond, 13 extra bits make 25).
constexpr uint32_t kMinKey = (3u << 16) | (0xFFFFu - 16384u);

// The token of match key k (length 3..18, distance 1..32768): length symbols
// 257..268 (7-bit codes, one extra bit from length 11), then the 5-bit
// distance code and its extra bits.  Distance code of d: with dm = d - 1 and
// e = max(0, floor(log2 dm) - 1) extra bits, code = 2e + (dm >> e).
// The length part is one byte, looked up by v_perm in two 8-byte tables
// (lengths 3..10: the reversed 7-bit codes of symbols 257..264; lengths
// 11..18: those of 265..268 with the extra bit above them); selector bytes
// 0x0C give zero bytes, so the result is that byte alone.
// The length part of a match of length L (3..258), its fixed code (7 bits
// for symbols 257..279, 8 for 280..285) and extra bits, as code | nbits << 26.
__device__ __forceinline__ uint32_t long_len_tok(uint32_t L)
{
    const uint32_t sym = huff_len_sym(L), e = huff_len_extra_bits(L);
    const uint32_t ex = (L - 3u) & ((1u << e) - 1u);
    const bool b8 = sym >= 280;
    const uint32_t code = b8 ? __builtin_bitreverse32(0xC0u + sym - 280u) >> 24 : __builtin_bitreverse32(sym - 256u) >> 25;
    const uint32_t nb = b8 ? 8u : 7u;
    return (code | (ex << nb)) | ((nb + e) << 26);
}

// The distance part of match key k: the 5-bit code and its extra bits.
__device__ __forceinline__ uint32_t dist_tok(uint32_t k)
{
    const uint32_t dm = 0xFFFEu - (k & 0xFFFFu);
    const uint32_t deb = 30u - __builtin_clz(dm | 2u);
    const uint32_t dc = (deb << 1) + (dm >> deb);
    const uint32_t dex = dm & ((1u << deb) - 1u);
    return ((__builtin_bitreverse32(dc) >> 27) | (dex << 5)) | ((5u + deb) << 26);
}

__device__ __forceinline__ uint32_t match_tok(uint32_t k)
{
    const uint32_t L = k >> 16;
    const bool lhi = L >= 11;
    const uint32_t lcode = lhi ? __builtin_amdgcn_perm(0x9818E868u, 0xA828C848u, L + (0x0C0C0C00u - 11u))
                               : __builtin_amdgcn_perm(0x08703050u, 0x10602040u, L + (0x0C0C0C00u - 3u));
    const uint32_t lnb = lhi ? 8u : 7u;
    const uint32_t dm = 0xFFFEu - (k & 0xFFFFu);
    const uint32_t deb = 30u - __builtin_clz(dm | 2u);
    const uint32_t dc = (deb << 1) + (dm >> deb);
    const uint32_t dex = dm & ((1u << deb) - 1u);
    const uint32_t dpart = (__builtin_bitreverse32(dc) >> 27) | (dex << 5);
    return lcode | (dpart << lnb) | ((lnb + 5 + deb) << 26);
}

// Bytes of the lane from position i equal to the KW dwords c, up to 4 * KW:
// the first differing bit over the dwords (v_ffbl gives ~0 for a zero
// dword, so min() carries the count over), in bytes.  i and KW are
// constants after unrolling.
__device__ __forceinline__ uint32_t cmp_run(const uint32_t* o, const uint32_t* c, int i, int KW)
{
    uint32_t r = 0;
#pragma unroll
    for (int k = 3; k >= 0; k--) {
        if (k >= KW) continue;
        const uint32_t x = bytes_at(o, i + 4 * k) ^ c[k];
        r = min((uint32_t)__builtin_ctzg(x, -1), 32u + r);
    }
    return r >> 3;
}

__device__ __forceinline__ void load_words(const uint8_t* p, uint32_t* c, int KW)
{
    typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
    typedef u32x3 u32x3_ua __attribute__((aligned(1)));
    if (KW == 4) { const u32x4 v = g128u(p); c[0] = v[0]; c[1] = v[1]; c[2] = v[2]; c[3] = v[3]; }
    else if (KW == 3) { const u32x3 v = *(R2P_GLOBAL const u32x3_ua*)p; c[0] = v[0]; c[1] = v[1]; c[2] = v[2]; }
    else if (KW == 2) { const uint64_t v = g64u(p); c[0] = (uint32_t)v; c[1] = (uint32_t)(v >> 32); }
    else c[0] = g32u(p);
}

// The hash of position i (a multiple of SS: only those are hashed).
template <int SS>
__device__ __forceinline__ uint32_t hash_at(const uint32_t* hh, int i) { return hh[i / SS]; }

// Best match key of every position of the lane (0 for a literal).
// Candidates: distance 1, bpp and one row up, whose runs come from the
// equality masks n* (bit j set where byte j differs from the byte that far
// back, or j is past the lane's end), and at every SS-th position
// one hashed earlier position -- T2's entry when it precedes the position,
// else T1's within the window.
// Positions go in groups of G so that a group's table reads, then its
// candidate loads, are in flight together; a scheduling barrier between
// groups bounds the registers.  Branch-free: an unusable hashed candidate
// loads from the window base and is then discarded.
template <int G, int SS>
__device__ __forceinline__ void lane_keys(uint32_t* key, const uint8_t* wb, const uint32_t* o, const uint32_t* hh,
                                          int lb, int le, uint32_t hmask, uint32_t n1, uint32_t nb, uint32_t nr,
                                          uint32_t kb, uint32_t kr, const uint32_t* T,
                                          uint32_t base1, uint32_t base2)
{
#pragma unroll
    for (int g0 = 0; g0 < kLane; g0 += G) {
        __builtin_amdgcn_sched_barrier(0);
        uint32_t C[G];
        bool hv[G];
#pragma unroll
        for (int u = 0; u < G; u++) {
            const int i = g0 + u;
            const uint32_t r1 = __builtin_ctz(n1 >> i), rb = __builtin_ctz(nb >> i), rr = __builtin_ctz(nr >> i);
            key[i] = max(max((r1 << 16) | (0xFFFFu - 1u), (rb << 16) | kb), (rr << 16) | kr);
            hv[u] = false;
            C[u] = 0;
            if (i % SS) continue;                   // (i is a constant: no code for these positions)
            const uint32_t h = hash_at<SS>(hh, i);
            const u32x2 e = *(const u32x2*)(T + 2 * h);                       // T1 and T2 entries, one read
            const uint32_t x2 = e[1] ^ base2;                                // position if the tag matches
            const uint32_t x1 = e[0] ^ base1;
            const uint32_t P = (uint32_t)(lb + i);
            const bool fv = x2 < P;                                          // earlier in this sub-block
            const bool tv = P - 1u - x1 < (uint32_t)kWindow;                 // 1 <= P - x1 <= 32768
            hv[u] = (fv | tv) & (((hmask >> i) & 1u) != 0);
            C[u] = hv[u] ? (fv ? x2 : x1) : 0u;
        }
        uint32_t c[G][4];
#pragma unroll
        for (int u = 0; u < G; u++)
            if ((g0 + u) % SS == 0) load_words(wb + C[u], c[u], min(kHashCmpDwords, (kLane - (g0 + u) + 3) / 4));
#pragma unroll
        for (int u = 0; u < G; u++) {
            const int i = g0 + u;
            if (i % SS) { key[i] = key[i] >= kMinKey ? key[i] : 0u; continue; }
            const uint32_t lh = min(cmp_run(o, c[u], i, min(kHashCmpDwords, (kLane - i + 3) / 4)), (uint32_t)max(le - i, 0));
            const uint32_t k = max(key[i], hv[u] ? mkey(lh, (uint32_t)(lb + i) - C[u]) : 0u);
            key[i] = k >= kMinKey ? k : 0u;
        }
    }
    __builtin_amdgcn_sched_barrier(0);
}

// Hides the values from the optimiser, so that expressions computed from
// them before a barrier are recomputed after it instead of being held in
// registers across it (which costs occupancy).
__device__ __forceinline__ void launder(uint32_t* v, int n)
{
#pragma unroll
    for (int i = 0; i < n; i++) asm volatile("" : "+v"(v[i]));
}

// The lane's bytes [p, p + 20) (the 4 after it feed the hashes of its last
// positions); in bounds by the buffers' tail slack.
__device__ __forceinline__ void load_own(const uint8_t* p, uint32_t* o)
{
    const u32x4 v = g128u(p);
    o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3];
    o[4] = g32u(p + 16);
}

// STATS: also accumulate per-slice match statistics (the strip form's
// LzStats); a separate instantiation so that the plain and batched forms do
// not carry their registers.  DYN: the parse's tokens and their symbol
// histogram go to memory (DeflateLaunch::tokens / tok_count / hist) for a
// dynamic-Huffman block per segment; no bits are emitted here.
template <int NT, bool STATS, bool DYN, int SS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_deflate(DeflateLaunch DL)
{
    constexpr int LANE = kLane;
    constexpr int SUB = NT * LANE;               // positions per sub-block
    constexpr int OBUF = (SUB * 9 + 31) / 32 + 8;
    constexpr int NW = NT / 64;
    // per hash h: T[2h] "T1", the most recent position before the sub-block
    // (tag: slice), and T[2h + 1] "T2", the first position inside the
    // sub-block (tag: ~sub-block) -- side by side, so a lookup is one read
    static_assert(kHash2Bits == kHashBits, "T1 and T2 share the index");
    __shared__ __attribute__((aligned(8))) uint32_t T[2 << kHashBits];
    __shared__ uint32_t obuf[OBUF];
    __shared__ uint32_t red[3 * NW + 64];
    __shared__ uint32_t st[4];
    __shared__ uint32_t next_slice;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    const uint32_t total = DL.counts ? DL.counts[1] : DL.total_slices;
    const int njobs = DL.counts ? (int)DL.counts[0] : DL.njobs;
    // T1 keeps the largest entry (atomicMax) and T2 the smallest (atomicMin),
    // so T1's tag counts up per slice and T2's counts down per sub-block.
    uint32_t tag1 = kTagMax, tag2 = 0;
    for (int i = tid; i < OBUF; i += NT) obuf[i] = 0;   // then kept zero by every flush
    // Slices are dealt to workgroups in runs per XCD (xcd_slice).  (Giving
    // each of the 8 XCDs one contiguous eighth of the slices cut the HBM reads
    // to the input size but ran 35% slower on a radiograph: neighbouring
    // slices cost alike, so the XCDs ended unevenly loaded.)  Persistent
    // workgroups with a queue take their first slice by blockIdx and the
    // rest from the queue's counter, so that the faster ones take more.
    for (uint32_t g0 = blockIdx.x; g0 < total;) {
        const uint32_t g = xcd_slice(g0, total);
        int j = 0;
        if (DL.jobs) {
            int lo = 0, hi = njobs - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (DL.slice_start[mid] <= g) lo = mid; else hi = mid - 1;
            }
            j = lo;
        }
        const DeflateJob J = DL.jobs ? DL.jobs[j] : DL.single;
        const int sl = (int)(g - (DL.jobs ? DL.slice_start[j] : 0u));
        const int spb = J.seg_bytes / J.slice_bytes;
        const int seg = lay_seg(sl, spb, J.half_from);
        const int nseg = J.len > 0 ? (int)((J.len + J.seg_bytes - 1) / J.seg_bytes) : 1;
        const uint8_t* d = J.data;
        const int64_t s0 = J.first + lay_start(sl, spb, J.half_from, J.slice_bytes);
        const int64_t s1 = min64(s0 + lay_bytes(sl, spb, J.half_from, J.slice_bytes), J.first + J.len);
        const int64_t dend = J.first + J.len;
        const bool last = (seg == nseg - 1) && J.final_block;
        const int64_t wlo = max64(J.dict_lo, s0 - (int64_t)kWindow);
        uint32_t* out = J.slice_out + lay_buf(sl, spb, J.half_from, J.slice_cap_dw, J.slice_cap_dw2);
        // Positions below are relative to wb = d + wlo (window base).
        const uint8_t* wb = d + wlo;
        const int s0_rel = (int)(s0 - wlo);
        const int s1_rel = (int)max64(s1 - wlo, s0_rel);
        const int end_rel = (int)min64(dend - wlo, 0x7FFFFFF0);
        const bool lz = J.lz77 != 0;
        const int bpp = J.bpp;
        const int dr = J.row_stride;

        __syncthreads();
        // a new slice: fresh tags (both tables cleared when they run out; a
        // slice has at most 64 KiB / SUB sub-blocks)
        if (tag1 >= kTagMax || tag2 <= (uint32_t)(65536 / SUB) + 1) {
            for (int i = tid; i < (1 << kHashBits); i += NT) { T[2 * i] = 0; T[2 * i + 1] = ~0u; }
            tag1 = 0;
            tag2 = kTagMax;
        }
        tag1++;
        if (tid < 4) st[tid] = 0;
        __syncthreads();
        if (lz) {
            // prime T1 with the (up to) 32 KiB before the slice: the window
            // positions P = 0 mod kPrimeStride, 8 per lane and pass
            static_assert(kPrimeStride == 8, "every other dword of 64 bytes");
            constexpr int PL = 64;
            constexpr int NP = (kWindow + NT * PL - 1) / (NT * PL);   // passes over the window
            for (int b0 = tid * PL; b0 < s0_rel; b0 += NP * NT * PL) {
                // 64 bytes from each pass's b (in bounds when b < s0_rel: the
                // slice and the tail slack follow), all loads in flight at once
                u32x4 v[NP][4];
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const int b = b0 + p * NT * PL;
#pragma unroll
                    for (int k = 0; k < 4; k++) v[p][k] = b < s0_rel ? g128u(wb + b + 16 * k) : (u32x4){0, 0, 0, 0};
                }
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const int bp = b0 + p * NT * PL;
                    // all eight positions inside the window with their 4 bytes
                    // (every lane but those at the window's edges): no test each
                    if (bp + PL <= s0_rel && bp + PL <= end_rel) {
#pragma unroll
                        for (int k = 0; k < 16; k += 2)
                            atomicMax(&T[2 * hash4(v[p][k >> 2][k & 3])], (tag1 << kPosBits) | (uint32_t)(bp + 4 * k));
                    } else {
#pragma unroll
                        for (int k = 0; k < 16; k += 2) {
                            const int P = bp + 4 * k;
                            if (P < s0_rel && P + 4 <= end_rel)
                                atomicMax(&T[2 * hash4(v[p][k >> 2][k & 3])], (tag1 << kPosBits) | (uint32_t)P);
                        }
                    }
                }
            }
        }

        // this thread's terms of the slice's Adler-32: its bytes' sum and
        // sum of (offset in the slice) * byte, mod 65521
        uint32_t ad_s = 0, ad_t = 0;

        // the segment's first slice opens its block: BFINAL | BTYPE=01 << 1
        const bool opens = sl == lay_first(seg, spb, J.half_from);
        uint32_t carry = opens ? (last ? 3u : 2u) : 0u;
        uint32_t carry_bits = opens ? 3u : 0u;
        int64_t out_dw = 0;
        uint32_t* hist = obuf;                     // DYN: the slice's histogram (obuf is zero)
        // The flush of a sub-block's bits waits for the next sub-block's first
        // barrier (or the slice end's), which then also orders it before the
        // next emission: one barrier fewer per sub-block.
        bool pending = false;                      // a sub-block's bits not yet flushed
        uint32_t pend_tot = 0;                     // their count (carry excluded)
        auto flush = [&]() {
            const uint32_t tb = carry_bits + pend_tot;
            const uint32_t full = tb >> 5;
            // the flush clears what it read, so the next emission finds obuf
            // zero; only thread 0 keeps the carry
            for (uint32_t i = tid; i < full; i += NT) { out[out_dw + i] = obuf[i]; obuf[i] = 0; }
            if (tid == 0) { carry = (tb & 31) ? obuf[full] : 0u; obuf[full] = 0; }
            carry_bits = tb & 31;
            out_dw += full;
            pending = false;
        };

        for (int q = s0_rel; q < s1_rel; q += SUB) {
            const int n = min(SUB, s1_rel - q);
            tag2--;                                // (obuf is all zero: the last flush cleared it)
            const int lb = q + tid * LANE;         // this lane's first position
            const int le = max(0, min(LANE, n - tid * LANE));
            uint32_t o[LANE / 4 + 1] = {0, 0, 0, 0, 0};
            if (le > 0) {
                load_own(wb + lb, o);
                uint32_t s16 = 0, w16 = 0;
#pragma unroll
                for (int k = 0; k < LANE / 4; k++) {
                    const uint32_t v = le >= LANE ? o[k] : o[k] & byte_mask(le - 4 * k);
                    const uint32_t sk = __builtin_amdgcn_udot4(v, 0x01010101u, 0u, false);
                    w16 = __builtin_amdgcn_udot4(v, 0x03020100u, w16, false) + (uint32_t)(4 * k) * sk;
                    s16 += sk;
                }
                ad_s += s16;                           // < 16 * 4080 per slice
                ad_t = (ad_t + (uint32_t)(lb - s0_rel) * s16 + w16) % 65521u;   // < 2^29 before the mod
            }
            // ---- A. hashes of every SS-th position (kept), the
            // sub-block's first of those positions per hash ------------------------
            // hashable positions: inside the lane, with 4 bytes before the end
            const int hl = min(le, end_rel - 3 - lb);
            const uint32_t hmask = hl <= 0 ? 0u : hl >= LANE ? 0xFFFFu : (1u << hl) - 1u;
            uint32_t hh[LANE / SS];
#pragma unroll
            for (int i = 0; i < LANE / SS; i++) hh[i] = 0;
            if (lz && le > 0) {
#pragma unroll
                for (int i = 0; i < LANE; i += SS) hh[i / SS] = hash4(bytes_at(o, i));
                if (hmask == 0xFFFFu) {            // every position hashable: no test each
#pragma unroll
                    for (int i = 0; i < LANE; i += SS)
                        atomicMin(&T[2 * hash_at<SS>(hh, i) + 1], (tag2 << kPosBits) | (uint32_t)(lb + i));
                } else {
#pragma unroll
                    for (int i = 0; i < LANE; i += SS)
                        if ((hmask >> i) & 1u)
                            atomicMin(&T[2 * hash_at<SS>(hh, i) + 1], (tag2 << kPosBits) | (uint32_t)(lb + i));
                }
            }
            __syncthreads();
            // ---- B. best match per position, greedy parse of the lane -----------
            if constexpr (!DYN) {
                if (pending) flush();
                if (tid == 0) obuf[0] = carry;
            }
            uint32_t key[LANE];
#pragma unroll
            for (int i = 0; i < LANE; i++) key[i] = 0;
            if (lz && le > 0) {
                uint32_t X[LANE / 4 + 2];
                uint64_t pr = 0;
                if (lb >= 8) pr = g64u(wb + lb - 8);
                else
                    for (int k = 8 - lb; k < 8; k++) pr |= (uint64_t)g8(wb + lb - 8 + k) << (8 * k);
                X[0] = (uint32_t)pr; X[1] = (uint32_t)(pr >> 32);
#pragma unroll
                for (int k = 0; k < LANE / 4; k++) X[2 + k] = o[k];
                // a distance reaching before the window base is no candidate
                const uint32_t m1 = eq_mask(X, 1) & (lb >= 1 ? ~0u : ~1u);
                uint32_t mb = 0, mr = 0;
                if (bpp > 1) mb = eq_mask(X, bpp) & (lb >= bpp ? ~0u : (~0u << (bpp - lb)));
                const bool row_ok = dr > 1 && dr != bpp && dr <= kWindow && lb >= dr;
                if (row_ok) {
                    const u32x4 r = g128u(wb + lb - dr);
                    uint32_t x[LANE / 4];
#pragma unroll
                    for (int k = 0; k < LANE / 4; k++) x[k] = o[k] ^ r[k];
                    mr = zero_bytes16(x);
                }
                // runs stop at the lane's end: past it every position "differs"
                const uint32_t cap = le >= LANE ? 0u : ~0u << le;
                lane_keys<R2P_MATCH_GROUP, SS>(key, wb, o, hh, lb, le, hmask, ~m1 | cap, ~mb | cap, ~mr | cap,
                                           0xFFFFu - (uint32_t)(bpp > 1 ? bpp : 1), 0xFFFFu - (uint32_t)(row_ok ? dr : 1),
                                           T, tag1 << kPosBits, tag2 << kPosBits);
            }
            // greedy parse of the lane: token starts (tmask)
            uint32_t tmask = 0;
            {
                uint32_t nextp = 0;
#pragma unroll
                for (int i = 0; i < LANE; i++) {
                    const uint32_t L = key[i] >> 16;   // 0: literal
                    const bool tok = ((uint32_t)i == nextp) & (i < le);
                    tmask |= (tok ? 1u : 0u) << i;
                    nextp = tok ? (uint32_t)i + max(L, 1u) : nextp;
                }
            }
            // ---- matches across lanes: a lane whose one token is a match of
            // all its 16 bytes at distance d, after a lane that is the same
            // (d included), continues that match: the lane heading such a run
            // absorbs up to 15 following lanes into its token (16 + 15 * 16 =
            // 256 <= 258 bytes), the next one heads a new run.  Runs stay
            // inside the wave.  (Flat areas and repeated rows become one token
            // per 256 bytes instead of one per 16.)
            const bool whole = tmask == 1u && (key[0] >> 16) == (uint32_t)LANE;
            const uint32_t kw = whole ? key[0] : 0u;
            const uint32_t kprev = (uint32_t)__shfl_up((int)kw, 1, 64);
            const bool cont = lane > 0 && whole && kprev != 0 && ((kprev ^ kw) & 0xFFFFu) == 0;
            const uint64_t C = __ballot(cont);
            const uint64_t below = lane ? (~C & ((1ull << lane) - 1ull)) : 0ull;
            const int head = below ? 63 - __builtin_clzll(below) : 0;       // lane heading my run (if cont)
            const bool absorbed = cont && ((lane - head) & 15) != 0;
            // the continuing lanes after a (sub)head that it absorbs: up to 15
            const uint64_t after = lane < 63 ? ~(C >> (lane + 1)) : 1ull;
            const uint32_t nabs = (whole && !absorbed) ? min(15u, (uint32_t)__builtin_ctzll(after)) : 0u;
            const uint32_t Lrun = (uint32_t)LANE + 16u * nabs;
            if (absorbed) tmask = 0;

            // each token's code (code | nbits << 26, 0 where no token starts;
            // a run's token as its length part, then its distance part in the
            // next slot) and the lane's bit count
            uint32_t tk[LANE];
            uint32_t nbits = 0;
            const uint32_t ext_len = DYN ? 0u : long_len_tok(Lrun);
            const uint32_t ext_dist = DYN ? 0u : dist_tok(kw);
#pragma unroll
            for (int i = 0; i < LANE; i++) {
                if ((i & 3) == 0) __builtin_amdgcn_sched_barrier(0);
                const bool tok = (tmask >> i) & 1u;
                const uint32_t b = (o[i >> 2] >> (8 * (i & 3))) & 0xFFu;
                if constexpr (DYN) {                   // raw token; nbits counts tokens
                    // (a match: 0x8000 | length - 3, distance - 1 in the high half)
                    const uint32_t L = (i == 0 && nabs) ? Lrun : key[i] >> 16;
                    tk[i] = !tok ? kTokNone : L ? 0x8000u | (L - 3u) | (0xFFFEu - (key[i] & 0xFFFFu)) << 16 : b;
                    nbits += tok ? 1u : 0u;
                } else {
                    const uint32_t L = key[i] >> 16;
                    uint32_t t = tok ? (L ? match_tok(key[i]) : lit_tok(b)) : 0u;
                    if (i == 0) t = nabs ? ext_len : t;
                    if (i == 1) t = nabs ? ext_dist : t;
                    tk[i] = t;
                    nbits += t >> 26;
                }
            }
            // ---- block scan of bit counts ----------------------------------------------
            const uint32_t incl = wave_incl_scan(nbits);
            if (lane == 63) red[wave] = incl;
            if (STATS && J.slice_stats) {          // match statistics (strip form only)
                uint32_t m_cnt = 0, m_bytes = 0, m_dist = 0, m_max = 0;
#pragma unroll
                for (int i = 0; i < LANE; i++) {
                    const uint32_t Lt = (tmask >> i) & 1u ? (nabs && i == 0 ? Lrun : key[i] >> 16) : 0u;
                    m_cnt += Lt ? 1u : 0u; m_bytes += Lt; m_dist += Lt ? 0xFFFFu - (key[i] & 0xFFFFu) : 0u;
                    m_max = max(m_max, Lt);
                }
                atomicAdd(&st[0], m_cnt); atomicAdd(&st[1], m_bytes);
                atomicAdd(&st[2], m_dist); atomicMax(&st[3], m_max);
            }
            __syncthreads();                       // every lookup of T1/T2 done, red complete
            uint32_t wbase = 0, tot = 0;
            for (int w2 = 0; w2 < NW; w2++) { if (w2 < wave) wbase += red[w2]; tot += red[w2]; }
            const uint32_t off = carry_bits + wbase + incl - nbits;
            // the sub-block's positions (every SS-th) enter T1; T2's
            // entries expire with its tag
            if (lz && le > 0) {
                if (hmask == 0xFFFFu) {
#pragma unroll
                    for (int i = 0; i < LANE; i += SS)
                        atomicMax(&T[2 * hash_at<SS>(hh, i)], (tag1 << kPosBits) | (uint32_t)(lb + i));
                } else {
#pragma unroll
                    for (int i = 0; i < LANE; i += SS)
                        if ((hmask >> i) & 1u) atomicMax(&T[2 * hash_at<SS>(hh, i)], (tag1 << kPosBits) | (uint32_t)(lb + i));
                }
            }
            if constexpr (DYN) {
                // ---- a 16-bit slot per position, and the symbols' histogram ----
                // A slot holds a literal, 0x8000 | length - 3 where a match
                // starts with its distance - 1 in the next slot (a match never
                // starts at a lane's last position), kTokNone elsewhere.  The
                // slot pairs (16l + 2k, 16l + 2k + 1) of each 4096-position unit
                // of the slice go to dword 4(l + 256q) + j, k = 4q + j: every
                // 16-byte store instruction covers contiguous lanes, and so
                // does every load of k_huff_emit.
                static_assert(LANE == 16, "the token layout assumes 16 positions per lane");
                if (le > 0) {
                    uint32_t d[LANE / 2];
#pragma unroll
                    for (int k = 0; k < LANE / 2; k++) {
                        const uint32_t lo = 2 * k > 0 && (tk[2 * k - 1] & 0x8000u) ? tk[2 * k - 1] >> 16 : tk[2 * k] & 0xFFFFu;
                        const uint32_t hi = (tk[2 * k] & 0x8000u) ? tk[2 * k] >> 16 : tk[2 * k + 1] & 0xFFFFu;
                        d[k] = lo | hi << 16;
                    }
                    const int p0 = lb - s0_rel;
                    uint32_t* tp = J.tokens + (s0 - J.first + (p0 & ~4095)) / 2 + 4 * ((p0 >> 4) & 255);
#pragma unroll
                    for (int q = 0; q < 2; q++)
                        *(u32x4*)(tp + 1024 * q) = u32x4{d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]};
                }
#pragma unroll
                for (int i = 0; i < LANE; i++) {
                    if (tk[i] == kTokNone) continue;
                    if (tk[i] & 0x8000u) {
                        atomicAdd(&hist[huff_len_sym((tk[i] & 0xFFu) + 3u)], 1u);
                        atomicAdd(&hist[286 + huff_dist_sym((tk[i] >> 16) + 1u)], 1u);
                    } else {
                        atomicAdd(&hist[tk[i]], 1u);
                    }
                }
                __syncthreads();                   // the sub-block's T1/T2 and red uses done
                continue;
            }
            // ---- emission ------------------------------------------------------------------
            // A lane's first word may hold the previous lane's last bits and
            // its last partial word the next lane's first ones: those two are
            // ORed in (obuf starts zeroed, each bit has one writer); every
            // other word the lane completes is its own and is stored.
            {
                const uint32_t w0 = off >> 5;
                uint32_t wi = w0, nacc = off & 31;
                uint64_t acc = 0;
#pragma unroll
                for (int i = 0; i < LANE; i++) {
                    acc |= (uint64_t)(tk[i] & 0x3FFFFFFu) << nacc;
                    nacc += tk[i] >> 26;
                    if (nacc >= 32) {
                        if (wi == w0) atomicOr(&obuf[wi], (uint32_t)acc);
                        else obuf[wi] = (uint32_t)acc;
                        wi++;
                        acc >>= 32;
                        nacc -= 32;
                    }
                }
                if (nacc > 0) atomicOr(&obuf[wi], (uint32_t)acc);
            }
            pending = true;
            pend_tot = tot;
        }

        // ---- slice end: the partial last word and the bit count ---------------
        __syncthreads();
        if constexpr (!DYN) {
            if (pending) flush();
        }
        if constexpr (DYN) {
            for (int i = tid; i < kHuffSyms; i += NT) {    // into the segment's histogram, obuf zero again
                if (hist[i]) atomicAdd(&J.hist[(int64_t)seg * kHuffSyms + i], hist[i]);
                hist[i] = 0;
            }
            if (tid == 0) J.tok_count[sl] = (uint32_t)(s1 - s0);
        }
        if (!DYN && tid == 0) {
            if (carry_bits) out[out_dw] = carry;
            J.slice_bits[sl] = (uint32_t)(out_dw * 32 + carry_bits);
            if (STATS && J.slice_stats) {
                uint32_t* ss = J.slice_stats + 4 * (int64_t)sl;
                ss[0] = st[0]; ss[1] = st[1]; ss[2] = st[2]; ss[3] = st[3];
            }
        }
        const uint32_t ad = adler_finish(ad_s, ad_t, s1 - s0, red);
        if (tid == 0) J.slice_adler[sl] = ad;
        if (DL.queue) {
            // (the loop's first barrier keeps this write after every read of
            // the previous value)
            if (tid == 0) next_slice = gridDim.x + atomicAdd(DL.queue, 1u);
            __syncthreads();
            g0 = next_slice;
        } else {
            g0 += gridDim.x;
        }
    }
}

// ---------------------------------------------------------------------------
// Dynamic-Huffman blocks (RFC 1951 3.2.7), one per segment
//
// k_huff_tables: one workgroup per segment reads its histogram (summed by the
// slices' atomics), builds length-limited codes for the literal/length and
// distance alphabets and for the code-length alphabet (the lengths are sent
// one symbol each, without run-length codes: a little more header for a
// header written in parallel), and writes the emit table and the block
// header -- or the fixed code and its 3-bit header when that is no longer
// (so a segment never grows past its fixed-Huffman size).
// k_huff_emit: one workgroup per slice turns its token slots into bits with
// the segment's table (the segment's first slice starts with the header, its
// last ends with the end-of-block code), in rounds of 16 slots per thread:
// per-slot code lookups, a block scan of the bit counts, ors into LDS words,
// flush.
// ---------------------------------------------------------------------------

constexpr int kHuffThreads = 256;        // k_huff_tables workgroup

// Sum of v over the workgroup (all threads call it); red holds a u32 per wave.
__device__ uint32_t block_sum(uint32_t v, uint32_t* red)
{
    v = wave_total(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
    __syncthreads();
    return t;
}

// Ors the n (<= 28) bits of v into LDS words at bit offset off: one 64-bit
// or, a second one when the bits cross into the next 8-byte word.
__device__ __forceinline__ void put_bits64(uint32_t* obuf, uint32_t v, uint32_t n, uint32_t off)
{
    const uint32_t w = off >> 6, sh = off & 63;
    unsigned long long* o64 = (unsigned long long*)obuf;
    atomicOr(&o64[w], (unsigned long long)v << sh);
    if (sh + n > 64) atomicOr(&o64[w + 1], (unsigned long long)(v >> (64 - sh)));
}
// Ors the n (<= 28) bits of v into LDS words at bit offset off.
__device__ __forceinline__ void put_bits(uint32_t* obuf, uint32_t v, uint32_t n, uint32_t off)
{
    const uint32_t w = off >> 5, sh = off & 31;
    atomicOr(&obuf[w], v << sh);
    if (sh + n > 32) atomicOr(&obuf[w + 1], v >> (32 - sh));
}

// Maximum of v over the 64 lanes of the wave (the DPP pattern of wave_total), uniform.
__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// The order key of symbol s with frequency f > 0: its Shannon length
// L = ceil(log2(T / f)) clamped to [1, LIMIT], packed as f << L (< 2T < 2^18),
// then 15 - L, then 511 - s, so that the larger key is the larger f * 2^L
// (bits saved per unit of Kraft sum by a shortening), then the larger f, then
// the lower symbol.  L comes back as 15 - ((key >> 9) & 15).
template <int LIMIT>
__device__ __forceinline__ uint32_t shannon_key(uint32_t f, uint32_t T, int s)
{
    const uint32_t q = (T + f - 1) / f;
    const uint32_t L = min(max(q <= 1 ? 1u : 32u - __builtin_clz(q - 1), 1u), (uint32_t)LIMIT);
    return (f << L) << 13 | (15u - L) << 9 | (511u - (uint32_t)s);
}
__device__ __forceinline__ uint32_t key_len(uint32_t k) { return 15u - ((k >> 9) & 15u); }

constexpr int kHuffChunks = 5;           // 64-symbol chunks of the largest alphabet (286)
constexpr int kHuffPhases = 2;           // shortening passes of the length construction

// The length construction's second half, by one wave, on the symbols in
// registers in the order of their keys: lane p of chunk c holds the length
// L[c] of the symbol of order 64c + p (0: none), used symbols in all.  The
// lengths start at the Shannon lengths against a T that keeps the Kraft sum
// <= 1 with the clamp.
//  1. kHuffPhases walks of the order: a symbol is shortened by one bit when
//     its term fits in the remaining slack, and stays in the next walk only
//     if it was.  A walk is a few rounds of 64-lane prefix sums: the lanes
//     before the first one that does not fit all move, the slack drops below
//     that lane's term, and the round repeats from the next lane with the
//     smaller terms only.
//  2. Any slack left is a multiple of the longest length's term: codes of the
//     longest length move up one bit (first in the order first) until the
//     sum is exactly 1 -- inflate takes complete codes only.  (Each step
//     leaves a sum that is a multiple of the next shorter length's term, and
//     all codes at length 1 would sum to >= 1.)
// On the images measured the result is within 0.4 % of the optimal
// (two-queue Huffman) code.
template <int LIMIT, int NC>
__device__ __forceinline__ void walk_lengths(uint32_t (&L)[NC], uint32_t used)
{
    constexpr uint32_t U = 1u << LIMIT;                       // the Kraft sum in units of 2^-LIMIT
    const int lane = threadIdx.x & 63;
    bool el[NC];
    uint32_t kk = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        el[c] = L[c] > 1;
        kk += L[c] ? U >> L[c] : 0u;
    }
    uint32_t S = U - wave_total(kk);                          // the slack
    for (int ph = 0; ph < kHuffPhases && S; ph++) {
        uint32_t bound = S;                                   // the largest term that can still fit
#pragma unroll
        for (int c = 0; c < NC; c++) {
            bool moved = false;
            int start = 0;
            while (S && start < 64 && c * 64 < (int)used) {
                const bool adm = el[c] && lane >= start && (U >> L[c]) <= bound;
                const uint32_t incl = wave_incl_scan(adm ? U >> L[c] : 0u);
                const uint64_t fail = __ballot(adm && incl > S);
                const int f = fail ? (int)__builtin_ctzll(fail) : 64;
                const bool take = adm && lane < f;
                L[c] -= take;
                moved |= take;
                S -= f ? (uint32_t)__builtin_amdgcn_readlane((int)incl, f - 1) : 0u;
                bound = S;
                start = f + 1;
            }
            el[c] = moved && L[c] > 1;
        }
    }
    uint32_t lmax = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) lmax = max(lmax, L[c]);
    lmax = S ? wave_max(lmax) : 0u;
    for (int l = (int)lmax; l > 1 && S; l--) {
        const uint32_t want = S >> (LIMIT - l);
        uint32_t before = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (c * 64 >= (int)used) break;
            const uint64_t bm = __ballot(L[c] == (uint32_t)l);
            const uint32_t idx = before + (uint32_t)__builtin_popcountll(bm & ((1ull << lane) - 1ull));
            if (L[c] == (uint32_t)l && idx < want) L[c] -= 1;
            before += (uint32_t)__builtin_popcountll(bm);
        }
        S -= min(before, want) << (LIMIT - l);
    }
}

// Code lengths of at most LIMIT bits for an alphabet of N <= 63 symbols held
// by one wave (symbol = lane, f its frequency; 0 past N); returns the lane's
// length.  At least two symbols get a code (a used one alone, or none, gets
// partners of count 1).  The Shannon lengths are taken against
// T = ceil(total * U / (U - m)) (U = 2^LIMIT, m = symbols with
// f * U < 2 * total, which bounds the clamped ones), so the Kraft sum stays
// <= 1 with the clamp: clamped terms add at most m / U, the others at most
// total / T = (U - m) / U.  The order is a rank among the wave's keys; lane p
// takes the symbol of rank p (a forward permute) for walk_lengths, and each
// symbol reads its length back from the lane of its rank.
template <int LIMIT, int N>
__device__ uint32_t wave_lengths(uint32_t f)
{
    constexpr uint32_t U = 1u << LIMIT;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t used = (uint32_t)__builtin_popcountll(__ballot(f != 0));
    if (used < 2) {
        const uint64_t z = __ballot(f == 0 && lane < N);
        if (f == 0 && lane < N && (uint32_t)__builtin_popcountll(z & lt) < 2u - used) f = 1;
        used = 2;
    }
    const uint32_t total = wave_total(f);
    const uint32_t m = (uint32_t)__builtin_popcountll(__ballot(f != 0 && (uint64_t)f * U < 2ull * total));
    const uint32_t T = (uint32_t)(((uint64_t)total * U + (U - m - 1)) / (U - m));
    const uint32_t key = f ? shannon_key<LIMIT>(f, T, lane) : 0u;
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < N; j++) r += (uint32_t)__builtin_amdgcn_readlane((int)key, j) > key;
    // unused symbols all have rank `used` (< 64): that lane is not read
    const uint32_t v = (uint32_t)__builtin_amdgcn_ds_permute((int)(r * 4), (int)(f ? key_len(key) : 0u));
    uint32_t L[1] = {(uint32_t)lane < used ? v : 0u};
    walk_lengths<LIMIT, 1>(L, used);
    const uint32_t mine = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(r * 4), (int)L[0]);
    return f ? mine : 0u;
}

// Canonical code (RFC 1951 3.2.2) of a wave-held alphabet (symbol = lane,
// l its length, 0: unused), bit-reversed for the LSB-first stream:
// code | l << 16.  A symbol's code is its length's first code plus the
// number of lower symbols of that length (one ballot per length).
__device__ uint32_t wave_codes(uint32_t l)
{
    const int lane = threadIdx.x & 63;
    uint64_t mine = 0;
    uint32_t bc = 0;                                           // lane b: codes of length b
#pragma unroll
    for (int b = 1; b < 16; b++) {
        const uint64_t m = __ballot(l == (uint32_t)b);
        if (lane == b) bc = (uint32_t)__builtin_popcountll(m);
        mine = l == (uint32_t)b ? m : mine;
    }
    uint32_t code = 0;                                         // lane b: the first code of length b
#pragma unroll
    for (int k = 1; k < 15; k++) {
        const uint32_t ck = (uint32_t)__builtin_amdgcn_readlane((int)bc, k);
        code += lane > k && lane < 16 ? ck << (lane - k) : 0u;
    }
    const uint32_t fl = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(l * 4), (int)code);
    const uint32_t c = fl + (uint32_t)__builtin_popcountll(mine & ((1ull << lane) - 1ull));
    return l ? (__builtin_bitreverse32(c) >> (32 - l)) | (l << 16) : 0u;
}

// The job of global item g (a segment or slice) and g's index inside it:
// the last job whose start is <= g.
__device__ __forceinline__ int huff_job_of(const HuffLaunch& HL, const uint32_t* start, uint32_t g, int njobs)
{
    if (!HL.jobs) return 0;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Fixed-code length of literal/length symbol s (RFC 1951 3.2.6).
__device__ __forceinline__ uint32_t fixed_len(int s) { return s < 144 ? 8u : s < 256 ? 9u : s < 280 ? 7u : 8u; }

// The order in which a dynamic header sends the code-length code's lengths (RFC 1951 3.2.7).
__device__ const uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

// The code tables and block header of segment seg of job H, by the workgroup.
__device__ __forceinline__ void huff_tables_segment(const HuffJob& H, int seg)
{
    constexpr int NW = kHuffThreads / 64;
    __shared__ uint32_t freq[kHuffSyms];
    __shared__ uint8_t len[kHuffSyms];
    __shared__ uint32_t slot[288];
    __shared__ uint32_t red[16], first[16], cnt[kHuffChunks * 16];
    __shared__ __attribute__((aligned(16))) uint32_t bcnt[256], bstart[256];
    __shared__ uint32_t cfreq[19], table[kHuffSyms];
    __shared__ uint32_t hdr[kHuffHdrWords];
    __shared__ uint32_t s_hlit, s_hdist, s_dynamic;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool bfinal = seg == H.nseg - 1 && H.final_block;
    // the segment's histogram (its slices' atomic sums), left zero for the
    // next run
    static_assert(kHuffSyms <= 2 * kHuffThreads, "two symbols per thread");
    {
        uint32_t* hp = H.hist + (int64_t)seg * kHuffSyms + tid;
        const bool two = tid + kHuffThreads < kHuffSyms;
        const uint32_t f0 = hp[0], f1 = two ? hp[kHuffThreads] : 0u;
        hp[0] = 0;
        if (two) hp[kHuffThreads] = 0;
        freq[tid] = f0;
        if (two) freq[tid + kHuffThreads] = f1 + (tid + kHuffThreads == 256 ? 1u : 0u);   // the end-of-block code
    }
    for (int i = tid; i < 256; i += kHuffThreads) bcnt[i] = 0;
    for (int i = tid; i < 19; i += kHuffThreads) cfreq[i] = 0;
    for (int i = tid; i < kHuffHdrWords; i += kHuffThreads) hdr[i] = 0;
    __syncthreads();
    // literal/length code: counts, keys and ranks by the workgroup (at least
    // two symbols: the end-of-block code always has one); T as in
    // wave_lengths with m bounded by the used symbols
    uint32_t used, total;
    for (;;) {
        uint32_t v = 0;
        for (int s = tid; s < 286; s += kHuffThreads) v += freq[s] + (freq[s] ? 1u << 23 : 0u);
        v = block_sum(v, red);                                // total < 2^23
        used = v >> 23;
        total = v & ((1u << 23) - 1);
        if (used >= 2) break;
        if (tid == 0)
            for (int i = 0; used < 2 && i < 286; i++)
                if (freq[i] == 0) { freq[i] = 1; used++; }
        __syncthreads();
    }
    const uint32_t T = (uint32_t)(((uint64_t)total * 32768u + (32768u - used - 1)) / (32768u - used));
    // The order of the keys by a bucket sort: 256 buckets by f << L
    // (< 2T), descending; a bucket's start from a scan of the counts; a
    // symbol's rank is its bucket's start plus the keys of its bucket above
    // its own.  Symbols tid and tid + 256.
    uint32_t kk[2] = {0, 0}, bk[2] = {0, 0}, bi[2] = {0, 0};
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int s = tid + kHuffThreads * i;
        if (s >= 286) continue;
        const uint32_t f = freq[s];
        len[s] = 0;
        if (!f) continue;
        kk[i] = shannon_key<15>(f, T, s);
        bk[i] = 255u - min(255u, ((kk[i] >> 13) * 128u) / T);
        bi[i] = atomicAdd(&bcnt[bk[i]], 1u);
    }
    __syncthreads();
    if (wave == 0) {
        const uint4 c = *(const uint4*)&bcnt[4 * lane];
        const uint32_t t4 = c.x + c.y + c.z + c.w, ex = wave_incl_scan(t4) - t4;
        *(uint4*)&bstart[4 * lane] = uint4{ex, ex + c.x, ex + c.x + c.y, ex + c.x + c.y + c.z};
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; i++)
        if (kk[i]) table[bstart[bk[i]] + bi[i]] = kk[i];        // (table is free until the codes)
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; i++) {
        if (!kk[i]) continue;
        // (equal counts share a bucket: up to a few hundred keys, read 8 at a
        // time so that the loads are in flight together)
        const uint32_t b0 = bstart[bk[i]], b1 = b0 + bcnt[bk[i]];
        uint32_t r = b0, j = b0;
        for (; j + 8 <= b1; j += 8) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = table[j + u];
#pragma unroll
            for (int u = 0; u < 8; u++) r += v[u] > kk[i];
        }
        for (; j < b1; j++) r += table[j] > kk[i];
        slot[r] = (uint32_t)(tid + kHuffThreads * i) | key_len(kk[i]) << 16;
    }
    __syncthreads();
    if (wave == 0) {
        // the literal/length lengths: walk in rank order, lengths out, HLIT
        uint32_t sy[kHuffChunks], L[kHuffChunks];
#pragma unroll
        for (int c = 0; c < kHuffChunks; c++) {
            const int p = c * 64 + lane;
            const uint32_t v = p < (int)used ? slot[p] : 0u;
            sy[c] = v & 0xFFFFu;
            L[c] = v >> 16;
        }
        walk_lengths<15, kHuffChunks>(L, used);
        uint32_t mx = 0;
#pragma unroll
        for (int c = 0; c < kHuffChunks; c++)
            if (L[c]) { len[sy[c]] = (uint8_t)L[c]; mx = max(mx, sy[c] + 1); }
        mx = wave_max(mx);
        if (lane == 0) s_hlit = max(257u, mx);
    } else if (wave == 1) {
        // the distance code, by this wave alone: lengths, codes, HDIST
        const uint32_t l = wave_lengths<15, 30>(lane < 30 ? freq[286 + lane] : 0u);
        const uint32_t c = wave_codes(l);
        if (lane < 30) { len[286 + lane] = (uint8_t)l; table[286 + lane] = c; }
        const uint32_t mx = wave_max(l ? (uint32_t)lane + 1 : 0u);
        if (lane == 0) s_hdist = max(1u, mx);
    }
    __syncthreads();
    const int hlit = (int)s_hlit, hdist = (int)s_hdist, N = hlit + hdist;
    // code sizes of this code and of the fixed one (extra bits are the same
    // for both); the code-length histogram (the lengths are sent one symbol
    // each, no run-length codes); the literal/length code's counts per length
    {
        uint32_t dyn = 0, fix = 0;
        for (int s = tid; s < kHuffSyms; s += kHuffThreads) {
            const uint32_t f = freq[s], l = len[s];
            dyn += f * l;
            fix += f * (s < 286 ? fixed_len(s) : 5u);
        }
        for (int j = tid; j < N; j += kHuffThreads) atomicAdd(&cfreq[j < hlit ? len[j] : len[286 + j - hlit]], 1u);
        dyn = wave_total(dyn);
        fix = wave_total(fix);
        if (lane == 0) { red[4 + wave] = dyn; red[8 + wave] = fix; }
    }
    uint32_t myl[2] = {0, 0}, below[2] = {0, 0};
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int ch = wave + NW * i;
        if (ch >= kHuffChunks) continue;
        const int s = ch * 64 + lane;
        const uint32_t l = s < 286 ? len[s] : 0u;
        uint64_t mine = 0;
#pragma unroll
        for (int b = 1; b < 16; b++) {
            const uint64_t m = __ballot(l == (uint32_t)b);
            if (lane == b) cnt[ch * 16 + b] = (uint32_t)__builtin_popcountll(m);
            mine = l == (uint32_t)b ? m : mine;
        }
        myl[i] = l;
        below[i] = (uint32_t)__builtin_popcountll(mine & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if (wave == 0) {
        {   // the literal/length code's first codes per length
            uint32_t bc = 0;
            if (lane >= 1 && lane < 16)
                for (int ch = 0; ch < kHuffChunks; ch++) bc += cnt[ch * 16 + lane];
            uint32_t code = 0;
#pragma unroll
            for (int k = 1; k < 15; k++) {
                const uint32_t ck = (uint32_t)__builtin_amdgcn_readlane((int)bc, k);
                code += lane > k && lane < 16 ? ck << (lane - k) : 0u;
            }
            if (lane < 16) first[lane] = code;
        }
        // the code-length code, its order of lengths, HCLEN
        const uint32_t cl = wave_lengths<7, 19>(lane < 19 ? cfreq[lane] : 0u);
        const uint32_t cc = wave_codes(cl);
        const uint32_t clo = (uint32_t)__builtin_amdgcn_ds_bpermute(lane < 19 ? 4 * (int)kClOrder[lane] : 0, (int)cl);
        const uint64_t nz = __ballot(lane < 19 && clo != 0);
        const int hclen = max(4, nz ? 64 - (int)__builtin_clzll(nz) : 0);
        // the code lengths' codes and bits
        uint32_t lc[5], hb = 0;
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const int j = 64 * i + lane;
            const uint32_t l = j < N ? len[j < hlit ? j : 286 + j - hlit] : 0u;
            const uint32_t c = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(4 * l), (int)cc);
            lc[i] = j < N ? c : 0u;
            hb += lc[i] >> 16;
        }
        hb = wave_total(hb);
        const uint32_t dyn = red[4] + red[5] + red[6] + red[7];
        const uint32_t fix = red[8] + red[9] + red[10] + red[11] + 3;
        const uint32_t fixed_part = 3 + 14 + 3 * (uint32_t)hclen;
        const bool dynamic = dyn + fixed_part + hb < fix;
        if (dynamic) {
            if (lane == 0)
                atomicOr(&hdr[0], (bfinal ? 1u : 0u) | (2u << 1) | (uint32_t)(hlit - 257) << 3 |
                                      (uint32_t)(hdist - 1) << 8 | (uint32_t)(hclen - 4) << 13);
            if (lane < hclen) put_bits(hdr, clo, 3, 17 + 3 * (uint32_t)lane);
            uint32_t base = fixed_part;
#pragma unroll
            for (int i = 0; i < 5; i++) {
                const uint32_t nb = lc[i] >> 16;
                const uint32_t incl = wave_incl_scan(nb);
                if (nb) put_bits(hdr, lc[i] & 0xFFFFu, nb, base + incl - nb);
                base += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            }
            if (lane == 0) H.hdr_bits[seg] = base;
        } else if (lane == 0) {
            hdr[0] = (bfinal ? 1u : 0u) | (1u << 1);
            H.hdr_bits[seg] = 3;
        }
        if (lane == 0) s_dynamic = dynamic;
    }
    __syncthreads();
    if (s_dynamic) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int ch = wave + NW * i;
            const int s = ch * 64 + lane;
            if (ch >= kHuffChunks || s >= 286) continue;
            const uint32_t l = myl[i];
            uint32_t r = below[i];
            for (int c2 = 0; c2 < ch; c2++) r += l ? cnt[c2 * 16 + l] : 0u;
            table[s] = l ? (__builtin_bitreverse32(first[l] + r) >> (32 - l)) | (l << 16) : 0u;
        }
    } else {
        // the fixed code (RFC 1951 3.2.6): its canonical codes directly
        for (int s = tid; s < 286; s += kHuffThreads) {
            const uint32_t l = fixed_len(s);
            const uint32_t c = s < 144 ? 0x30u + s : s < 256 ? 0x190u + (s - 144) : s < 280 ? (uint32_t)(s - 256) : 0xC0u + (s - 280);
            table[s] = (__builtin_bitreverse32(c) >> (32 - l)) | (l << 16);
        }
        for (int d = tid; d < 30; d += kHuffThreads) table[286 + d] = (__builtin_bitreverse32((uint32_t)d) >> 27) | (5u << 16);
    }
    __syncthreads();
    for (int i = tid; i < kHuffHdrWords; i += kHuffThreads) H.hdr[(int64_t)seg * kHuffHdrWords + i] = hdr[i];
    // the emit table: value | bit count << 24 per symbol, then per match
    // length 3..258 its code with the extra bits already in place
    uint32_t* C = H.codes + (int64_t)seg * kHuffTab;
    for (int s = tid; s < kHuffSyms; s += kHuffThreads) C[s] = (table[s] & 0xFFFFu) | (table[s] >> 16) << 24;
    for (int k = tid; k < 256; k += kHuffThreads) {
        const uint32_t L = (uint32_t)k + 3u, lc = table[huff_len_sym(L)], le = huff_len_extra_bits(L), ln = lc >> 16;
        C[kHuffSyms + k] = ((lc & 0xFFFFu) | ((L - 3u) & ((1u << le) - 1u)) << ln) | (ln + le) << 24;
    }
}

// One workgroup per segment (the single job: the grid is its segments), or
// workgroups taking the job table's segments in turn.
template <bool JOBS>
__global__ __launch_bounds__(kHuffThreads) void k_huff_tables(HuffLaunch HL)
{
    if (!JOBS) {
        huff_tables_segment(HL.single, blockIdx.x);
        return;
    }
    const uint32_t total = HL.counts ? HL.counts[1] : HL.total_segs;
    const int njobs = HL.counts ? (int)HL.counts[0] : HL.njobs;
    for (uint32_t gs = blockIdx.x; gs < total; gs += gridDim.x) {
        const int j = huff_job_of(HL, HL.seg_start, gs, njobs);
        __syncthreads();                                      // the previous segment is done with the LDS
        huff_tables_segment(HL.jobs[j], (int)(gs - HL.seg_start[j]));
    }
}

constexpr int kEmitThreads = 256;
constexpr int kEmitPer = 16;                                   // positions per thread and round
// a round's bits: at most 16 per position (a literal's 15, a match's 48 over
// >= 3 positions; one match may run past the round), the carry and the header
constexpr int kEmitWords = kEmitThreads * kEmitPer * 16 / 32 + kHuffHdrWords + 8;
static_assert(kEmitThreads * kEmitPer == 4096, "a round is one unit of k_deflate's token layout");

// One workgroup per slice: rounds of 4096 positions (one unit of the slot
// layout k_deflate writes), thread e's 16 slots (positions 16e..) read as two
// contiguous-lane 16-byte loads issued one round ahead, each slot's code
// looked up (a literal's, a match length's with its extra bits, or in the slot
// after a match start the distance's with its extra bits), a block scan of the
// bit counts, every code ored into LDS at its offset, flush of the whole
// words.  Slot npos of the segment's last slice is the end-of-block code.
// The single image's form runs at five waves per SIMD (96 registers): five
// workgroups per CU hold 1280 slices at once, so an image of up to that many
// slices has no second round of workgroups on a nearly idle GPU (the
// radiograph's 1153 slices: 30.8 -> 29.3 us); the job-table form keeps four,
// where five would spill 14 registers inside its loop.
template <bool JOBS>
__global__ __launch_bounds__(kEmitThreads) __attribute__((amdgpu_waves_per_eu(JOBS ? 4 : 5))) void k_huff_emit(HuffLaunch HL)
{
    constexpr int NT = kEmitThreads, NW = NT / 64;
    __shared__ uint32_t T[kHuffTab + 1];                      // + kTokNone's zero entry
    __shared__ __attribute__((aligned(8))) uint32_t obuf[kEmitWords];
    __shared__ uint32_t red[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < kEmitWords; i += NT) obuf[i] = 0;
    if (tid == 0) T[kTokNone] = 0;
    const uint32_t total = HL.counts ? HL.counts[2] : HL.total_slices;
    const int njobs = HL.counts ? (int)HL.counts[0] : HL.njobs;
    for (uint32_t gsl = blockIdx.x; gsl < total; gsl += gridDim.x) {
        const int j = JOBS ? huff_job_of(HL, HL.slice_start, gsl, njobs) : 0;
        const HuffJob H = JOBS ? HL.jobs[j] : HL.single;
        const int sl = (int)(gsl - (JOBS ? HL.slice_start[j] : 0u));
        const int seg = lay_seg(sl, H.spb, H.half_from);
        const bool first = sl == lay_first(seg, H.spb, H.half_from);
        const bool lastsl = sl == lay_first(seg + 1, H.spb, H.half_from) - 1 || sl == H.nslice - 1;
        const uint32_t npos = H.tok_count[sl];
        const uint32_t ntok = npos + (lastsl ? 1u : 0u);     // + the end-of-block code
        const uint32_t* tk = H.tokens + lay_start(sl, H.spb, H.half_from, H.slice_bytes) / 2;
        const uint32_t slots = (uint32_t)lay_bytes(sl, H.spb, H.half_from, H.slice_bytes);
        u32x4 nx[2];
        // (loads bounded by the slice's slots, not its length: the first
        // round's loads do not wait for the count; slots past the length are
        // replaced below)
        auto fetch = [&](uint32_t r) {
            const bool any = r * NT * kEmitPer + (uint32_t)tid * kEmitPer < slots;
#pragma unroll
            for (int q = 0; q < 2; q++)
                nx[q] = any ? *(const u32x4*)(tk + r * NT * kEmitPer / 2 + 4 * (tid + NT * q))
                            : u32x4{kTokNone * 0x10001u, kTokNone * 0x10001u, kTokNone * 0x10001u, kTokNone * 0x10001u};
        };
        fetch(0);
        __syncthreads();                                       // the previous slice is done with T
        for (int i = tid; i < kHuffTab; i += NT) T[i] = H.codes[(int64_t)seg * kHuffTab + i];
        // bits already in obuf: the segment's header before its first slice
        uint32_t pre = 0;
        if (first) {
            pre = H.hdr_bits[seg];
            for (uint32_t i = tid; i < (pre + 31) / 32; i += NT) obuf[i] = H.hdr[(int64_t)seg * kHuffHdrWords + i];
        }
        __syncthreads();
        uint32_t* out = H.slice_out + lay_buf(sl, H.spb, H.half_from, H.slice_cap_dw, H.slice_cap_dw2);
        int64_t out_dw = 0;
        uint32_t carry_bits = pre;
        const uint32_t rounds = max(1u, (ntok + NT * kEmitPer - 1) / (NT * kEmitPer));
        for (uint32_t r = 0; r < rounds; r++) {
            const uint32_t t0 = r * NT * kEmitPer + tid * kEmitPer;
            uint32_t v[kEmitPer];
#pragma unroll
            for (int i = 0; i < kEmitPer; i++) v[i] = (nx[i >> 3][(i >> 1) & 3] >> (16 * (i & 1))) & 0xFFFFu;
            if (r + 1 < rounds) fetch(r + 1);
            if (t0 + kEmitPer > npos) {                       // the slice's end, and past it
#pragma unroll
                for (int i = 0; i < kEmitPer; i++)
                    if (t0 + i >= npos) v[i] = t0 + i == npos && lastsl ? 256u : kTokNone;
            }
            uint32_t a[kEmitPer], n[kEmitPer];
            uint32_t nbits = 0;
#pragma unroll
            for (int i = 0; i < kEmitPer; i++) {
                const bool isd = i > 0 && (v[i - 1] & 0x8000u);   // a distance - 1
                const uint32_t e = 30u - __builtin_clz(v[i] | 2u);
                const uint32_t idx = isd ? 286u + (e << 1) + (v[i] >> e)
                                         : (v[i] & 0x8000u) ? (uint32_t)kHuffSyms + (v[i] & 0xFFu) : v[i];
                const uint32_t c = T[idx], cn = c >> 24;
                a[i] = (c & 0xFFFFFFu) | (isd ? (v[i] & ((1u << e) - 1u)) << cn : 0u);
                n[i] = cn + (isd ? e : 0u);
                nbits += n[i];
            }
            const uint32_t incl = wave_incl_scan(nbits);
            if (lane == 63) red[wave] = incl;
            __syncthreads();
            uint32_t wbase = 0, tot = 0;
            for (int w2 = 0; w2 < NW; w2++) { if (w2 < wave) wbase += red[w2]; tot += red[w2]; }
            uint32_t off = carry_bits + wbase + incl - nbits;
#pragma unroll
            for (int i = 0; i < kEmitPer; i++) {
                if (n[i]) put_bits64(obuf, a[i], n[i], off);
                off += n[i];
            }
            __syncthreads();
            const uint32_t tb = carry_bits + tot;
            const uint32_t full = tb >> 5;
            for (uint32_t i = tid; i < full; i += NT) { out[out_dw + i] = obuf[i]; obuf[i] = 0; }
            __syncthreads();
            if (tid == 0) { obuf[0] = (tb & 31) ? obuf[full] : 0u; if (full) obuf[full] = 0; }
            __syncthreads();
            carry_bits = tb & 31;
            out_dw += full;
        }
        if (tid == 0) {
            if (carry_bits) out[out_dw] = obuf[0];
            obuf[0] = 0;
            H.slice_bits[sl] = (uint32_t)(out_dw * 32 + carry_bits);
        }
    }
}

// ---------------------------------------------------------------------------
// k_scan: per job, exclusive scan of segment lengths, payload size and the
// zlib trailer's Adler-32 (combine of the per-segment values).
// ---------------------------------------------------------------------------


__device__ __forceinline__ int64_t png_payload(const AssembleJob& A, uint64_t segbytes)
{
    return A.mode == 1 ? (int64_t)segbytes + 6 : (int64_t)segbytes;
}
// Bytes written for a payload of z bytes: mode 1 adds the signature, IHDR,
// the 12 bytes of every IDAT chunk and IEND.
__device__ __forceinline__ uint64_t out_total(const AssembleJob& A, int64_t z)
{
    if (A.mode != 1) return (uint64_t)z;
    const int64_t nch = (z + A.chunk_bytes - 1) / A.chunk_bytes;
    return (uint64_t)(8 + 25 + nch * 12 + z + 12);
}

// Byte length of a segment with tb token bits (block header included): EOB
// (7 bits, unless the bits hold it: dynamic blocks), and unless it is the
// stream's last, the 3-bit stored-block header, the pad to a byte and
// 00 00 FF FF.
__device__ __forceinline__ uint32_t seg_len_bytes(uint32_t tb, bool last, bool eob_in_bits)
{
    return (tb + (eob_in_bits ? 0u : 7u) + (last ? 0u : 3u) + 7) / 8 + (last ? 0u : 4u);
}

// Slice i's terms of the whole input's Adler-32 (sum form, mod 65521): for
// a slice with sums A_i (A_i - 1 = am) and B_i and R bytes after it,
// A = 1 + sum(am) and B = sum(B_i + am * R).
__device__ __forceinline__ void adler_terms(const AssembleJob& A, int i, uint32_t& sa, uint32_t& sb)
{
    constexpr uint32_t M = 65521;
    const int64_t end = min64(lay_start(i, A.spb, A.half_from, A.slice_bytes) + lay_bytes(i, A.spb, A.half_from, A.slice_bytes),
                              A.total_in);
    const uint32_t R = (uint32_t)((uint64_t)(A.total_in - end) % M);
    const uint32_t a = A.slice_adler[i];
    const uint32_t am = (a & 0xFFFFu) ? (a & 0xFFFFu) - 1u : M - 1u;
    sa += am;
    sa = sa >= M ? sa - M : sa;
    sb = (sb + (a >> 16) + (am * R) % M) % M;
}

__global__ __launch_bounds__(256) void k_scan(AssembleLaunch AL)
{
    __shared__ uint64_t wsum[4];
    __shared__ uint32_t ared[8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (AL.counts && blockIdx.x >= AL.counts[0]) return;
    const AssembleJob A = AL.jobs ? AL.jobs[blockIdx.x] : AL.single;
    // per segment: bit offsets of its slices, token bits, then the byte
    // length with its trailer (EOB, and unless last the 3-bit stored header,
    // pad, 00 00 FF FF); exclusive scan of the lengths
    uint64_t base = 0;
    for (int c0 = 0; c0 < A.nseg; c0 += 256) {
        const int i = c0 + tid;
        uint32_t v = 0;
        if (i < A.nseg) {
            const int k0 = lay_first(i, A.spb, A.half_from), k1 = min(lay_first(i + 1, A.spb, A.half_from), A.nslice);
            uint32_t tb = 0;
            for (int kb = k0; kb < k1; kb += 8) {          // eight loads in flight
                uint32_t bits[8];
#pragma unroll
                for (int u = 0; u < 8; u++) bits[u] = kb + u < k1 ? A.slice_bits[kb + u] : 0u;
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (kb + u < k1) { A.slice_boff[kb + u] = tb; tb += bits[u]; }
            }
            A.seg_tbits[i] = tb;
            v = seg_len_bytes(tb, (i == A.nseg - 1) && A.final_block, A.eob_in_bits);
        }
        const uint32_t incl = wave_incl_scan(v);
        __syncthreads();
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint64_t wb = 0, tot = 0;
        for (int w = 0; w < 4; w++) { if (w < wave) wb += wsum[w]; tot += wsum[w]; }
        if (i < A.nseg) A.seg_off[i] = base + wb + incl - v;
        base += tot;
    }
    // Adler-32 of the whole input from the slices' values: for pieces with
    // sums (A_i - 1) and R_i bytes after them, A = 1 + sum(A_i - 1) and
    // B = sum(B_i + (A_i - 1) R_i) (mod 65521) -- a plain sum, so every
    // slice is one independent term
    constexpr uint32_t M = 65521;
    uint32_t sa = 0, sb = 0;
    for (int i = tid; i < A.nslice; i += 256) adler_terms(A, i, sa, sb);
    sa = wave_total(sa) % M;                       // 64 terms < M: no overflow
    sb = wave_total(sb) % M;
    if (lane == 0) { ared[wave] = sa; ared[4 + wave] = sb; }
    __syncthreads();
    const uint32_t adler_a = (1u + ared[0] + ared[1] + ared[2] + ared[3]) % M;
    const uint32_t adler_b = (ared[4] + ared[5] + ared[6] + ared[7]) % M;
    if (tid == 0) {
        A.seg_off[A.nseg] = base;
        const uint32_t adler = A.total_in ? (adler_b << 16) | adler_a : 1u;
        if (A.adler_out) *A.adler_out = adler;
        const int64_t z = png_payload(A, base);
        if (A.running) {
            const uint64_t b0 = *A.running;
            *A.base_slot = b0;
            *A.running = b0 + (uint64_t)z;
        }
        const uint64_t len = out_total(A, z);
        *A.out_len = len;
        if (A.host_res) { A.host_res[0] = len; A.host_res[1] = adler; }
    }
}

// ---------------------------------------------------------------------------
// k_assemble: one workgroup per output chunk.
// mode 1 (PNG): chunk c holds zlib payload bytes [c*C, (c+1)*C) of
//   [78 FLG][segments...][adler BE] as [len BE]["IDAT"][data][crc BE];
//   chunk 0 also writes the signature and IHDR, the last chunk IEND.
// mode 0 (raw): chunk c copies payload bytes [c*C, (c+1)*C) to out.
// The chunk's data is cut into 1024 pieces of 64 bytes counted back from its
// end (the first pieces of a short chunk are partly or wholly empty).
// Thread t moves piece t from the segment buffers to the output in 16-byte
// groups (unaligned loads, stores coalesced through the wave's LDS stage;
// groups across a slice or segment boundary are gathered by the wave) and
// folds it into R(piece), the CRC register run from 0 with no final
// inversion (a whole piece as two 32-byte chains, R(A|B) below with
// |B| = 32).  R is linear and
// leading zeros do not change it, so every piece acts as exactly 64 bytes
// and the pieces combine in a tree R(A|B) = R(A) * x^(8|B|) + R(B) whose
// shift at level k is one fixed multiplication by x^(8*64*2^k), done as four
// lookups in the workgroup's LDS copy of the table.  The chunk CRC is
// crc32("IDAT"|D) = ~(g * x^(8|D|) + R(D)) with g the register after
// "IDAT": the piece holding the chunk's first byte starts its chain from g
// instead of 0 (its real bytes run from there), which puts that term in
// place for any chunk length.
// ---------------------------------------------------------------------------
constexpr int kAsmBytes = 64;                  // payload bytes per thread
constexpr int kAsmShiftLevels = 10;            // tree levels: 64 B to 32 KiB
constexpr uint32_t kIdatRegister = 0xCA50F9E1u;    // CRC-32 register after "IDAT": ~crc32("IDAT")
static_assert((64 << (kAsmShiftLevels - 1)) * 2 == 65536, "the tree spans one 64 KiB chunk");
static_assert(kAssembleThreads * kAsmBytes == 65536, "one 64 KiB chunk per workgroup");

__device__ __forceinline__ void put_be32(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)(v >> 24); p[1] = (uint8_t)(v >> 16); p[2] = (uint8_t)(v >> 8); p[3] = (uint8_t)v;
}

typedef R2P_GLOBAL u32x4_ua gu32x4_ua_w;
__device__ __forceinline__ void st128u(uint8_t* p, u32x4 v) { *(gu32x4_ua_w*)p = v; }

// 32 bits of slice k's bit string from bit q (bits past its end are junk;
// the slice buffers have slack words for the second read).
__device__ __forceinline__ uint32_t slice_read32(const AssembleJob& A, int k, uint32_t q)
{
    const uint32_t* w = A.slice_out + lay_buf(k, A.spb, A.half_from, A.slice_cap_dw, A.slice_cap_dw2) + (q >> 5);
    return __builtin_amdgcn_alignbit(w[1], w[0], q & 31);
}

__device__ __forceinline__ uint32_t low_bits(uint32_t v, uint32_t n) { return n >= 32 ? v : v & ((1u << n) - 1); }

// 32 bits of segment s's token string from bit p (zero past its end): from
// the slice holding p and, across its end, the next one (a slice shorter
// than 32 bits can only be a segment's last).  k is a slice cursor at or
// before that slice, advanced in place.
__device__ uint32_t seg_bits32(const AssembleJob& A, int s, uint32_t p, int& k)
{
    const uint32_t tb = A.seg_tbits[s];
    if (p >= tb) return 0u;
    const int k1 = min(lay_first(s + 1, A.spb, A.half_from), A.nslice);
    k = max(k, lay_first(s, A.spb, A.half_from));
    while (k + 1 < k1 && A.slice_boff[k + 1] <= p) k++;
    const uint32_t off = p - A.slice_boff[k];
    const uint32_t avail = A.slice_bits[k] - off;
    uint32_t v = low_bits(slice_read32(A, k, off), avail);
    if (avail < 32 && k + 1 < k1) v |= low_bits(slice_read32(A, k + 1, 0), A.slice_bits[k + 1]) << avail;
    return low_bits(v, tb - p);
}

// NQ * 16 bytes of segment s from its bit p (a multiple of 8), into v: the
// token bits of slice k (the slice holding p, or the segment's last) and,
// past its end, of slices k + 1 and k + 2; zero past the segment's token bits
// (EOB, the stored block's header and the pad are zero bits), then 0xFF for
// the sync flush's last two bytes.  False (v untouched) when a fourth slice
// would be needed.  Every load is issued up front: slice k + 1 is read from
// ceil(n1 / 32) dwords before its start (slice k's buffer), so that dword i
// of the result takes its high part from loaded word i at one lane-uniform
// shift, and slice k + 2 likewise from ceil((n1 + n2) / 32) dwords before its
// start (the short slices of a flat tail are often under 64 bytes; THREE
// off: a third slice is not read, and false when it would be needed).
template <int NQ, bool THREE = false>
__device__ __forceinline__ bool seg_bytes(const AssembleJob& A, int s, uint32_t p, int k, u32x4* v)
{
    constexpr int ND = 4 * NQ;
    const int k1 = min(lay_first(s + 1, A.spb, A.half_from), A.nslice);
    const uint32_t bk = A.slice_boff[k], ek = bk + A.slice_bits[k];
    const uint32_t n1 = p < ek ? ek - p : 32u * ND;          // bits from slice k
    const bool has2 = k + 1 < k1;
    const uint32_t n2 = has2 ? A.slice_bits[k + 1] : 0u;
    const bool need3 = n1 < 32u * ND && has2 && n1 + n2 < 32u * ND && k + 2 < k1;
    if (!THREE && need3) return false;
    const uint32_t n3 = THREE && need3 ? A.slice_bits[k + 2] : 0u;
    if (THREE && need3 && n1 + n2 + n3 < 32u * ND && k + 3 < k1) return false;
    const uint32_t off = p - bk;
    const uint32_t* w = A.slice_out + lay_buf(k, A.spb, A.half_from, A.slice_cap_dw, A.slice_cap_dw2) + (off >> 5);
    uint32_t t[ND + 1];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const u32x4 x = *(R2P_GLOBAL const u32x4_ua*)(w + 4 * q);
        t[4 * q] = x[0]; t[4 * q + 1] = x[1]; t[4 * q + 2] = x[2]; t[4 * q + 3] = x[3];
    }
    t[ND] = w[ND];
    uint32_t d[ND];
    const uint32_t sh = off & 31;
#pragma unroll
    for (int i = 0; i < ND; i++) d[i] = __builtin_amdgcn_alignbit(t[i + 1], t[i], sh);
    if (n1 < 32u * ND) {
        uint32_t u[ND + 1];
#pragma unroll
        for (int i = 0; i <= ND; i++) u[i] = 0;
        if (has2) {
            const uint32_t c = (n1 + 31) >> 5;
            const uint32_t* w2 = A.slice_out + lay_buf(k + 1, A.spb, A.half_from, A.slice_cap_dw, A.slice_cap_dw2) - c;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const u32x4 x = *(R2P_GLOBAL const u32x4_ua*)(w2 + 4 * q);
                u[4 * q] = x[0]; u[4 * q + 1] = x[1]; u[4 * q + 2] = x[2]; u[4 * q + 3] = x[3];
            }
            u[ND] = w2[ND];
        }
        const uint32_t sh2 = (32u - (n1 & 31u)) & 31u;
#pragma unroll
        for (int i = 0; i < ND; i++) {
            const uint32_t vu = __builtin_amdgcn_alignbit(u[i + 1], u[i], sh2);
            const int r = (int)n1 - 32 * i;                    // bits of dword i from slice k
            const uint32_t m = r >= 32 ? ~0u : r <= 0 ? 0u : (1u << r) - 1u;
            d[i] = (d[i] & m) | (vu & ~m);
        }
    }
    if (THREE && need3) {
        const uint32_t n12 = n1 + n2;
        const uint32_t c = (n12 + 31) >> 5;
        const uint32_t* w3 = A.slice_out + lay_buf(k + 2, A.spb, A.half_from, A.slice_cap_dw, A.slice_cap_dw2) - c;
        uint32_t u[ND + 1];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const u32x4 x = *(R2P_GLOBAL const u32x4_ua*)(w3 + 4 * q);
            u[4 * q] = x[0]; u[4 * q + 1] = x[1]; u[4 * q + 2] = x[2]; u[4 * q + 3] = x[3];
        }
        u[ND] = w3[ND];
        const uint32_t sh3 = (32u - (n12 & 31u)) & 31u;
#pragma unroll
        for (int i = 0; i < ND; i++) {
            const uint32_t vu = __builtin_amdgcn_alignbit(u[i + 1], u[i], sh3);
            const int r = (int)n12 - 32 * i;                   // bits of dword i from slices k, k + 1
            const uint32_t m = r >= 32 ? ~0u : r <= 0 ? 0u : (1u << r) - 1u;
            d[i] = (d[i] & m) | (vu & ~m);
        }
    }
    const uint32_t tb = A.seg_tbits[s];
    if (p + 32u * ND > tb) {
#pragma unroll
        for (int i = 0; i < ND; i++) {
            const int r = (int)tb - (int)p - 32 * i;          // token bits left in dword i
            d[i] &= r >= 32 ? ~0u : r <= 0 ? 0u : (1u << r) - 1u;
        }
    }
    const uint32_t seglen = (uint32_t)(A.seg_off[s + 1] - A.seg_off[s]);
    const uint32_t bo = p >> 3;
    if (!((s == A.nseg - 1) && A.final_block) && bo + 4u * ND > seglen - 2) {
#pragma unroll
        for (int i = 0; i < 4 * ND; i++)
            if (bo + i >= seglen - 2) d[i >> 2] |= 0xFFu << (8 * (i & 3));
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = (u32x4){d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]};
    return true;
}

// Payload bytes [q, q + n) (n <= 16) of mode 1: [78 FLG][segments][adler BE]
// (mode 0: the segments), assembled 4 bytes at a time where they come from
// one segment; s and k are segment and slice cursors at or before q.
__device__ u32x4 payload16(const AssembleJob& A, int64_t hdr, int64_t segbytes, int64_t q, int n, int& s, int& k)
{
    unsigned __int128 acc = 0;
    int t = 0;
    while (t < n) {
        const int64_t x = q + t;
        if (x < hdr) {
            acc |= (unsigned __int128)(x == 0 ? 0x78u : (uint32_t)A.zflg) << (8 * t);
            t++;
            continue;
        }
        const int64_t a = x - hdr;
        if (a >= segbytes) {
            acc |= (unsigned __int128)((*A.adler_out >> (24 - 8 * (int)(a - segbytes))) & 0xFFu) << (8 * t);
            t++;
            continue;
        }
        while ((int64_t)A.seg_off[s + 1] <= a) s++;
        const uint32_t bo = (uint32_t)(a - (int64_t)A.seg_off[s]);
        const uint32_t seglen = (uint32_t)(A.seg_off[s + 1] - A.seg_off[s]);
        const bool last = (s == A.nseg - 1) && A.final_block;
        const int m = (int)min((uint32_t)min(4, n - t), seglen - bo);     // bytes from this segment
        uint32_t v = seg_bits32(A, s, 8 * bo, k);
        if (!last && bo + m > seglen - 2) {                                  // ... 00 00 FF FF
            for (int i = 0; i < m; i++)
                if (bo + i >= seglen - 2) v |= 0xFFu << (8 * i);
        }
        acc |= (unsigned __int128)low_bits(v, 8 * m) << (8 * t);
        t += m;
    }
    u32x4 r;
    r[0] = (uint32_t)acc; r[1] = (uint32_t)(acc >> 32); r[2] = (uint32_t)(acc >> 64); r[3] = (uint32_t)(acc >> 96);
    return r;
}

// Last segment s in [0, nseg) with seg_off[s] <= t (seg_off[0] = 0 <= t),
// found by the whole wave: each round probes 64 evenly spaced candidates
// and keeps the span after the last one at or below t, so a few dependent
// loads replace a per-lane binary search.  t must be the same on all lanes.
__device__ int wave_find_seg(const uint64_t* seg_off, int nseg, int64_t t)
{
    const int lane = threadIdx.x & 63;
    int lo = 0, n = nseg;
    while (n > 1) {
        const int step = (n + 63) >> 6;
        const int idx = lo + lane * step;
        const bool p = idx < lo + n && (int64_t)seg_off[idx] <= t;
        const int cnt = __popcll(__ballot(p));          // a prefix of the lanes (lane 0 always)
        const int nlo = lo + (cnt - 1) * step;
        n = min(step, lo + n - nlo);
        lo = nlo;
    }
    return lo;
}


// The 64 bytes in w moved up by o bytes (0 < o < 64), zeros below: a
// barrel shift by whole dwords (per-lane selects), then a byte align.
__device__ __forceinline__ void shift_up_bytes(uint32_t (&w)[16], int o)
{
    const int oq = o >> 2, orr = o & 3;
#pragma unroll
    for (int bit = 8; bit >= 1; bit >>= 1) {
        const bool on = (oq & bit) != 0;
#pragma unroll
        for (int d = 15; d >= 0; d--) w[d] = on ? (d >= bit ? w[d - bit] : 0u) : w[d];
    }
    if (orr) {
#pragma unroll
        for (int d = 15; d >= 0; d--) w[d] = __builtin_amdgcn_alignbyte(w[d], d > 0 ? w[d - 1] : 0u, 4 - orr);
    }
}

// Stage slot of 16-byte group jg of lane l's piece (slots are 16 bytes, four
// per lane): the group index is XORed with bits 1..2 of the lane, so the
// lanes' writes of one group (64 bytes apart) and the coalesced reads of
// consecutive groups both fall on distinct LDS banks within each eight lanes.
__device__ __forceinline__ int stage_slot(int l, int jg) { return l * 4 + (jg ^ ((l >> 1) & 3)); }

// SELF: one job of at most kAsmScanSlices slices and kAsmScanSegs segments,
// whose scan every workgroup computes for itself into LDS (k_scan's work,
// without its launch): slice bit counts and offsets, segment token bits and
// byte offsets, and the Adler-32; the workgroup of the first chunk stores
// the Adler-32 and the output length.
template <bool SELF>
__global__ __launch_bounds__(kAssembleThreads) void k_assemble(AssembleLaunch AL)
{
    __shared__ uint32_t tab[4][256];
    __shared__ uint32_t tab32[4][256];                     // the 32-byte shift
    __shared__ uint32_t shf[kAsmShiftLevels][4][256];      // the tree's shifts, 64 B .. 32 KiB
    __shared__ uint32_t red[kAssembleThreads / 64];
    __shared__ u32x4 stage[kAssembleThreads / 64][256];    // 4 KiB per wave, by stage_slot
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the CRC tables (one entry of each per thread): SELF stores them in its
    // scan, before the scan's first barrier, so that their loads and the
    // scan's are in flight together; otherwise they are stored here
    static_assert(kAssembleThreads == 1024, "one entry of each table per thread");
    const uint32_t t_crc = (&c_crc[0][0])[tid], t_s32 = (&c_crcshift32[0][0])[tid];
    uint32_t t_shf[kAsmShiftLevels];
#pragma unroll
    for (int k = 0; k < kAsmShiftLevels; k++) t_shf[k] = (&c_crcshift[k][0][0])[tid];
    auto store_tables = [&]() {
        (&tab[0][0])[tid] = t_crc;
        (&tab32[0][0])[tid] = t_s32;
#pragma unroll
        for (int k = 0; k < kAsmShiftLevels; k++) (&shf[k][0][0])[tid] = t_shf[k];
    };
    if constexpr (!SELF) {
        store_tables();
        __syncthreads();
    }

    AssembleJob S = AL.single;                     // SELF: the job, its scan arrays in LDS
    if constexpr (SELF) {
        __shared__ uint32_t s_bits[kAsmScanSlices], s_boff[kAsmScanSlices], s_tbits[kAsmScanSegs];
        __shared__ uint64_t s_off[kAsmScanSegs + 1];
        __shared__ uint64_t s_wsum[kAssembleThreads / 64];
        __shared__ uint32_t s_ared[2 * (kAssembleThreads / 64)], s_adler;
        constexpr uint32_t M = 65521;
        const int i = tid;                         // segment i (nseg <= kAsmScanSegs == threads)
        uint32_t v = 0, sa = 0, sb = 0;
        if (i < S.nseg) {
            const int k0 = lay_first(i, S.spb, S.half_from), k1 = min(lay_first(i + 1, S.spb, S.half_from), S.nslice);
            uint32_t tb = 0;
            for (int kb = k0; kb < k1; kb += 8) {          // eight loads in flight
                uint32_t bits[8];
#pragma unroll
                for (int u = 0; u < 8; u++) bits[u] = kb + u < k1 ? S.slice_bits[kb + u] : 0u;
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (kb + u < k1) { s_bits[kb + u] = bits[u]; s_boff[kb + u] = tb; tb += bits[u]; }
            }
            s_tbits[i] = tb;
            v = seg_len_bytes(tb, (i == S.nseg - 1) && S.final_block, S.eob_in_bits);
        }
        for (int k = tid; k < S.nslice; k += kAssembleThreads) adler_terms(S, k, sa, sb);
        store_tables();
        const uint32_t incl = wave_incl_scan(v);
        sa = wave_total(sa) % M;
        sb = wave_total(sb) % M;
        if (lane == 63) s_wsum[wave] = incl;
        if (lane == 0) { s_ared[wave] = sa; s_ared[kAssembleThreads / 64 + wave] = sb; }
        __syncthreads();
        uint64_t wb = 0, tot = 0;
        for (int w = 0; w < kAssembleThreads / 64; w++) { if (w < wave) wb += s_wsum[w]; tot += s_wsum[w]; }
        if (i < S.nseg) s_off[i] = wb + incl - v;
        if (tid == 0) {
            s_off[S.nseg] = tot;
            uint32_t aa = 1u, ab = 0u;
            for (int w = 0; w < kAssembleThreads / 64; w++) { aa += s_ared[w]; ab += s_ared[kAssembleThreads / 64 + w]; }
            s_adler = S.total_in ? ((ab % M) << 16) | (aa % M) : 1u;
            if (blockIdx.x == 0) {
                if (S.adler_out) *S.adler_out = s_adler;
                const uint64_t len = out_total(S, png_payload(S, tot));
                *S.out_len = len;
                if (S.host_res) { S.host_res[0] = len; S.host_res[1] = s_adler; }
            }
        }
        __syncthreads();
        S.slice_bits = s_bits; S.slice_boff = s_boff; S.seg_tbits = s_tbits; S.seg_off = s_off;
        S.adler_out = &s_adler;
    }

    const uint32_t total = AL.counts ? AL.counts[1] : AL.total_chunks;
    const int njobs = AL.counts ? (int)AL.counts[0] : AL.njobs;
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        int j = 0;
        if (AL.jobs) {
            int lo = 0, hi = njobs - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (AL.chunk_start[mid] <= g) lo = mid; else hi = mid - 1;
            }
            j = lo;
        }
        const AssembleJob A = SELF ? S : AL.jobs ? AL.jobs[j] : AL.single;
        const int c = (int)(g - (AL.jobs ? AL.chunk_start[j] : 0u));
        const int64_t segbytes = (int64_t)A.seg_off[A.nseg];
        const int64_t z = png_payload(A, (uint64_t)segbytes);
        const int64_t C = A.chunk_bytes;
        const int64_t nch = (z + C - 1) / C;
        if (c >= nch) continue;                    // uniform per workgroup
        const int64_t z0 = c * C, z1 = min64(z, z0 + C);
        const int n = (int)(z1 - z0);
        const int64_t hdr = A.mode == 1 ? 2 : 0;
        uint8_t* o = A.mode == 1 ? A.out + 33 + c * (C + 12) : nullptr;
        // output address of payload byte zz is dst + zz
        uint8_t* dst = A.mode == 1 ? o + 8 - z0 : A.out + (A.base_slot ? *A.base_slot : 0ull);

        // piece tid: [e - 64, e) with e counted back from the chunk end
        const int64_t e = z1 - (int64_t)(kAssembleThreads - 1 - tid) * kAsmBytes;
        const int64_t b = max64(z0, e - kAsmBytes);
        // the segment of the wave's first byte, then each lane's by a short walk
        const int64_t wfirst = max64(z0, z1 - (int64_t)(kAssembleThreads - 64 * wave) * kAsmBytes);
        const int sw = wave_find_seg(A.seg_off, A.nseg, min64(max64(wfirst - hdr, 0), max64(segbytes - 1, 0)));
        // the chain of the piece holding the chunk's first byte starts at
        // the register after "IDAT"
        const uint32_t r_in = (A.mode == 1 && b == z0 && b < e) ? kIdatRegister : 0u;
        uint32_t r = r_in;
        // a full 64-byte piece: two independent 32-byte chains, joined by the
        // 32-byte shift
        auto crc_piece = [&](const u32x4* v) {
            uint32_t r0 = r_in, r1 = 0;
#pragma unroll
            for (int jg = 0; jg < 2; jg++)
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    r0 = crc_step4(r0, v[jg][u], tab);
                    r1 = crc_step4(r1, v[jg + 2][u], tab);
                }
            return crc_shift(r0, tab32) ^ r1;
        };
        bool whole = false;
        uint32_t slow = 0;                         // this piece's groups that the wave gathers
        int sc = 0, kc = 0;                        // segment and slice cursors at the first of them
        if (b < e) {
            int s = sw;
            const int64_t a0 = b - hdr;
            if (a0 > 0 && a0 < segbytes)
                while ((int64_t)A.seg_off[s + 1] <= a0) s++;
            int k = lay_first(s, A.spb, A.half_from);     // slice cursor
            // the usual case: all 64 bytes from one slice's token bits, read
            // with one cursor lookup and four loads in flight
            whole = e - b == kAsmBytes && a0 >= 0 && a0 + kAsmBytes <= segbytes &&
                         (int64_t)A.seg_off[s + 1] >= a0 + kAsmBytes;
            u32x4 v[4];
            // a piece across exactly one segment boundary: segment s's last
            // o bytes, then segment s + 1's first 64 - o moved up into place
            const bool cross = SELF && e - b == kAsmBytes && a0 >= 0 && a0 + kAsmBytes <= segbytes &&
                               (int64_t)A.seg_off[s + 1] < a0 + kAsmBytes &&
                               (int64_t)A.seg_off[s + 2] >= a0 + kAsmBytes;
            if (whole || cross) {
                const uint32_t pw = 8 * (uint32_t)(a0 - (int64_t)A.seg_off[s]);
                const int k1 = min(lay_first(s + 1, A.spb, A.half_from), A.nslice);
                while (k + 1 < k1 && A.slice_boff[k + 1] <= pw) k++;
                whole = seg_bytes<4, SELF>(A, s, pw, k, v);
            }
            if (cross && whole) {
                u32x4 v2[4];
                whole = seg_bytes<4, SELF>(A, s + 1, 0u, lay_first(s + 1, A.spb, A.half_from), v2);
                if (whole) {
                    const int o = (int)((int64_t)A.seg_off[s + 1] - a0);
                    uint32_t w1[16], w2[16];
#pragma unroll
                    for (int q = 0; q < 4; q++)
#pragma unroll
                        for (int i = 0; i < 4; i++) { w1[4 * q + i] = v[q][i]; w2[4 * q + i] = v2[q][i]; }
                    shift_up_bytes(w2, o);
#pragma unroll
                    for (int d = 0; d < 16; d++) {
                        const int r = o - 4 * d;                 // bytes of dword d from segment s
                        const uint32_t m = r >= 4 ? ~0u : r <= 0 ? 0u : (1u << (8 * r)) - 1u;
                        w1[d] = (w1[d] & m) | (w2[d] & ~m);
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) v[q] = (u32x4){w1[4 * q], w1[4 * q + 1], w1[4 * q + 2], w1[4 * q + 3]};
                }
            }
            if (whole) {
#pragma unroll
                for (int jg = 0; jg < 4; jg++) stage[wave][stage_slot(lane, jg)] = v[jg];
                r = crc_piece(v);
            }
            // otherwise (a piece across a segment boundary, with framing
            // bytes, short, or over three slices) its 16-byte groups are
            // gathered by the wave below, all loads of a round in flight
            // together
            if (!whole) {
                sc = s; kc = k;
                slow = (1u << (int)((e - b + 15) / 16)) - 1u;
            }
        }
        // The gathered groups, up to 16 at a time for the whole wave: lanes
        // 4j .. 4j + 3 take the j-th pending group (in lane, then group
        // order), lane 4j + i assembling dword i of it (the lanes' loads in
        // flight together, where one lane would wait for each group in turn)
        // and storing it into the group's stage slot.
        for (uint32_t pend = slow;;) {
            uint64_t ml = __ballot(pend != 0);
            if (ml == 0) break;
            const int js = lane >> 2, di = lane & 3;
            int src = 0, jg = 0, cnt = 0, cl = 0;
            uint32_t cur = 0;
            for (int j = 0; j < 16; j++) {         // wave-uniform picks
                if (cur == 0) {
                    if (ml == 0) break;
                    cl = __builtin_ctzll(ml);
                    ml &= ml - 1;
                    cur = (uint32_t)__builtin_amdgcn_readlane((int)pend, cl);
                }
                const int gj = __builtin_ctz(cur);
                cur &= cur - 1;
                if (js == j) { src = cl; jg = gj; }
                if (lane == cl) pend &= ~(1u << gj);
                cnt++;
            }
            // (every shuffle runs on all lanes: ds_bpermute reads zero from a
            // source lane that is off in EXEC)
            const int64_t q = (int64_t)__shfl((long long)b, src, 64) + 16 * jg;
            const int64_t e_src = (int64_t)__shfl((long long)e, src, 64);
            int s2 = __shfl(sc, src, 64), k2 = __shfl(kc, src, 64);
            const int nn = js < cnt ? (int)min64(4, e_src - q - 4 * di) : 0;
            if (nn > 0)
                ((uint32_t*)&stage[wave][stage_slot(src, jg)])[di] = payload16(A, hdr, segbytes, q + 4 * di, nn, s2, k2)[0];
        }
        // the CRC of a piece that is not whole, over its staged groups in order
        if (b < e && !whole && e - b == kAsmBytes) {
            u32x4 v[4];
#pragma unroll
            for (int jg = 0; jg < 4; jg++) v[jg] = stage[wave][stage_slot(lane, jg)];
            r = crc_piece(v);
        } else if (b < e && !whole) {
#pragma unroll 1
            for (int jg = 0; jg < 4; jg++) {
                const int64_t q = b + 16 * jg;
                if (q >= e) break;
                const u32x4 v = stage[wave][stage_slot(lane, jg)];
                const int nb = (int)min64(16, e - q);
                if (nb == 16) {
                    r = crc_step4(r, v[0], tab);
                    r = crc_step4(r, v[1], tab);
                    r = crc_step4(r, v[2], tab);
                    r = crc_step4(r, v[3], tab);
                } else {
#pragma unroll
                    for (int i = 0; i < 16; i++)
                        if (i < nb) r = tab[0][(r ^ ((v[i >> 2] >> (8 * (i & 3))) & 0xFFu)) & 0xFF] ^ (r >> 8);
                }
            }
        }
        // stores from the wave's LDS stage: a wave whose 64 pieces are all whole
        // writes consecutive 16 bytes from consecutive lanes; otherwise piece-wise
        __builtin_amdgcn_wave_barrier();
        {
            const int64_t wb0 = z1 - (int64_t)(kAssembleThreads - 64 * wave) * kAsmBytes;   // wave's first byte
            if (wb0 >= z0) {
#pragma unroll
                for (int jg = 0; jg < 4; jg++) st128u(dst + wb0 + 1024 * jg + 16 * lane, stage[wave][stage_slot(16 * jg + (lane >> 2), lane & 3)]);
            } else {
                for (int jg = 0; jg < 4; jg++) {
                    const int64_t q = b + 16 * jg;
                    if (q >= e) break;
                    const u32x4 v = stage[wave][stage_slot(lane, jg)];
                    const int nb = (int)min64(16, e - q);
                    if (nb == 16) st128u(dst + q, v);
                    else
                        for (int i = 0; i < nb; i++) dst[q + i] = (uint8_t)(v[i >> 2] >> (8 * (i & 3)));
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        // tree: levels 0..5 inside the wave, 6..9 across the 16 waves
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const uint32_t right = __shfl_down(r, 1 << k, 64);
            if ((lane & ((2 << k) - 1)) == 0) r = crc_shift(r, shf[k]) ^ right;
        }
        __syncthreads();                           // previous chunk done with red
        if (lane == 0) red[wave] = r;
        __syncthreads();
        if (wave == 0) {
            // (the lane index and the IHDR CRC below are made opaque at
            // their use, so that the values derived from them are not kept
            // live, and spilled, across the chunk loop)
            int li = lane;
            asm volatile("" : "+v"(li));
            r = li < kAssembleThreads / 64 ? red[li] : 0u;
            // 16 waves of 4 KiB: the shifts by 4, 8, 16 and 32 KiB
            static_assert(kAssembleThreads == 1024, "four cross-wave levels");
#pragma unroll
            for (int k = 6; k < 10; k++) {
                const uint32_t right = __shfl_down(r, 1 << (k - 6), 64);
                if ((lane & ((2 << (k - 6)) - 1)) == 0) r = crc_shift(r, shf[k]) ^ right;
            }
        }
        if (A.mode == 1 && tid == 0) {
            if (c == 0) {
                const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
                for (int i = 0; i < 8; i++) A.out[i] = sig[i];
                uint8_t ih[17] = {'I', 'H', 'D', 'R'};
                put_be32(ih + 4, A.width); put_be32(ih + 8, A.height);
                ih[12] = A.bit_depth; ih[13] = A.color_type; ih[14] = 0; ih[15] = 0; ih[16] = 0;
                put_be32(A.out + 8, 13);
                for (int i = 0; i < 17; i++) A.out[12 + i] = ih[i];
                uint32_t ihdr_crc = A.ihdr_crc;
                asm volatile("" : "+v"(ihdr_crc));
                put_be32(A.out + 29, ihdr_crc);
            }
            put_be32(o, (uint32_t)n);
            o[4] = 'I'; o[5] = 'D'; o[6] = 'A'; o[7] = 'T';
            put_be32(o + 8 + n, ~r);
            if (c == nch - 1) {
                uint8_t* ie = o + 12 + n;
                put_be32(ie, 0);
                ie[4] = 'I'; ie[5] = 'E'; ie[6] = 'N'; ie[7] = 'D';
                put_be32(ie + 8, 0xAE426082u);
            }
        }
    }
}

// Writes up to 16 literal bytes, then (optionally) a big-endian CRC read
// from device memory right after them.  Used for IDAT chunk framing.
struct PutBytes { uint8_t b[16]; };
__global__ void k_put_bytes(uint8_t* dst, PutBytes pb, int n, const uint32_t* crc_src)
{
    if (threadIdx.x == 0) {
        for (int i = 0; i < n; i++) dst[i] = pb.b[i];
        if (crc_src) put_be32(dst + n, *crc_src);
    }
}

// ---------------------------------------------------------------------------
// Host-side launch wrappers (called from r2p_capi.cpp)
// ---------------------------------------------------------------------------
hipError_t launch_build_lut(const LutBuild& lb, void* out, hipStream_t s)
{
    (void)hipGetLastError();
    k_build_lut<<<(lb.entries + 255) / 256, 256, 0, s>>>(lb, out);
    return hipGetLastError();
}

template <int K>
static hipError_t launch_filter_k(const FilterLaunch& L, int grid, size_t lds, hipStream_t s)
{
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_filter<K>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    k_filter<K><<<grid, kFilterThreads, lds, s>>>(L);
    return hipGetLastError();
}

template <int K>
static hipError_t launch_filter_split_k(const FilterLaunch& L, int grid, size_t lds, hipStream_t s)
{
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_filter_split<K>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    k_filter_split<K><<<grid, kFilterRowThreads, lds, s>>>(L);
    return hipGetLastError();
}

// Instantiation by the widest row: k_filter's lane l of a wave takes items
// l, l + 64, ... of a row (K of them, rows up to K KiB); k_filter_split
// spreads a row over the lanes of 8 / band_rows waves.
hipError_t launch_filter(const FilterLaunch& L, int grid, size_t lds, hipStream_t s)
{
    (void)hipGetLastError();
    const int items = (L.row_cap + 15) / 16;
    if (!filter_split(L.row_cap)) {
        const int k = (items + 63) / 64;
        if (k <= 1) return launch_filter_k<1>(L, grid, lds, s);
        if (k <= 2) return launch_filter_k<2>(L, grid, lds, s);
        if (k <= 3) return launch_filter_k<3>(L, grid, lds, s);
        if (k <= 4) return launch_filter_k<4>(L, grid, lds, s);
        return launch_filter_k<6>(L, grid, lds, s);
    }
    const int k = filter_items_per_lane(items, L.band_rows);
    if (L.band_rows < 1 || L.band_rows > kFilterRowThreads / 64 || (L.band_rows & (L.band_rows - 1)) ||
        k > kFilterMaxK)
        return hipErrorInvalidValue;
    if (k <= 2) return launch_filter_split_k<2>(L, grid, lds, s);
    if (k <= 3) return launch_filter_split_k<3>(L, grid, lds, s);
    if (k <= 4) return launch_filter_split_k<4>(L, grid, lds, s);
    if (k <= 5) return launch_filter_split_k<5>(L, grid, lds, s);
    if (k <= 6) return launch_filter_split_k<6>(L, grid, lds, s);
    return launch_filter_split_k<8>(L, grid, lds, s);
}

hipError_t launch_filter_tiled(const FilterLaunch& L, int grid, size_t lds, hipStream_t s)
{
    (void)hipGetLastError();
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_filter_tiled, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    k_filter_tiled<<<grid, kFilterThreads, lds, s>>>(L);
    return hipGetLastError();
}

// Workgroup sizes: 256 (4 K positions per sub-block, three workgroups per
// CU) and 512 (8 K positions, one or two per CU).  Both need ~165 VGPRs, so
// 1024-thread groups (at most 128) would spill.
hipError_t launch_deflate(const DeflateLaunch& DL, int lane_bytes, int wg_threads, int grid, hipStream_t s)
{
    (void)lane_bytes;
    (void)hipGetLastError();
    const bool stats = !DL.jobs && DL.single.slice_stats;
    if (DL.dense) {             // hashed search at every 2nd position
        if (DL.dyn)                 k_deflate<256, false, true, 2><<<grid, 256, 0, s>>>(DL);
        else if (stats)             k_deflate<256, true, false, 2><<<grid, 256, 0, s>>>(DL);
        else if (wg_threads == 256) k_deflate<256, false, false, 2><<<grid, 256, 0, s>>>(DL);
        else                        k_deflate<512, false, false, 2><<<grid, 512, 0, s>>>(DL);
    } else {                    // at every 4th
        if (DL.dyn)                 k_deflate<256, false, true, 4><<<grid, 256, 0, s>>>(DL);
        else if (stats)             k_deflate<256, true, false, 4><<<grid, 256, 0, s>>>(DL);
        else if (wg_threads == 256) k_deflate<256, false, false, 4><<<grid, 256, 0, s>>>(DL);
        else                        k_deflate<512, false, false, 4><<<grid, 512, 0, s>>>(DL);
    }
    return hipGetLastError();
}

hipError_t launch_scan(const AssembleLaunch& AL, int njobs, hipStream_t s)
{
    (void)hipGetLastError();
    k_scan<<<njobs, 256, 0, s>>>(AL);
    return hipGetLastError();
}

hipError_t launch_huffman(const HuffLaunch& HL, int grid_tables, int grid_emit, hipStream_t s)
{
    (void)hipGetLastError();
    if (HL.jobs) {
        k_huff_tables<true><<<grid_tables, kHuffThreads, 0, s>>>(HL);
        k_huff_emit<true><<<grid_emit, kEmitThreads, 0, s>>>(HL);
    } else {
        k_huff_tables<false><<<grid_tables, kHuffThreads, 0, s>>>(HL);
        k_huff_emit<false><<<grid_emit, kEmitThreads, 0, s>>>(HL);
    }
    return hipGetLastError();
}

hipError_t launch_assemble(const AssembleLaunch& AL, int grid, hipStream_t s)
{
    (void)hipGetLastError();
    k_assemble<false><<<grid, kAssembleThreads, 0, s>>>(AL);
    return hipGetLastError();
}

bool assemble_scans_itself(const AssembleJob& A)
{
    return A.nslice <= kAsmScanSlices && A.nseg <= kAsmScanSegs && !A.running;
}

hipError_t launch_assemble_self_scan(const AssembleLaunch& AL, int grid, hipStream_t s)
{
    (void)hipGetLastError();
    k_assemble<true><<<grid, kAssembleThreads, 0, s>>>(AL);
    return hipGetLastError();
}

hipError_t launch_adler(const uint8_t* p, int64_t n, uint32_t* part, uint32_t* result, hipStream_t s)
{
    (void)hipGetLastError();
    const int nb = (int)((n + kChecksumBlock - 1) / kChecksumBlock);
    if (nb > 0) k_adler_partial<<<nb, 256, 0, s>>>(p, n, part);
    k_adler_final<<<1, kFinalThreads, 0, s>>>(part, nb, result);
    return hipGetLastError();
}

hipError_t launch_crc(const uint8_t* p, int64_t n, uint32_t* part, uint32_t* result, hipStream_t s)
{
    (void)hipGetLastError();
    const int64_t nb = (n + kCrcBlock - 1) / kCrcBlock;
    int lm = 0;                                     // blocks per combine thread: 2^lm
    while (((int64_t)kFinalThreads << lm) < nb) lm++;
    if (kBlockLevel + lm + 9 >= kCrcShiftLevels) return hipErrorInvalidValue;
    if (nb > 0) k_crc_partial<<<nb < kCrcPartialGrid ? (int)nb : kCrcPartialGrid, 256, 0, s>>>(p, n, part);
    k_crc_final<<<1, kFinalThreads, 0, s>>>(part, (int)nb, lm, r2p_crc32_combine(0xFFFFFFFFu, 0u, (uint64_t)n), result);
    return hipGetLastError();
}

hipError_t launch_put_bytes(uint8_t* dst, const uint8_t* bytes16, int n, const uint32_t* crc_src,
                            hipStream_t s)
{
    (void)hipGetLastError();
    PutBytes pb;
    for (int i = 0; i < 16; i++) pb.b[i] = i < n ? bytes16[i] : 0;
    k_put_bytes<<<1, 64, 0, s>>>(dst, pb, n, crc_src);
    return hipGetLastError();
}

hipError_t upload_tables(const uint32_t* lit, const uint32_t* len, const uint32_t* dist_rev,
                         const uint32_t* crc4, const uint32_t* x8pow, const uint32_t* crc_shift,
                         const uint32_t* crc_shift32)
{
    hipError_t e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_lit), lit, sizeof(uint32_t) * 256)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_len), len, sizeof(uint32_t) * 259)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_dist_rev), dist_rev, sizeof(uint32_t) * 32)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_crc), crc4, sizeof(uint32_t) * 1024)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_crcshift), crc_shift, sizeof(uint32_t) * kCrcShiftLevels * 1024)) != hipSuccess)
        return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_crcshift32), crc_shift32, sizeof(uint32_t) * 1024)) != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_x8pow), x8pow, sizeof(uint32_t) * 64);
}

}  // namespace r2p
