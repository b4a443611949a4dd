// r2p_capi.cpp -- the C-ABI (include/r2p.h): device setup, the plain
// per-image encoder, kernel-level entry points for the drop-in facade.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include <zlib.h>

#include "r2p_common.h"

namespace r2p {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char tl_err[512];

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(tl_err, sizeof(tl_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char* what)
{
    set_error("%s: %s", what, hipGetErrorString(e));
    return R2P_ERR_HIP;
}

// ---------------------------------------------------------------------------
// constant tables (RFC 1951 section 3.2.5-3.2.6, PNG Annex D)
// ---------------------------------------------------------------------------
static uint32_t rev_bits(uint32_t v, int n)
{
    uint32_t r = 0;
    for (int i = 0; i < n; i++) { r = (r << 1) | (v & 1); v >>= 1; }
    return r;
}

// Fixed literal/length code of symbol s (RFC 1951 3.2.6), bit-reversed for
// an LSB-first writer; *nb = code length.
static uint32_t fixed_code(int s, int* nb)
{
    if (s < 144) { *nb = 8; return rev_bits(0x30 + s, 8); }
    if (s < 256) { *nb = 9; return rev_bits(0x190 + (s - 144), 9); }
    if (s < 280) { *nb = 7; return rev_bits(s - 256, 7); }
    *nb = 8;
    return rev_bits(0xC0 + (s - 280), 8);
}

static void build_tables(uint32_t lit[256], uint32_t len[259], uint32_t dist_rev[32],
                         uint32_t crc4[4][256], uint32_t x8pow[64], uint32_t shift[kCrcShiftLevels][4][256],
                         uint32_t shift32[4][256])
{
    for (int b = 0; b < 256; b++) {
        int nb;
        const uint32_t c = fixed_code(b, &nb);
        lit[b] = c | ((uint32_t)nb << 16);
    }
    // length codes 257..285: base lengths and extra bits
    static const int base[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                                 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
    static const int extra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2,
                                  3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
    len[0] = len[1] = len[2] = 0;
    for (int L = 3; L <= 258; L++) {
        int c = 28;
        if (L < 258) { c = 0; while (c + 1 < 28 && base[c + 1] <= L) c++; }
        int nb;
        const uint32_t code = fixed_code(257 + c, &nb);
        const uint32_t ev = (uint32_t)(L - base[c]);
        len[L] = (code | (ev << nb)) | ((uint32_t)(nb + extra[c]) << 16);
    }
    for (int d = 0; d < 32; d++) dist_rev[d] = rev_bits((uint32_t)d, 5);
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc4[0][n] = c;
    }
    for (int t = 1; t < 4; t++)
        for (int n = 0; n < 256; n++)
            crc4[t][n] = (crc4[t - 1][n] >> 8) ^ crc4[0][crc4[t - 1][n] & 0xFF];
    // x^(8*2^k) mod P in the reflected representation: start at x^8, square.
    auto mulmod = [](uint32_t a, uint32_t b) {
        uint32_t p = 0;
        for (int k = 0; k < 32; k++) {
            if (a & (0x80000000u >> k)) p ^= b;
            b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
        }
        return p;
    };
    uint32_t x = 0x00800000u;  // x^8
    for (int k = 0; k < 64; k++) { x8pow[k] = x; x = mulmod(x, x); }
    // shift[k]: the register after 64 * 2^k zero bytes (c -> c * x^(8*64*2^k)),
    // split by input byte so that it is four lookups.
    for (int k = 0; k < kCrcShiftLevels; k++)
        for (int b = 0; b < 4; b++)
            for (uint32_t v = 0; v < 256; v++) shift[k][b][v] = mulmod(x8pow[6 + k], v << (8 * b));
    // the same for 32 zero bytes (the assembly joins two 32-byte CRC chains)
    for (int b = 0; b < 4; b++)
        for (uint32_t v = 0; v < 256; v++) shift32[b][v] = mulmod(x8pow[5], v << (8 * b));
}

static std::mutex g_init_mu;
static bool g_dev_ready[64];
static int g_cu[64];

// Compute units of the current device (256 on MI355X).
int cu_count()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64 || g_cu[dev] <= 0) return 256;
    return g_cu[dev];
}

// Scratch of r2p_encode_device_noalloc: every buffer of one plain encode in
// one static device array (each device holds its own copy until the process
// ends).  In host code `scratch` names a host-side stand-in, not a device
// address: r2p_init looks up each device's address once, and only that
// address is used.
static __device__ __align__(256) uint8_t scratch[R2P_NOALLOC_SCRATCH_BYTES];
static void* g_scratch[64];

// ---------------------------------------------------------------------------
// geometry / planning
// ---------------------------------------------------------------------------
int resolve(const r2p_image* img, Geometry* g)
{
    if (!img) { set_error("null image"); return R2P_ERR_ARG; }
    if (img->width == 0 || img->height == 0 || img->channels < 1 || img->channels > 4 ||
        (img->in_bits != 8 && img->in_bits != 16)) {
        set_error("bad image geometry %ux%u ch=%d bits=%d", img->width, img->height, img->channels,
                  img->in_bits);
        return R2P_ERR_ARG;
    }
    int ob = img->in_bits;
    switch (img->pix_mode) {
    case R2P_PIX_READY: break;
    case R2P_PIX_SWAP16:
        if (img->in_bits != 16) { set_error("SWAP16 needs 16-bit input"); return R2P_ERR_ARG; }
        break;
    case R2P_PIX_DICOM:
        if (img->dicom.bits_allocated != img->in_bits) {
            set_error("DICOM bits_allocated %d != in_bits %d", img->dicom.bits_allocated, img->in_bits);
            return R2P_ERR_ARG;
        }
        ob = img->dicom.bits_allocated;
        break;
    case R2P_PIX_DICOM8:
        if (img->in_bits != 16) { set_error("DICOM8 needs 16-bit input"); return R2P_ERR_ARG; }
        ob = 8;
        break;
    default: set_error("bad pix_mode %d", img->pix_mode); return R2P_ERR_ARG;
    }
    g->out_bits = ob;
    g->bpp = img->channels * (ob / 8);
    g->row_bytes = (size_t)img->width * (size_t)g->bpp;
    g->filtered = (size_t)img->height * (g->row_bytes + 1);
    g->in_row = (size_t)img->width * img->channels * (img->in_bits / 8);
    g->pitch = img->src_pitch ? img->src_pitch : (int64_t)g->in_row;
    if (g->pitch < (int64_t)g->in_row) { set_error("src_pitch < row bytes"); return R2P_ERR_ARG; }
    // channels_to_color_type (pipeline.cpp:80-88)
    g->color_type = img->channels == 1 ? 0 : img->channels == 3 ? 2 : img->channels == 4 ? 6 : 2;
    if (img->channels == 2) g->color_type = 4;
    return R2P_OK;
}

bool q15_coeffs(const r2p_dicom_params& p, int16_t* a_fp, int32_t* b_fp)
{
    const int shift = p.high_bit - p.bits_stored + 1;
    if (shift != 0) return false;
    float A_f, B_f;
    if (p.apply_window && p.window_width > 0.f) {
        const float lo = p.window_center - p.window_width * 0.5f;
        const float slope = p.apply_rescale ? p.rescale_slope : 1.f;
        const float intercept = p.apply_rescale ? p.rescale_intercept : 0.f;
        A_f = slope * 255.f / p.window_width;
        B_f = (intercept - lo) * 255.f / p.window_width;
    } else if (!p.apply_window) {
        const float slope = p.apply_rescale ? p.rescale_slope : 1.f;
        const float intercept = p.apply_rescale ? p.rescale_intercept : 0.f;
        const float range = (float)(1u << p.bits_stored);
        if (p.pixel_rep == 1) {
            const float half = range * 0.5f;
            A_f = slope * 255.f / range;
            B_f = (intercept + half) * 255.f / range;
        } else {
            A_f = slope * 255.f / (range - 1.f);
            B_f = intercept * 255.f / (range - 1.f);
        }
    } else {
        return false;
    }
    const float A_q15 = A_f * 32768.f;
    if (A_q15 < -32767.f || A_q15 > 32767.f) return false;
    *a_fp = (int16_t)(int)A_q15;
    float bq;
    if (A_f >= 0.f) bq = (A_q15 == (float)(int)A_q15) ? B_f * 32768.f : B_f * 32768.f + 0.5f;
    else            bq = B_f * 32768.f - 0.5f;
    *b_fp = (int32_t)bq;
    return true;
}

static int check_dicom(const r2p_dicom_params& p, int need_alloc)
{
    if (p.bits_allocated != need_alloc) { set_error("bits_allocated %d unsupported here", p.bits_allocated); return R2P_ERR_ARG; }
    if (p.bits_stored < 1 || p.bits_stored > p.bits_allocated) { set_error("bits_stored %d out of range", p.bits_stored); return R2P_ERR_ARG; }
    if (p.high_bit < p.bits_stored - 1 || p.high_bit >= p.bits_allocated) {
        set_error("high_bit %d inconsistent with bits_stored %d", p.high_bit, p.bits_stored);
        return R2P_ERR_ARG;
    }
    if (p.pixel_rep != 0 && p.pixel_rep != 1) { set_error("pixel_rep %d", p.pixel_rep); return R2P_ERR_ARG; }
    return R2P_OK;
}

int plan_window(const r2p_image* img, WindowPlan* wp)
{
    *wp = WindowPlan{};
    const r2p_dicom_params& p = img->dicom;
    switch (img->pix_mode) {
    case R2P_PIX_READY: wp->wmode = WM_COPY; return R2P_OK;
    case R2P_PIX_SWAP16: wp->wmode = WM_SWAP16; return R2P_OK;
    case R2P_PIX_DICOM: {
        int rc = check_dicom(p, img->in_bits);
        if (rc) return rc;
        if (p.bits_allocated == 8) {
            wp->wmode = WM_LUT8_FROM8;
            wp->lut_kind = 0; wp->lut_entries = 256; wp->lut_bytes = 256; wp->lds_bytes = 256;
            return R2P_OK;
        }
        const int sh = p.high_bit - p.bits_stored + 1;
        wp->wmode = WM_LUT16;
        wp->shift = sh > 0 ? sh : 0;
        wp->mask = p.bits_stored < 16 ? (1u << p.bits_stored) - 1u : 0xFFFFu;
        wp->lut_kind = 0;
        wp->lut_entries = 1 << p.bits_stored;
        wp->lut_bytes = 2 * wp->lut_entries;
        wp->lds_bytes = wp->lut_bytes;
        if (wp->lut_bytes > 65536) {
            // 16 bits stored: the 128 KiB table stays in global memory (read
            // through the caches) instead of taking the LDS the rows need
            wp->wmode = WM_LUT16G;
            wp->lds_bytes = 0;
        }
        return R2P_OK;
    }
    case R2P_PIX_DICOM8: {
        int rc = check_dicom(p, 16);
        if (rc) return rc;
        int16_t a; int32_t b;
        if (q15_coeffs(p, &a, &b)) {
            wp->wmode = WM_Q15; wp->q15_a = a; wp->q15_b = b; wp->pixel_rep = p.pixel_rep;
            return R2P_OK;
        }
        wp->wmode = WM_LUT8_FROM16;
        wp->shift = p.high_bit - p.bits_stored + 1;
        wp->lut_kind = 1;
        wp->lut_entries = 1 << p.bits_stored;
        wp->mask = (uint32_t)(wp->lut_entries - 1);
        wp->lut_bytes = wp->lut_entries;
        wp->lds_bytes = wp->lut_bytes;
        return R2P_OK;
    }
    default: set_error("bad pix_mode"); return R2P_ERR_ARG;
    }
}

LutCache::~LutCache()
{
    if (d_lut) (void)hipFree(d_lut);
}

int LutCache::ensure(const WindowPlan& wp, const r2p_dicom_params& pp, hipStream_t s)
{
    if (wp.lut_kind < 0) return R2P_OK;
    if (valid && kind == wp.lut_kind && entries == wp.lut_entries && memcmp(&p, &pp, sizeof(pp)) == 0)
        return R2P_OK;
    const size_t need = ((size_t)wp.lut_bytes + 15) & ~(size_t)15;
    if (need > cap) {
        if (d_lut) R2P_HIP(hipFree(d_lut));
        d_lut = nullptr;
        R2P_HIP(hipMalloc(&d_lut, std::max<size_t>(need, 4096)));
        cap = std::max<size_t>(need, 4096);
    }
    LutBuild lb;
    lb.p = pp;
    lb.entries = wp.lut_entries;
    lb.kind = wp.lut_kind;
    R2P_HIP(launch_build_lut(lb, d_lut, s));
    valid = true; kind = wp.lut_kind; entries = wp.lut_entries; p = pp;
    return R2P_OK;
}

size_t png_bound_bytes(size_t filtered, int segment_bytes, int chunk_bytes)
{
    const size_t nseg = (filtered + segment_bytes - 1) / segment_bytes;
    const size_t z = filtered + filtered / 8 + 16 * nseg + 16;
    const size_t nch = (z + chunk_bytes - 1) / chunk_bytes + 1;
    return 8 + 25 + 12 * nch + z + 12 + 64;
}

uint32_t ihdr_crc(uint32_t width, uint32_t height, uint8_t bit_depth, uint8_t color_type)
{
    const uint8_t ih[17] = {'I', 'H', 'D', 'R',
                            (uint8_t)(width >> 24), (uint8_t)(width >> 16), (uint8_t)(width >> 8), (uint8_t)width,
                            (uint8_t)(height >> 24), (uint8_t)(height >> 16), (uint8_t)(height >> 8), (uint8_t)height,
                            bit_depth, color_type, 0, 0, 0};
    return (uint32_t)crc32(0L, ih, 17);
}

size_t filter_lds_bytes(int band_rows, int row_cap, int lut_cap)
{
    return (size_t)lut_cap + (size_t)(band_rows + 1) * (size_t)(row_cap + 32);
}

size_t filter_tiled_lds_bytes(int lut_cap)
{
    return (size_t)lut_cap + 2 * (size_t)(kTileBytes + 32);
}

int filter_band_rows(int requested, int64_t total_rows, int row_cap, int lut_cap)
{
    (void)total_rows;
    auto fits = [&](int br) { return filter_lds_bytes(br, row_cap, lut_cap) <= (size_t)kFilterLdsMax; };
    if (!filter_split(row_cap)) {
        // k_filter: one row per wave of the workgroup by default -- the
        // band's rows are filtered in one pass, and the small LDS footprint
        // keeps several workgroups per CU (on the mixed batch, 8 and 16 rows
        // measured slower)
        int h = requested > 0 ? requested : kFilterThreads / 64;
        while (h > 1 && !fits(h)) h--;
        return h;
    }
    // k_filter_split: a band of BR rows is filtered by 8 / BR waves per row,
    // K items per lane: the height that leaves the fewest idle lanes (then
    // the fewest items per lane), with its BR + 1 rows in LDS beside the table
    const int items = (row_cap + 15) / 16;
    if (requested > 0) {
        int br = 1;
        while (br * 2 <= std::min(requested, kFilterRowThreads / 64)) br *= 2;
        while (br > 1 && (!fits(br) || filter_items_per_lane(items, br) > kFilterMaxK)) br >>= 1;
        return br;
    }
    int best = 1, best_k = 1 << 30;
    double best_u = -1.0;
    for (int br = kFilterRowThreads / 64; br >= 1; br >>= 1) {
        const int k = filter_items_per_lane(items, br);
        if (!fits(br) || k > kFilterMaxK) continue;
        const double u = (double)items / ((double)k * 64.0 * (kFilterRowThreads / 64 / br));
        if (u > best_u + 1e-9 || (u > best_u - 1e-9 && k < best_k)) { best = br; best_u = u; best_k = k; }
    }
    return best;
}


int filter_grid(const FilterLaunch& L)
{
    // k_filter_split is persistent (kFilterSplitWgPerCu workgroups per CU,
    // each a run of bands); k_filter takes one band per workgroup, or loops
    // over the bands of a job table whose count is read on the device
    if (filter_split(L.row_cap)) {
        const uint32_t g = (uint32_t)(kFilterSplitWgPerCu * cu_count());
        return (int)(L.counts ? g : std::min<uint32_t>(L.total_bands, g));
    }
    return (int)(L.counts ? 4096u : std::min<uint32_t>(L.total_bands, 8192u));
}

int resolve_slice_bytes(const r2p_options& o, size_t filtered)
{
    if (o.slice_bytes > 0) return o.slice_bytes;
    int sl = 16384;
    while (sl > 4096 && ((size_t)sl * 512 > filtered || o.segment_bytes % sl)) sl >>= 1;
    return sl;
}

int slice_cap_for(int slice_bytes, int huffman)
{
    // a dynamic-Huffman slice may take up to 16 bits per byte (15-bit
    // literal codes; a match of 3 or more bytes is at most 48 bits) plus the
    // segment's header
    return huffman ? (16 * slice_bytes + 32 * kHuffHdrWords + 31) / 32 + 8 : slice_cap_dwords(slice_bytes);
}

void slice_arena_bound(const r2p_options& o, size_t filtered, size_t* max_slices, size_t* max_dwords)
{
    size_t ns = 0, nd = 0;
    for (int sl = 4096; sl <= 65536; sl <<= 1) {
        if (o.slice_bytes > 0 && sl != o.slice_bytes) continue;
        if (o.slice_bytes == 0 && sl > 16384) continue;
        ns = std::max(ns, slice_count_bound(filtered, sl));
        nd = std::max(nd, slice_dword_bound(filtered, sl, slice_cap_for(sl, o.huffman)));
    }
    *max_slices = ns;
    *max_dwords = nd;
}

void fill_defaults(r2p_options* o)
{
    if (o->segment_bytes <= 0) o->segment_bytes = 65536;
    o->segment_bytes = std::max(4096, (o->segment_bytes + 4095) & ~4095);
    // a slice is a power of two from 4 KiB up that divides the segment (0: auto, per stream)
    if (o->slice_bytes < 0) o->slice_bytes = 0;
    int sl = o->slice_bytes;
    if (sl && (sl > o->segment_bytes || (sl & (sl - 1)) || sl < 4096 || o->segment_bytes % sl)) {
        sl = 4096;
        while (sl * 2 <= std::min(o->segment_bytes, 16384) && o->segment_bytes % (sl * 2) == 0) sl *= 2;
    }
    o->slice_bytes = sl;
    if (o->idat_chunk_bytes <= 0) o->idat_chunk_bytes = 65536;
    if (o->idat_chunk_bytes > 65536) o->idat_chunk_bytes = 65536;
    o->lane_bytes = 16;
    if (o->wg_threads != 256 && o->wg_threads != 512) o->wg_threads = 256;
    if (o->lz77 < 0) o->lz77 = 1;
    if (o->lz77 > 2) o->lz77 = 2;
    o->huffman = o->huffman ? 1 : 0;
}

}  // namespace r2p

using namespace r2p;

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* r2p_last_error(void) { return tl_err; }
const char* r2p_version(void) { return "r2p-gfx950 0.1"; }

int r2p_init(int device)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) { set_error("no HIP device"); return R2P_ERR_NODEV; }
    if (device < 0 || device >= n || device >= 64) { set_error("device %d out of range", device); return R2P_ERR_ARG; }
    R2P_HIP(hipSetDevice(device));
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_dev_ready[device]) return R2P_OK;
    hipDeviceProp_t prop;
    R2P_HIP(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx9", 4) != 0) {
        set_error("device %d is %s, this build targets gfx950", device, prop.gcnArchName);
        return R2P_ERR_NODEV;
    }
    static uint32_t lit[256], len[259], dist_rev[32], crc4[4][256], x8pow[64], shift[kCrcShiftLevels][4][256],
        shift32[4][256];
    build_tables(lit, len, dist_rev, crc4, x8pow, shift, shift32);
    R2P_HIP(upload_tables(lit, len, dist_rev, &crc4[0][0], x8pow, &shift[0][0][0], &shift32[0][0]));
    g_cu[device] = prop.multiProcessorCount;
    void* sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !sp) return hip_fail(se != hipSuccess ? se : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = sp;
    g_dev_ready[device] = true;
    return R2P_OK;
}

static int current_init()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    return r2p_init(dev);
}

int r2p_image_out_bits(const r2p_image* img)
{
    Geometry g;
    return resolve(img, &g) == R2P_OK ? g.out_bits : -1;
}
size_t r2p_image_row_bytes(const r2p_image* img)
{
    Geometry g;
    return resolve(img, &g) == R2P_OK ? g.row_bytes : 0;
}
size_t r2p_image_filtered_bytes(const r2p_image* img)
{
    Geometry g;
    return resolve(img, &g) == R2P_OK ? g.filtered : 0;
}
size_t r2p_image_input_bytes(const r2p_image* img)
{
    Geometry g;
    if (resolve(img, &g) != R2P_OK) return 0;
    return (size_t)g.pitch * (img->height - 1) + g.in_row;
}
size_t r2p_png_bound(const r2p_image* img, int segment_bytes, int idat_chunk_bytes)
{
    Geometry g;
    if (resolve(img, &g) != R2P_OK) return 0;
    r2p_options o{};
    o.segment_bytes = segment_bytes; o.idat_chunk_bytes = idat_chunk_bytes;
    fill_defaults(&o);
    return png_bound_bytes(g.filtered, o.segment_bytes, o.idat_chunk_bytes);
}

void r2p_options_default(r2p_options* o)
{
    memset(o, 0, sizeof(*o));
    o->segment_bytes = 65536;
    o->lz77 = 1;
    o->idat_chunk_bytes = 65536;
    o->zlib_level = 3;       // PipelineConfig::deflate_level default (pipeline.h:17)
    o->band_rows = 0;
    o->lane_bytes = 16;
    o->wg_threads = 256;
    o->slice_bytes = 0;      // auto: by the stream's size (resolve_slice_bytes)
    o->huffman = 0;
}

uint32_t r2p_adler32_combine(uint32_t a1, uint32_t a2, uint64_t len2)
{
    const uint64_t M = 65521;
    const uint64_t A1 = a1 & 0xFFFF, B1 = a1 >> 16, A2 = a2 & 0xFFFF, B2 = a2 >> 16;
    const uint64_t n = len2 % M;
    const uint64_t A = (A1 + A2 + M - 1) % M;
    const uint64_t B = (B1 + B2 + n * ((A1 + M - 1) % M)) % M;
    return (uint32_t)((B << 16) | A);
}

uint32_t r2p_crc32_combine(uint32_t c1, uint32_t c2, uint64_t len2)
{
    auto mulmod = [](uint32_t a, uint32_t b) {
        uint32_t p = 0;
        for (int k = 0; k < 32; k++) {
            if (a & (0x80000000u >> k)) p ^= b;
            b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
        }
        return p;
    };
    uint32_t r = 0x80000000u, sq = 0x00800000u;
    for (uint64_t n = len2; n; n >>= 1) {
        if (n & 1) r = mulmod(r, sq);
        sq = mulmod(sq, sq);
    }
    return mulmod(r, c1) ^ c2;
}

// ---------------------------------------------------------------------------
// filter rows (kernel-level)
// ---------------------------------------------------------------------------
static int launch_filter_job(const FilterJob& J, int band_rows, hipStream_t s)
{
    if (J.row_bytes > kMaxRowBytesLds) {
        // rows too wide for LDS: one row per workgroup, in tiles
        FilterLaunch L{};
        L.njobs = 1;
        L.band_rows = 1;
        L.lut_cap = (J.lut_bytes + 15) & ~15;
        L.row_cap = kTileBytes;
        L.total_bands = (uint32_t)J.rows;
        L.single = J;
        L.single.band_rows = 1;
        R2P_HIP(launch_filter_tiled(L, (int)std::min<uint32_t>(L.total_bands, 4096), filter_tiled_lds_bytes(L.lut_cap), s));
        return R2P_OK;
    }
    FilterLaunch L{};
    L.jobs = nullptr;
    L.band_start = nullptr;
    L.njobs = 1;
    L.row_cap = (J.row_bytes + 15) & ~15;
    L.lut_cap = (J.lut_bytes + 15) & ~15;
    L.band_rows = filter_band_rows(band_rows, J.rows, L.row_cap, L.lut_cap);
    const uint32_t bands = (uint32_t)((J.rows + L.band_rows - 1) / L.band_rows);
    L.total_bands = bands;
    L.single = J;
    L.single.band_rows = L.band_rows;
    const size_t lds = filter_lds_bytes(L.band_rows, L.row_cap, L.lut_cap);
    R2P_HIP(launch_filter(L, filter_grid(L), lds, s));
    return R2P_OK;
}

int r2p_filter_rows(const r2p_image* img, const void* d_src, int rows, const void* d_prior,
                    void* d_filtered, void* d_last_row, int band_rows, void* stream)
{
    int rc = current_init();
    if (rc) return rc;
    Geometry g;
    if ((rc = resolve(img, &g))) return rc;
    if (rows <= 0 || !d_src || !d_filtered) { set_error("bad filter arguments"); return R2P_ERR_ARG; }
    WindowPlan wp;
    if ((rc = plan_window(img, &wp))) return rc;
    static thread_local LutCache cache;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = cache.ensure(wp, img->dicom, s))) return rc;
    FilterJob J{};
    J.src = (const uint8_t*)d_src; J.dst = (uint8_t*)d_filtered;
    J.prior = (const uint8_t*)d_prior; J.last_row = (uint8_t*)d_last_row;
    J.lut = cache.d_lut; J.src_pitch = g.pitch; J.row_bytes = (int32_t)g.row_bytes; J.rows = rows;
    J.bpp = g.bpp; J.wmode = wp.wmode; J.shift = wp.shift; J.mask = wp.mask;
    J.q15_a = wp.q15_a; J.q15_b = wp.q15_b; J.pixel_rep = wp.pixel_rep;
    J.lut_bytes = wp.lut_kind >= 0 ? wp.lds_bytes : 0;
    return launch_filter_job(J, band_rows, s);
}

// ---------------------------------------------------------------------------
// checksums
// ---------------------------------------------------------------------------
size_t r2p_checksum_work_bytes(size_t len)
{
    return 4 * ((len + kChecksumBlock - 1) / kChecksumBlock + 1) + 256;
}

int r2p_adler32_device(const void* d_data, size_t len, void* d_work, uint32_t* d_result, void* stream)
{
    int rc = current_init();
    if (rc) return rc;
    if ((!d_data && len) || !d_work || !d_result) { set_error("bad checksum arguments"); return R2P_ERR_ARG; }
    R2P_HIP(launch_adler((const uint8_t*)d_data, (int64_t)len, (uint32_t*)d_work, d_result, (hipStream_t)stream));
    return R2P_OK;
}

int r2p_crc32_device(const void* d_data, size_t len, void* d_work, uint32_t* d_result, void* stream)
{
    int rc = current_init();
    if (rc) return rc;
    if ((!d_data && len) || !d_work || !d_result) { set_error("bad checksum arguments"); return R2P_ERR_ARG; }
    R2P_HIP(launch_crc((const uint8_t*)d_data, (int64_t)len, (uint32_t*)d_work, d_result, (hipStream_t)stream));
    return R2P_OK;
}

}  // extern "C"

// ===========================================================================
// Plain per-image encoder
// ===========================================================================
struct r2p_encoder {
    r2p_options opt{};
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};
    size_t max_filtered = 0;
    int max_seg = 0;
    size_t max_slice = 0;            // slice buffers for any slice size the options allow
    size_t slice_arena_dw = 0;
    uint8_t* d_filt = nullptr;
    uint32_t* d_slice_out = nullptr;
    uint32_t* d_slice_bits = nullptr;
    uint32_t* d_slice_adler = nullptr;
    uint32_t* d_slice_boff = nullptr;
    uint32_t* d_seg_tbits = nullptr;
    uint64_t* d_seg_off = nullptr;
    uint32_t* d_tokens = nullptr;    // dynamic-Huffman mode: tokens, counts, histograms,
    uint32_t* d_tok_count = nullptr; // per-segment codes and block headers
    uint32_t* d_hist = nullptr;
    uint32_t* d_codes = nullptr;
    uint32_t* d_hdr = nullptr;
    uint32_t* d_hdr_bits = nullptr;
    uint64_t* d_misc = nullptr;      // [0] png length, [1] adler
    uint64_t* h_misc = nullptr;      // pinned mirror, written by the last kernel (coherent, mapped)
    uint64_t* dh_misc = nullptr;     // its device address
    uint8_t* d_src_stage = nullptr;  // encode_host staging
    size_t src_stage_cap = 0;
    uint8_t* d_png_stage = nullptr;
    size_t png_stage_cap = 0;
    LutCache lut;
    uint32_t last_nseg = 0;
    bool timing = false;             // record the per-kernel events
    bool timed = false;              // the last encode recorded them
};

static int enc_grow(uint8_t** p, size_t* cap, size_t need)
{
    if (need <= *cap) return R2P_OK;
    if (*p) R2P_HIP(hipFree(*p));
    *p = nullptr;
    R2P_HIP(hipMalloc((void**)p, need));
    *cap = need;
    return R2P_OK;
}

// Device buffers of one plain encode.
struct EncodeWork {
    uint8_t* filt = nullptr;           // filtered bytes + kTailSlack readable bytes
    uint32_t* slice_out = nullptr;     // nslice * slice_cap_dw dwords
    uint32_t* slice_bits = nullptr;    // nslice
    uint32_t* slice_adler = nullptr;   // nslice
    uint32_t* slice_boff = nullptr;    // nslice
    uint32_t* seg_tbits = nullptr;     // nseg + 1
    uint64_t* seg_off = nullptr;       // nseg + 2
    uint32_t* tokens = nullptr;        // dynamic-Huffman blocks only
    uint32_t* tok_count = nullptr;
    uint32_t* hist = nullptr;
    uint32_t* codes = nullptr;
    uint32_t* hdr = nullptr;
    uint32_t* hdr_bits = nullptr;
    int slice_cap_dw = 0;
    int slice_cap_dw2 = 0;             // a half slice's (slice_geom)
    const void* lut = nullptr;         // the image's device table, if its window needs one
    uint64_t* out_len = nullptr;       // device: PNG length
    uint32_t* adler_out = nullptr;     // device: Adler-32 of the filtered bytes
    uint64_t* host_res = nullptr;      // optional mapped host copy of both
    hipEvent_t* ev = nullptr;          // optional: five events around the stages
    bool halve_tail = false;           // the slice size was left to the encoder (slice_geom)
};

// The plain encode's launches: window + filter, DEFLATE (and the dynamic
// blocks' tables and emission), the segment scan when the assembly cannot
// do it itself, assembly of the complete PNG.
static int encode_into(const r2p_options& o, const EncodeWork& w, const r2p_image* img, const Geometry& g,
                       const WindowPlan& wp, const void* d_src, void* d_png, size_t png_cap, void* d_filtered_out,
                       hipStream_t s, uint32_t* nseg_out)
{
    int rc;
    // The deflate stage reads past the data end (kTailSlack), so it always
    // works on its own buffer; a caller's copy is made afterwards.
    uint8_t* filt = w.filt;
    if (w.ev) R2P_HIP(hipEventRecord(w.ev[0], s));
    FilterJob J{};
    J.src = (const uint8_t*)d_src; J.dst = filt; J.prior = nullptr; J.last_row = nullptr;
    J.lut = w.lut; J.src_pitch = g.pitch; J.row_bytes = (int32_t)g.row_bytes;
    J.rows = (int32_t)img->height; J.bpp = g.bpp; J.wmode = wp.wmode; J.shift = wp.shift;
    J.mask = wp.mask; J.q15_a = wp.q15_a; J.q15_b = wp.q15_b; J.pixel_rep = wp.pixel_rep;
    J.lut_bytes = wp.lut_kind >= 0 ? wp.lds_bytes : 0;
    if ((rc = launch_filter_job(J, o.band_rows, s))) return rc;
    if (w.ev) R2P_HIP(hipEventRecord(w.ev[1], s));
    if (d_filtered_out)
        R2P_HIP(hipMemcpyAsync(d_filtered_out, filt, g.filtered, hipMemcpyDeviceToDevice, s));

    const SliceGeom sg = slice_geom((int64_t)g.filtered, o.segment_bytes, o.slice_bytes, w.halve_tail);
    const int nseg = sg.nseg;
    DeflateLaunch DL{};
    DeflateJob& D = DL.single;
    D.data = filt; D.first = 0; D.len = (int64_t)g.filtered; D.dict_lo = 0;
    D.slice_out = w.slice_out; D.slice_bits = w.slice_bits; D.slice_adler = w.slice_adler;
    D.slice_stats = nullptr; D.seg_bytes = o.segment_bytes; D.slice_bytes = o.slice_bytes;
    D.slice_cap_dw = w.slice_cap_dw; D.slice_cap_dw2 = w.slice_cap_dw2;
    D.final_block = 1; D.bpp = g.bpp; D.row_stride = (int32_t)(g.row_bytes + 1);
    D.lz77 = o.lz77; D.nslice = sg.nslice; D.half_from = sg.half_from;
    if (o.huffman) { D.tokens = w.tokens; D.tok_count = w.tok_count; D.hist = w.hist; }
    DL.dyn = o.huffman;
    DL.dense = o.lz77 == 2;
    DL.jobs = nullptr; DL.njobs = 1; DL.total_slices = (uint32_t)sg.nslice;
    R2P_HIP(launch_deflate(DL, o.lane_bytes, o.wg_threads, sg.nslice, s));
    if (o.huffman) {                       // the segments' codes, then the slices' bits
        HuffLaunch HL{};
        HuffJob& H = HL.single;
        H.tokens = w.tokens; H.tok_count = w.tok_count; H.hist = w.hist;
        H.codes = w.codes; H.hdr = w.hdr; H.hdr_bits = w.hdr_bits;
        H.slice_out = w.slice_out; H.slice_bits = w.slice_bits; H.slice_cap_dw = w.slice_cap_dw;
        H.slice_cap_dw2 = w.slice_cap_dw2;
        H.slice_bytes = o.slice_bytes; H.spb = sg.spb; H.half_from = sg.half_from; H.nslice = sg.nslice; H.nseg = nseg;
        H.final_block = 1;
        HL.jobs = nullptr; HL.njobs = 1; HL.total_segs = (uint32_t)nseg; HL.total_slices = (uint32_t)sg.nslice;
        R2P_HIP(launch_huffman(HL, nseg, std::min(sg.nslice, 8192), s));
    }
    if (w.ev) R2P_HIP(hipEventRecord(w.ev[2], s));

    AssembleLaunch AL{};
    AssembleJob& A = AL.single;
    A.slice_out = w.slice_out; A.slice_bits = w.slice_bits; A.slice_adler = w.slice_adler;
    A.slice_boff = w.slice_boff; A.seg_tbits = w.seg_tbits;
    A.seg_off = w.seg_off; A.out = (uint8_t*)d_png; A.out_len = w.out_len;
    A.adler_out = w.adler_out; A.out_cap = (int64_t)png_cap;
    A.host_res = w.host_res;
    A.total_in = (int64_t)g.filtered; A.nseg = nseg; A.nslice = sg.nslice; A.spb = sg.spb; A.half_from = sg.half_from;
    A.slice_cap_dw = w.slice_cap_dw; A.slice_cap_dw2 = w.slice_cap_dw2; A.slice_bytes = o.slice_bytes;
    A.chunk_bytes = o.idat_chunk_bytes; A.mode = 1; A.final_block = 1; A.eob_in_bits = o.huffman;
    A.width = img->width; A.height = img->height; A.bit_depth = (uint8_t)g.out_bits;
    A.color_type = (uint8_t)g.color_type;
    A.ihdr_crc = ihdr_crc(A.width, A.height, A.bit_depth, A.color_type);
    const int lvl = o.zlib_level;
    A.zflg = lvl == 1 ? 0x01 : (lvl >= 2 && lvl <= 5) ? 0x5E : (lvl >= 6 && lvl <= 7) ? 0x9C : 0xDA;
    const size_t zmax = g.filtered + g.filtered / 8 + 16 * (size_t)nseg + 16;
    A.max_chunks = (int32_t)((zmax + o.idat_chunk_bytes - 1) / o.idat_chunk_bytes);
    A.running = nullptr; A.base_slot = nullptr;
    AL.jobs = nullptr; AL.njobs = 1; AL.total_chunks = (uint32_t)A.max_chunks;
    const bool self_scan = assemble_scans_itself(A);
    if (!self_scan) R2P_HIP(launch_scan(AL, 1, s));
    if (w.ev) R2P_HIP(hipEventRecord(w.ev[3], s));
    // one workgroup per CU is resident (1024 threads, ~110 VGPRs): a grid of
    // at most that many loops over the chunks in one round, where a grid of
    // the 9/8 bound would leave a second round of workgroups with no chunk
    const int grid = std::min(A.max_chunks, cu_count());
    if (self_scan) R2P_HIP(launch_assemble_self_scan(AL, grid, s));
    else R2P_HIP(launch_assemble(AL, grid, s));
    if (w.ev) R2P_HIP(hipEventRecord(w.ev[4], s));
    if (nseg_out) *nseg_out = (uint32_t)nseg;
    return R2P_OK;
}

extern "C" {

int r2p_encoder_create(size_t max_filtered_bytes, const r2p_options* opt, r2p_encoder** out)
{
    int rc = current_init();
    if (rc) return rc;
    if (!out || max_filtered_bytes == 0) { set_error("bad encoder arguments"); return R2P_ERR_ARG; }
    r2p_encoder* e = new r2p_encoder();
    if (opt) e->opt = *opt; else r2p_options_default(&e->opt);
    fill_defaults(&e->opt);
    e->max_filtered = max_filtered_bytes;
    e->max_seg = (int)((max_filtered_bytes + e->opt.segment_bytes - 1) / e->opt.segment_bytes);
    slice_arena_bound(e->opt, max_filtered_bytes, &e->max_slice, &e->slice_arena_dw);
    hipError_t err;
#define TRY(x) do { if ((err = (x)) != hipSuccess) { r2p_encoder_destroy(e); return hip_fail(err, #x); } } while (0)
    TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    for (auto& v : e->ev) TRY(hipEventCreate(&v));
    TRY(hipMalloc((void**)&e->d_filt, max_filtered_bytes + kTailSlack));
    TRY(hipMalloc((void**)&e->d_slice_out, e->slice_arena_dw * 4));
    TRY(hipMalloc((void**)&e->d_slice_bits, (size_t)e->max_slice * 4));
    TRY(hipMalloc((void**)&e->d_slice_adler, (size_t)e->max_slice * 4));
    TRY(hipMalloc((void**)&e->d_slice_boff, (size_t)e->max_slice * 4));
    TRY(hipMalloc((void**)&e->d_seg_tbits, ((size_t)e->max_seg + 1) * 4));
    TRY(hipMalloc((void**)&e->d_seg_off, ((size_t)e->max_seg + 2) * 8));
    if (e->opt.huffman) {
        TRY(hipMalloc((void**)&e->d_tokens, (max_filtered_bytes + 65536) * 2));
        TRY(hipMalloc((void**)&e->d_tok_count, (size_t)e->max_slice * 4));
        TRY(hipMalloc((void**)&e->d_hist, ((size_t)e->max_seg + 1) * kHuffSyms * 4));
        TRY(hipMemset(e->d_hist, 0, ((size_t)e->max_seg + 1) * kHuffSyms * 4));   // kept zero by the kernels
        TRY(hipMalloc((void**)&e->d_codes, ((size_t)e->max_seg + 1) * kHuffTab * 4));
        TRY(hipMalloc((void**)&e->d_hdr, ((size_t)e->max_seg + 1) * kHuffHdrWords * 4));
        TRY(hipMalloc((void**)&e->d_hdr_bits, ((size_t)e->max_seg + 1) * 4));
    }
    TRY(hipMalloc((void**)&e->d_misc, 64));
    TRY(hipHostMalloc((void**)&e->h_misc, 64, hipHostMallocMapped | hipHostMallocCoherent));
    TRY(hipHostGetDevicePointer((void**)&e->dh_misc, e->h_misc, 0));
#undef TRY
    *out = e;
    return R2P_OK;
}

int r2p_encoder_destroy(r2p_encoder* e)
{
    if (!e) return R2P_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    (void)hipFree(e->d_filt); (void)hipFree(e->d_slice_out); (void)hipFree(e->d_slice_bits);
    (void)hipFree(e->d_slice_adler); (void)hipFree(e->d_slice_boff); (void)hipFree(e->d_seg_tbits);
    (void)hipFree(e->d_seg_off); (void)hipFree(e->d_misc);
    (void)hipFree(e->d_tokens); (void)hipFree(e->d_tok_count); (void)hipFree(e->d_hist);
    (void)hipFree(e->d_codes); (void)hipFree(e->d_hdr); (void)hipFree(e->d_hdr_bits);
    if (e->h_misc) (void)hipHostFree(e->h_misc);
    (void)hipFree(e->d_src_stage); (void)hipFree(e->d_png_stage);
    for (auto& v : e->ev) if (v) (void)hipEventDestroy(v);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return R2P_OK;
}

void* r2p_encoder_stream(r2p_encoder* e) { return e ? (void*)e->stream : nullptr; }

int r2p_encode_device(r2p_encoder* e, const r2p_image* img, const void* d_src, void* d_png,
                      size_t png_cap, void* d_filtered_out)
{
    if (!e || !img || !d_src || !d_png) { set_error("bad encode arguments"); return R2P_ERR_ARG; }
    Geometry g;
    int rc = resolve(img, &g);
    if (rc) return rc;
    if (g.filtered > e->max_filtered) {
        set_error("image needs %zu filtered bytes, encoder sized for %zu", g.filtered, e->max_filtered);
        return R2P_ERR_SIZE;
    }
    r2p_options o = e->opt;
    o.slice_bytes = resolve_slice_bytes(e->opt, g.filtered);
    const size_t bound = png_bound_bytes(g.filtered, o.segment_bytes, o.idat_chunk_bytes);
    if (png_cap < bound) { set_error("png capacity %zu < bound %zu", png_cap, bound); return R2P_ERR_SIZE; }
    WindowPlan wp;
    if ((rc = plan_window(img, &wp))) return rc;
    hipStream_t s = e->stream;
    if ((rc = e->lut.ensure(wp, img->dicom, s))) return rc;
    EncodeWork w;
    w.filt = e->d_filt; w.slice_out = e->d_slice_out; w.slice_bits = e->d_slice_bits;
    w.slice_adler = e->d_slice_adler; w.slice_boff = e->d_slice_boff; w.seg_tbits = e->d_seg_tbits;
    w.seg_off = e->d_seg_off; w.tokens = e->d_tokens; w.tok_count = e->d_tok_count; w.hist = e->d_hist;
    w.codes = e->d_codes; w.hdr = e->d_hdr; w.hdr_bits = e->d_hdr_bits;
    w.slice_cap_dw = slice_cap_for(o.slice_bytes, o.huffman);
    w.slice_cap_dw2 = slice_cap_for(o.slice_bytes / 2, o.huffman);
    w.lut = e->lut.d_lut;
    w.out_len = e->d_misc; w.adler_out = (uint32_t*)(e->d_misc + 1);
    w.host_res = e->dh_misc;               // the results reach h_misc without a copy
    w.ev = e->timing ? e->ev : nullptr;
    // (dynamic-Huffman streams keep whole slices: a highly compressible
    // tail, such as a saturated border, becomes slices of a few bytes, and
    // the assembly gathers pieces across many of them byte by byte)
    w.halve_tail = e->opt.slice_bytes == 0 && !o.huffman;
    uint32_t nseg = 0;
    if ((rc = encode_into(o, w, img, g, wp, d_src, d_png, png_cap, d_filtered_out, s, &nseg))) return rc;
    e->last_nseg = nseg;
    e->timed = e->timing;
    return R2P_OK;
}

// Carves the plain encode's buffers for F filtered bytes out of `base`;
// returns the bytes used.
static size_t noalloc_layout(uint8_t* base, size_t F, const r2p_options& o, const WindowPlan& wp, EncodeWork* w)
{
    const SliceGeom sg = slice_geom((int64_t)F, o.segment_bytes, o.slice_bytes, w->halve_tail);
    size_t off = 0;
    auto take = [&](size_t n) { uint8_t* p = base + off; off += (n + 255) & ~(size_t)255; return p; };
    w->slice_cap_dw = slice_cap_for(o.slice_bytes, 0);
    w->slice_cap_dw2 = slice_cap_for(o.slice_bytes / 2, 0);
    w->filt = take(F + kTailSlack);
    w->slice_out = (uint32_t*)take((size_t)lay_buf(sg.nslice, sg.spb, sg.half_from, w->slice_cap_dw, w->slice_cap_dw2) * 4);
    w->slice_bits = (uint32_t*)take((size_t)sg.nslice * 4);
    w->slice_adler = (uint32_t*)take((size_t)sg.nslice * 4);
    w->slice_boff = (uint32_t*)take((size_t)sg.nslice * 4);
    w->seg_tbits = (uint32_t*)take(((size_t)sg.nseg + 1) * 4);
    w->seg_off = (uint64_t*)take(((size_t)sg.nseg + 2) * 8);
    w->lut = wp.lut_kind >= 0 ? take((size_t)wp.lut_bytes) : nullptr;
    return off;
}

int r2p_encode_device_noalloc(const r2p_image* img, const void* d_src, void* d_png, size_t png_cap,
                              void* d_result, const r2p_options* opt, void* stream)
{
    int rc = current_init();              // (tables and the scratch address, once per device)
    if (rc) return rc;
    if (!img || !d_src || !d_png || !d_result) { set_error("bad encode arguments"); return R2P_ERR_ARG; }
    int dev = 0;
    R2P_HIP(hipGetDevice(&dev));
    uint8_t* base = (uint8_t*)g_scratch[dev];
    if (!base) { set_error("device %d has no scratch address", dev); return R2P_ERR_STATE; }
    r2p_options o;
    if (opt) o = *opt; else r2p_options_default(&o);
    fill_defaults(&o);
    if (o.huffman) { set_error("dynamic Huffman blocks need the encoder's own buffers (r2p_encode_device)"); return R2P_ERR_UNSUPPORTED; }
    Geometry g;
    if ((rc = resolve(img, &g))) return rc;
    WindowPlan wp;
    if ((rc = plan_window(img, &wp))) return rc;
    const size_t bound = png_bound_bytes(g.filtered, o.segment_bytes, o.idat_chunk_bytes);
    if (png_cap < bound) { set_error("png capacity %zu < bound %zu", png_cap, bound); return R2P_ERR_SIZE; }
    EncodeWork w;
    w.halve_tail = o.slice_bytes == 0;
    o.slice_bytes = resolve_slice_bytes(o, g.filtered);
    const size_t need = noalloc_layout(base, g.filtered, o, wp, &w);
    if (need > (size_t)R2P_NOALLOC_SCRATCH_BYTES) {
        set_error("image needs %zu bytes of scratch (%zu filtered bytes); the no-allocation encoder holds %zu",
                  need, g.filtered, (size_t)R2P_NOALLOC_SCRATCH_BYTES);
        return R2P_ERR_SIZE;
    }
    hipStream_t s = (hipStream_t)stream;
    // Nothing in the scratch has to start at zero: every kernel writes what
    // it later reads (tables, filtered bytes, slice bits and offsets).
    if (wp.lut_kind >= 0) {
        LutBuild lb;
        lb.p = img->dicom;
        lb.entries = wp.lut_entries;
        lb.kind = wp.lut_kind;
        R2P_HIP(launch_build_lut(lb, (void*)w.lut, s));
    }
    w.out_len = (uint64_t*)d_result;
    w.adler_out = (uint32_t*)((uint8_t*)d_result + 8);
    w.host_res = nullptr;
    return encode_into(o, w, img, g, wp, d_src, d_png, png_cap, nullptr, s, nullptr);
}

int r2p_encoder_sync(r2p_encoder* e, size_t* png_len, uint32_t* adler, uint32_t* nseg)
{
    if (!e) { set_error("null encoder"); return R2P_ERR_ARG; }
    R2P_HIP(hipStreamSynchronize(e->stream));
    if (png_len) *png_len = (size_t)e->h_misc[0];
    if (adler) *adler = (uint32_t)e->h_misc[1];
    if (nseg) *nseg = e->last_nseg;
    return R2P_OK;
}

int r2p_encode_host(r2p_encoder* e, const r2p_image* img, const void* h_src, void* h_png,
                    size_t png_cap, size_t* png_len)
{
    if (!e || !img || !h_src || !h_png) { set_error("bad encode arguments"); return R2P_ERR_ARG; }
    Geometry g;
    int rc = resolve(img, &g);
    if (rc) return rc;
    const size_t in_bytes = (size_t)g.pitch * (img->height - 1) + g.in_row;
    const size_t bound = png_bound_bytes(g.filtered, e->opt.segment_bytes, e->opt.idat_chunk_bytes);
    if ((rc = enc_grow(&e->d_src_stage, &e->src_stage_cap, in_bytes))) return rc;
    if ((rc = enc_grow(&e->d_png_stage, &e->png_stage_cap, bound))) return rc;
    R2P_HIP(hipMemcpyAsync(e->d_src_stage, h_src, in_bytes, hipMemcpyHostToDevice, e->stream));
    if ((rc = r2p_encode_device(e, img, e->d_src_stage, e->d_png_stage, e->png_stage_cap, nullptr))) return rc;
    size_t n = 0;
    if ((rc = r2p_encoder_sync(e, &n, nullptr, nullptr))) return rc;
    if (n > png_cap) { set_error("png of %zu bytes exceeds host capacity %zu", n, png_cap); return R2P_ERR_SIZE; }
    R2P_HIP(hipMemcpy(h_png, e->d_png_stage, n, hipMemcpyDeviceToHost));
    if (png_len) *png_len = n;
    return R2P_OK;
}

int r2p_encoder_set_timing(r2p_encoder* e, int on)
{
    if (!e) { set_error("null encoder"); return R2P_ERR_ARG; }
    e->timing = on != 0;
    return R2P_OK;
}

int r2p_encoder_timings(r2p_encoder* e, r2p_timings* t)
{
    if (!e || !t || !e->timed) { set_error("no timings: the last encode ran with timing off"); return R2P_ERR_STATE; }
    R2P_HIP(hipEventSynchronize(e->ev[4]));
    R2P_HIP(hipEventElapsedTime(&t->filter_ms, e->ev[0], e->ev[1]));
    R2P_HIP(hipEventElapsedTime(&t->deflate_ms, e->ev[1], e->ev[2]));
    R2P_HIP(hipEventElapsedTime(&t->scan_ms, e->ev[2], e->ev[3]));
    R2P_HIP(hipEventElapsedTime(&t->assemble_ms, e->ev[3], e->ev[4]));
    R2P_HIP(hipEventElapsedTime(&t->total_ms, e->ev[0], e->ev[4]));
    return R2P_OK;
}

}  // extern "C"
