// r2p_batch.cpp -- batched form: many images of uneven size and bit depth
// encoded by one fixed launch sequence (window+filter -> deflate -> scan ->
// assemble) over device job tables, captured once in a HIP graph.
//
// The kernels are persistent: fixed grids loop over the work items (row
// bands, 64 KiB segments, IDAT chunks) of all images, and read the item
// counts from device memory, so the captured graph replays unchanged for any
// batch that fits the capacities fixed at create time.  Per batch the host
// only rewrites the pinned job tables, which the graph's first node uploads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <set>
#include <thread>
#include <unordered_map>
#include <vector>

#include "r2p_common.h"

using namespace r2p;

namespace {

struct LutKey {
    r2p_dicom_params p;
    int kind, entries;
    bool operator<(const LutKey& o) const
    {
        if (kind != o.kind) return kind < o.kind;
        if (entries != o.entries) return entries < o.entries;
        return memcmp(&p, &o.p, sizeof(p)) < 0;
    }
};

constexpr int kLutSlot = 131072;     // bytes per table slot (the largest table: 16 bits stored, u16 out)
constexpr int kNarrowRow = 4096;     // rows up to this many bytes go to the first filter launch

}  // namespace

struct r2p_batch {
    r2p_options opt{};
    int max_images = 0;
    size_t max_input = 0, max_filtered = 0;
    size_t max_segs = 0, max_slices = 0, slice_arena_dw = 0, max_out = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};
    // device arenas
    uint8_t* d_filt = nullptr;
    uint32_t* d_slice_out = nullptr;
    uint32_t* d_slice_bits = nullptr;
    uint32_t* d_slice_adler = nullptr;
    uint32_t* d_slice_boff = nullptr;
    uint32_t* d_seg_tbits = nullptr;
    uint64_t* d_seg_off = nullptr;
    uint8_t* d_out = nullptr;
    uint8_t* d_lut = nullptr;
    // device tables (one allocation) and their pinned host mirror
    void* d_tab = nullptr;
    void* h_tab = nullptr;
    size_t tab_bytes = 0;
    FilterJob* d_fj = nullptr; DeflateJob* d_dj = nullptr; AssembleJob* d_aj = nullptr; HuffJob* d_hj = nullptr;
    uint32_t* d_band_start = nullptr; uint32_t* d_slice_start = nullptr; uint32_t* d_chunk_start = nullptr;
    uint32_t* d_seg_start = nullptr;
    uint32_t* d_counts = nullptr;      // launch sizes: per launch [jobs, total, first] (see r2p_batch_set)
    // dynamic-Huffman blocks (allocated while the options ask for them):
    // token slots, counts and histograms per slice, codes and block headers
    // per segment, and the slice arena at the larger per-slice capacity
    uint32_t* d_tokens = nullptr;
    uint32_t* d_tok_count = nullptr;
    uint32_t* d_hist = nullptr;
    uint32_t* d_codes = nullptr;
    uint32_t* d_hdr = nullptr;
    uint32_t* d_hdr_bits = nullptr;
    size_t dyn_tok_dw = 0, dyn_segs = 0, dyn_slices = 0;
    uint64_t* d_res = nullptr;         // per image: png length, adler
    uint64_t* h_res = nullptr;         // pinned
    size_t off_fj, off_dj, off_aj, off_hj, off_bs, off_ss, off_cs, off_gs, off_cnt;
    // current batch
    int n = 0;
    std::vector<size_t> out_off;
    std::vector<std::pair<LutKey, int>> luts;
    // filter launches: rows up to kNarrowRow bytes, then wider rows, each
    // with its own items per lane, band height and LDS size
    int band_rows = 1, row_cap = 16, lut_cap = 0;
    int band_rows_w = 1, row_cap_w = 16;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_row_cap = -1, g_lut_cap = -1, g_band = -1, g_lane = -1, g_wg = -1, g_row_cap_w = -1, g_band_w = -1,
        g_huff = -1;
    bool ran = false;
    bool timed_kernels = false;
    std::vector<r2p_image> shadow_img;
    std::vector<const void*> shadow_src;
    // slice size per size class of the filtered stream (the autotuner's
    // choice; 0: by resolve_slice_bytes) and the one each image got
    int class_slice[3] = {0, 0, 0};
    std::vector<int> job_slice;
};

// Size classes of the autotuner: filtered streams under 2 MiB, under 8 MiB, larger.
static int size_class(size_t filtered) { return filtered < ((size_t)2 << 20) ? 0 : filtered < ((size_t)8 << 20) ? 1 : 2; }

static void free_dynamic(r2p_batch* b)
{
    (void)hipFree(b->d_tokens); (void)hipFree(b->d_tok_count); (void)hipFree(b->d_hist);
    (void)hipFree(b->d_codes); (void)hipFree(b->d_hdr); (void)hipFree(b->d_hdr_bits);
    b->d_tokens = b->d_tok_count = b->d_hist = b->d_codes = b->d_hdr = b->d_hdr_bits = nullptr;
    b->dyn_tok_dw = b->dyn_segs = b->dyn_slices = 0;
}

// The dynamic-Huffman buffers for the current options (none without them):
// per slice up to slice_bytes 16-bit token slots (a job's slices cover its
// stream rounded up to a slice: filtered + 64 KiB per image bounds them) and
// a count; per segment (of the options' size) a histogram (zeroed here, kept
// zero by the kernels), the codes and the header; and a slice arena at
// slice_cap_for(.., 1) dwords per slice.
static int ensure_dynamic(r2p_batch* b)
{
    if (!b->opt.huffman) return R2P_OK;
    const size_t n = (size_t)b->max_images;
    const size_t tok_dw = (b->max_filtered + n * 65536) / 2 + 1024;
    const size_t segs = b->max_filtered / (size_t)b->opt.segment_bytes + n + 1;
    const size_t slices = b->max_slices;
    const size_t arena = (b->max_filtered + n * 65536) / 2 + (size_t)(slice_cap_for(4096, 1) - 2048) * slices + 1024;
    if (b->d_tokens && b->dyn_tok_dw >= tok_dw && b->dyn_segs >= segs && b->dyn_slices >= slices &&
        b->slice_arena_dw >= arena)
        return R2P_OK;
    free_dynamic(b);
    hipError_t e;
#define TRY(x) do { if ((e = (x)) != hipSuccess) { free_dynamic(b); return hip_fail(e, #x); } } while (0)
    TRY(hipMalloc((void**)&b->d_tokens, tok_dw * 4));
    TRY(hipMalloc((void**)&b->d_tok_count, slices * 4));
    TRY(hipMalloc((void**)&b->d_hist, segs * kHuffSyms * 4));
    TRY(hipMemset(b->d_hist, 0, segs * kHuffSyms * 4));     // kept zero by the kernels
    TRY(hipMalloc((void**)&b->d_codes, segs * kHuffTab * 4));
    TRY(hipMalloc((void**)&b->d_hdr, segs * kHuffHdrWords * 4));
    TRY(hipMalloc((void**)&b->d_hdr_bits, segs * 4));
    if (b->slice_arena_dw < arena) {
        (void)hipFree(b->d_slice_out);
        b->d_slice_out = nullptr;
        b->slice_arena_dw = 0;
        TRY(hipMalloc((void**)&b->d_slice_out, arena * 4));
        b->slice_arena_dw = arena;
    }
#undef TRY
    b->dyn_tok_dw = tok_dw; b->dyn_segs = segs; b->dyn_slices = slices;
    return R2P_OK;
}

// Enqueues one batch run.  Per-kernel events are recorded only for eager
// runs (timed); a captured graph is timed as a whole around its launch.
static int batch_enqueue(r2p_batch* b, hipStream_t s, bool timed)
{
    // 1. job tables
    R2P_HIP(hipMemcpyAsync(b->d_tab, b->h_tab, b->tab_bytes, hipMemcpyHostToDevice, s));
    // 2. window + filter over all row bands
    FilterLaunch L{};
    L.jobs = b->d_fj; L.band_start = b->d_band_start; L.njobs = 0; L.band_rows = b->band_rows;
    L.total_bands = 0; L.row_cap = b->row_cap; L.lut_cap = b->lut_cap; L.counts = b->d_counts;
    L.queue = b->d_counts + 17;           // zero: uploaded with the tables
    if (timed) R2P_HIP(hipEventRecord(b->ev[0], s));
    R2P_HIP(launch_filter(L, filter_grid(L), filter_lds_bytes(L.band_rows, L.row_cap, L.lut_cap), s));
    FilterLaunch LW = L;                          // the wide rows: bands after the narrow ones
    LW.band_rows = b->band_rows_w; LW.row_cap = b->row_cap_w; LW.counts = b->d_counts + 3;
    LW.queue = filter_split(LW.row_cap) ? nullptr : b->d_counts + 18;
    R2P_HIP(launch_filter(LW, filter_grid(LW), filter_lds_bytes(LW.band_rows, LW.row_cap, LW.lut_cap), s));
    FilterLaunch LH = L;                          // rows too wide for LDS: one row per workgroup, tiled
    LH.band_rows = 1; LH.row_cap = kTileBytes; LH.counts = b->d_counts + 10; LH.queue = nullptr;
    R2P_HIP(launch_filter_tiled(LH, 2048, filter_tiled_lds_bytes(LH.lut_cap), s));
    if (timed) R2P_HIP(hipEventRecord(b->ev[1], s));
    // 3. deflate over all slices
    DeflateLaunch DL{};
    DL.jobs = b->d_dj; DL.slice_start = b->d_slice_start; DL.counts = b->d_counts + 6; DL.dyn = b->opt.huffman;
    DL.dense = b->opt.lz77 == 2;
    DL.queue = b->d_counts + 16;          // zero: uploaded with the tables
    R2P_HIP(launch_deflate(DL, b->opt.lane_bytes, b->opt.wg_threads, 1024, s));
    if (b->opt.huffman) {                 // every segment's codes, then every slice's bits
        HuffLaunch HL{};
        HL.jobs = b->d_hj; HL.seg_start = b->d_seg_start; HL.slice_start = b->d_slice_start;
        HL.counts = b->d_counts + 13;
        R2P_HIP(launch_huffman(HL, 2048, 2048, s));
    }
    if (timed) R2P_HIP(hipEventRecord(b->ev[2], s));
    // 4. per-image scan, 5. assemble all chunks
    AssembleLaunch AL{};
    AL.jobs = b->d_aj; AL.chunk_start = b->d_chunk_start; AL.counts = b->d_counts + 8;
    R2P_HIP(launch_scan(AL, b->max_images, s));
    if (timed) R2P_HIP(hipEventRecord(b->ev[3], s));
    R2P_HIP(launch_assemble(AL, 2048, s));
    if (timed) R2P_HIP(hipEventRecord(b->ev[4], s));
    // 6. lengths and adlers back to pinned host memory
    R2P_HIP(hipMemcpyAsync(b->h_res, b->d_res, (size_t)b->max_images * 16, hipMemcpyDeviceToHost, s));
    return R2P_OK;
}

extern "C" {

int r2p_batch_create(int max_images, size_t max_input_bytes, size_t max_filtered_bytes,
                     const r2p_options* opt, r2p_batch** out)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    int rc = r2p_init(dev);
    if (rc) return rc;
    if (!out || max_images <= 0 || max_filtered_bytes == 0) { set_error("bad batch arguments"); return R2P_ERR_ARG; }
    r2p_batch* b = new r2p_batch();
    if (opt) b->opt = *opt; else r2p_options_default(&b->opt);
    fill_defaults(&b->opt);
    b->max_images = max_images;
    b->max_input = max_input_bytes;
    b->max_filtered = max_filtered_bytes;
    // arenas sized for any segment and slice size from 4 KiB up: per image
    // ceil(len / slice) buffers of slice_cap_dwords(slice) dwords
    b->max_segs = max_filtered_bytes / 4096 + (size_t)max_images + 1;
    b->max_slices = b->max_segs;
    // (+ 20 dwords per slice: two half slices of a full one, slice_geom)
    b->slice_arena_dw = max_filtered_bytes * 9 / 32 + 9 * (max_filtered_bytes / 4096) +
                        (size_t)max_images * (size_t)(slice_cap_dwords(65536) + 16) + 20 * b->max_slices + 1024;
    b->max_out = max_filtered_bytes + max_filtered_bytes / 8 + 32 * b->max_segs + 4096 * (size_t)max_images;
    b->off_fj = 0;
    b->off_dj = b->off_fj + sizeof(FilterJob) * (size_t)max_images;
    b->off_aj = b->off_dj + sizeof(DeflateJob) * (size_t)max_images;
    b->off_hj = b->off_aj + sizeof(AssembleJob) * (size_t)max_images;
    b->off_bs = b->off_hj + sizeof(HuffJob) * (size_t)max_images;
    b->off_ss = b->off_bs + 4 * ((size_t)max_images + 1);
    b->off_cs = b->off_ss + 4 * ((size_t)max_images + 1);
    b->off_gs = b->off_cs + 4 * ((size_t)max_images + 1);
    b->off_cnt = b->off_gs + 4 * ((size_t)max_images + 1);
    b->tab_bytes = b->off_cnt + 128;
    hipError_t e;
#define TRY(x) do { if ((e = (x)) != hipSuccess) { r2p_batch_destroy(b); return hip_fail(e, #x); } } while (0)
    TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (auto& v : b->ev) TRY(hipEventCreate(&v));
    TRY(hipMalloc((void**)&b->d_filt, max_filtered_bytes + 256 * (size_t)max_images));
    TRY(hipMalloc((void**)&b->d_slice_out, b->slice_arena_dw * 4));
    TRY(hipMalloc((void**)&b->d_slice_bits, b->max_slices * 4));
    TRY(hipMalloc((void**)&b->d_slice_adler, b->max_slices * 4));
    TRY(hipMalloc((void**)&b->d_slice_boff, b->max_slices * 4));
    TRY(hipMalloc((void**)&b->d_seg_tbits, b->max_segs * 4));
    TRY(hipMalloc((void**)&b->d_seg_off, (b->max_segs + (size_t)max_images + 1) * 8));
    TRY(hipMalloc((void**)&b->d_out, b->max_out));
    TRY(hipMalloc((void**)&b->d_lut, (size_t)kLutSlot * (size_t)max_images));
    TRY(hipMalloc(&b->d_tab, b->tab_bytes));
    TRY(hipHostMalloc(&b->h_tab, b->tab_bytes, hipHostMallocDefault));
    TRY(hipMalloc((void**)&b->d_res, (size_t)max_images * 16));
    TRY(hipHostMalloc((void**)&b->h_res, (size_t)max_images * 16, hipHostMallocDefault));
#undef TRY
    memset(b->h_tab, 0, b->tab_bytes);
    uint8_t* dt = (uint8_t*)b->d_tab;
    b->d_fj = (FilterJob*)(dt + b->off_fj);
    b->d_dj = (DeflateJob*)(dt + b->off_dj);
    b->d_aj = (AssembleJob*)(dt + b->off_aj);
    b->d_hj = (HuffJob*)(dt + b->off_hj);
    b->d_band_start = (uint32_t*)(dt + b->off_bs);
    b->d_slice_start = (uint32_t*)(dt + b->off_ss);
    b->d_chunk_start = (uint32_t*)(dt + b->off_cs);
    b->d_seg_start = (uint32_t*)(dt + b->off_gs);
    b->d_counts = (uint32_t*)(dt + b->off_cnt);
    if ((rc = ensure_dynamic(b))) { r2p_batch_destroy(b); return rc; }
    *out = b;
    return R2P_OK;
}

int r2p_batch_destroy(r2p_batch* b)
{
    if (!b) return R2P_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    (void)hipFree(b->d_filt); (void)hipFree(b->d_slice_out); (void)hipFree(b->d_slice_bits);
    (void)hipFree(b->d_slice_adler); (void)hipFree(b->d_slice_boff); (void)hipFree(b->d_seg_tbits);
    (void)hipFree(b->d_seg_off); (void)hipFree(b->d_out);
    (void)hipFree(b->d_lut); (void)hipFree(b->d_tab); (void)hipFree(b->d_res);
    free_dynamic(b);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    if (b->h_res) (void)hipHostFree(b->h_res);
    for (auto& v : b->ev) if (v) (void)hipEventDestroy(v);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return R2P_OK;
}

int r2p_batch_options(r2p_batch* b, const r2p_options* opt)
{
    if (!b || !opt) { set_error("bad options"); return R2P_ERR_ARG; }
    r2p_options o = *opt;
    fill_defaults(&o);
    R2P_HIP(hipStreamSynchronize(b->stream));   // a run in flight may use the buffers
    b->opt = o;
    b->class_slice[0] = b->class_slice[1] = b->class_slice[2] = 0;
    return ensure_dynamic(b);
}

int r2p_batch_job_options(r2p_batch* b, int i, r2p_options* opt)
{
    if (!b || !opt || i < 0 || i >= b->n || (size_t)i >= b->job_slice.size()) { set_error("bad batch job index"); return R2P_ERR_ARG; }
    *opt = b->opt;
    opt->slice_bytes = b->job_slice[(size_t)i];
    return R2P_OK;
}

int r2p_batch_set(r2p_batch* b, int n, const r2p_image* imgs, const void* const* d_src)
{
    if (!b || n <= 0 || !imgs || !d_src) { set_error("bad batch set arguments"); return R2P_ERR_ARG; }
    if (n > b->max_images) { set_error("%d images exceed batch capacity %d", n, b->max_images); return R2P_ERR_SIZE; }
    R2P_HIP(hipStreamSynchronize(b->stream));   // previous run no longer reads the tables
    const r2p_options& o = b->opt;
    uint8_t* ht = (uint8_t*)b->h_tab;
    FilterJob* fj = (FilterJob*)(ht + b->off_fj);
    DeflateJob* dj = (DeflateJob*)(ht + b->off_dj);
    AssembleJob* aj = (AssembleJob*)(ht + b->off_aj);
    HuffJob* hj = (HuffJob*)(ht + b->off_hj);
    uint32_t* gs = (uint32_t*)(ht + b->off_gs);
    uint32_t* bs = (uint32_t*)(ht + b->off_bs);
    uint32_t* ss = (uint32_t*)(ht + b->off_ss);
    uint32_t* cs = (uint32_t*)(ht + b->off_cs);
    uint32_t* cnt = (uint32_t*)(ht + b->off_cnt);

    // band height: auto from the batch's total rows unless fixed by the options
    int64_t total_rows = 0;
    size_t total_f = 0;
    std::vector<Geometry> geo((size_t)n);
    std::vector<WindowPlan> plan((size_t)n);
    for (int i = 0; i < n; i++) {
        int rc = resolve(&imgs[i], &geo[i]);
        if (rc) { set_error("image %d: %s", i, r2p_last_error()); return rc; }
        if ((rc = plan_window(&imgs[i], &plan[i]))) return rc;
        if (plan[i].lut_bytes > kLutSlot) { set_error("image %d table too large", i); return R2P_ERR_UNSUPPORTED; }
        total_rows += imgs[i].height;
        total_f += geo[i].filtered;
        if (!d_src[i]) { set_error("image %d has no source", i); return R2P_ERR_ARG; }
    }
    if (total_f > b->max_filtered) { set_error("batch needs %zu filtered bytes > %zu", total_f, b->max_filtered); return R2P_ERR_SIZE; }
    // the filter's three classes, in this order in the job table: rows of up
    // to kNarrowRow bytes, rows up to kMaxRowBytesLds, wider rows (tiled);
    // the first two get their own band heights
    auto row_class = [&](int i) {
        const size_t rc = (geo[i].row_bytes + 15) & ~(size_t)15;
        return rc <= (size_t)kNarrowRow ? 0 : rc <= (size_t)kMaxRowBytesLds ? 1 : 2;
    };
    std::vector<int> forder;
    {
        int rn = 16, rw = 16, lmax = 0;
        int64_t rows_n = 0, rows_w = 0;
        for (int i = 0; i < n; i++) {
            const int rc = (int)std::min<size_t>((geo[i].row_bytes + 15) & ~(size_t)15, (size_t)kMaxRowBytesLds);
            if (row_class(i) == 0) { rn = std::max(rn, rc); rows_n += imgs[i].height; }
            else if (row_class(i) == 1) { rw = std::max(rw, rc); rows_w += imgs[i].height; }
            if (plan[i].lut_kind >= 0) lmax = std::max(lmax, (plan[i].lds_bytes + 15) & ~15);
        }
        for (int c = 0; c < 3; c++)
            for (int i = 0; i < n; i++) if (row_class(i) == c) forder.push_back(i);
        b->row_cap = rn; b->row_cap_w = rw; b->lut_cap = lmax;
        b->band_rows = filter_band_rows(o.band_rows, rows_n, rn, lmax);
        b->band_rows_w = filter_band_rows(o.band_rows, rows_w, rw, lmax);
    }
    std::vector<int> fslot((size_t)n);
    for (int k = 0; k < n; k++) fslot[(size_t)forder[(size_t)k]] = k;

    // device tables, deduplicated by parameter set
    std::map<LutKey, int> lut_slot;
    b->luts.clear();
    size_t foff = 0, segs = 0, slices = 0, ooff = 0, slice_dw = 0, tok_dw = 0;
    uint32_t chunks = 0;
    b->out_off.assign((size_t)n, 0);
    b->job_slice.assign((size_t)n, 0);
    for (int i = 0; i < n; i++) {
        const Geometry& g = geo[i];
        const WindowPlan& wp = plan[i];
        const uint8_t* lut = nullptr;
        if (wp.lut_kind >= 0) {
            LutKey k{imgs[i].dicom, wp.lut_kind, wp.lut_entries};
            auto it = lut_slot.find(k);
            int slot;
            if (it == lut_slot.end()) {
                slot = (int)lut_slot.size();
                lut_slot[k] = slot;
                b->luts.push_back({k, slot});
            } else {
                slot = it->second;
            }
            lut = b->d_lut + (size_t)slot * kLutSlot;
        }
        if (foff + g.filtered > b->max_filtered + 256 * (size_t)i) { set_error("filtered arena exhausted"); return R2P_ERR_SIZE; }
        FilterJob& F = fj[fslot[(size_t)i]];
        F = FilterJob{};
        F.src = (const uint8_t*)d_src[i]; F.dst = b->d_filt + foff; F.prior = nullptr; F.last_row = nullptr;
        F.lut = lut; F.src_pitch = g.pitch; F.row_bytes = (int32_t)g.row_bytes; F.rows = (int32_t)imgs[i].height;
        F.bpp = g.bpp; F.wmode = wp.wmode; F.shift = wp.shift; F.mask = wp.mask; F.q15_a = wp.q15_a;
        F.q15_b = wp.q15_b; F.pixel_rep = wp.pixel_rep; F.lut_bytes = wp.lut_kind >= 0 ? wp.lds_bytes : 0;
        F.band_rows = row_class(i) == 0 ? b->band_rows : row_class(i) == 1 ? b->band_rows_w : 1;

        // the slice size of this image: its size class's tuned one, else by
        // its size -- then, as in the plain form with the slice size left to
        // it, the stream's last partial round of slices is cut in half, and
        // the job's options say "auto" (0)
        int sl = b->class_slice[size_class(g.filtered)];
        const bool auto_sl = (sl <= 0 || o.segment_bytes % sl) && o.slice_bytes == 0;
        if (sl <= 0 || o.segment_bytes % sl) sl = resolve_slice_bytes(o, g.filtered);
        b->job_slice[(size_t)i] = auto_sl ? 0 : sl;
        const int cap_dw = slice_cap_for(sl, o.huffman), cap_dw2 = slice_cap_for(sl / 2, o.huffman);
        const SliceGeom sg = slice_geom((int64_t)g.filtered, o.segment_bytes, sl, auto_sl && !o.huffman);
        const int nseg = sg.nseg;
        const size_t dw_i = (size_t)lay_buf(sg.nslice, sg.spb, sg.half_from, cap_dw, cap_dw2);
        // token slots: the slices' positions, each slice's region rounded up to its size
        const size_t tok_i = (size_t)(lay_start(sg.nslice - 1, sg.spb, sg.half_from, sl) +
                                      lay_bytes(sg.nslice - 1, sg.spb, sg.half_from, sl) + 1) / 2;
        if (slice_dw + dw_i > b->slice_arena_dw || slices + sg.nslice > b->max_slices ||
            segs + nseg > b->max_segs ||
            (o.huffman && (tok_dw + tok_i > b->dyn_tok_dw || segs + nseg > b->dyn_segs))) {
            set_error("slice arena exhausted"); return R2P_ERR_SIZE;
        }
        DeflateJob& D = dj[i];
        D = DeflateJob{};
        D.data = F.dst; D.first = 0; D.len = (int64_t)g.filtered; D.dict_lo = 0;
        D.slice_out = b->d_slice_out + slice_dw; D.slice_bits = b->d_slice_bits + slices;
        D.slice_adler = b->d_slice_adler + slices; D.slice_stats = nullptr; D.seg_bytes = o.segment_bytes;
        D.slice_bytes = sl; D.slice_cap_dw = cap_dw; D.slice_cap_dw2 = cap_dw2; D.final_block = 1; D.bpp = g.bpp;
        D.row_stride = (int32_t)(g.row_bytes + 1); D.lz77 = o.lz77; D.nslice = sg.nslice; D.half_from = sg.half_from;
        ss[i] = (uint32_t)slices;
        gs[i] = (uint32_t)segs;
        if (o.huffman) {
            D.tokens = b->d_tokens + tok_dw; D.tok_count = b->d_tok_count + slices;
            D.hist = b->d_hist + segs * kHuffSyms;
            HuffJob& H = hj[i];
            H = HuffJob{};
            H.tokens = D.tokens; H.tok_count = D.tok_count; H.hist = D.hist;
            H.codes = b->d_codes + segs * kHuffTab; H.hdr = b->d_hdr + segs * kHuffHdrWords;
            H.hdr_bits = b->d_hdr_bits + segs; H.slice_out = D.slice_out; H.slice_bits = D.slice_bits;
            H.slice_cap_dw = cap_dw; H.slice_cap_dw2 = cap_dw2; H.slice_bytes = sl; H.spb = sg.spb;
            H.half_from = sg.half_from;
            H.nslice = sg.nslice; H.nseg = nseg;
            H.final_block = 1;
            tok_dw += tok_i;
        }

        const size_t bound = png_bound_bytes(g.filtered, o.segment_bytes, o.idat_chunk_bytes);
        if (ooff + bound > b->max_out) { set_error("output arena exhausted"); return R2P_ERR_SIZE; }
        AssembleJob& A = aj[i];
        A = AssembleJob{};
        A.slice_out = D.slice_out; A.slice_bits = D.slice_bits; A.slice_adler = D.slice_adler;
        A.slice_boff = b->d_slice_boff + slices; A.seg_tbits = b->d_seg_tbits + segs;
        A.seg_off = b->d_seg_off + segs + (size_t)i; A.out = b->d_out + ooff;
        A.out_len = b->d_res + 2 * (size_t)i; A.adler_out = (uint32_t*)(b->d_res + 2 * (size_t)i + 1);
        A.out_cap = (int64_t)bound; A.total_in = (int64_t)g.filtered; A.nseg = nseg; A.nslice = sg.nslice;
        A.spb = sg.spb; A.half_from = sg.half_from; A.slice_cap_dw = cap_dw; A.slice_cap_dw2 = cap_dw2;
        A.slice_bytes = sl;
        A.chunk_bytes = o.idat_chunk_bytes; A.mode = 1; A.final_block = 1; A.eob_in_bits = o.huffman;
        A.width = imgs[i].width; A.height = imgs[i].height; A.bit_depth = (uint8_t)g.out_bits;
        A.color_type = (uint8_t)g.color_type;
        A.ihdr_crc = ihdr_crc(A.width, A.height, A.bit_depth, A.color_type);
        const int lvl = o.zlib_level;
        A.zflg = lvl == 1 ? 0x01 : (lvl >= 2 && lvl <= 5) ? 0x5E : (lvl >= 6 && lvl <= 7) ? 0x9C : 0xDA;
        const size_t zmax = g.filtered + g.filtered / 8 + 16 * (size_t)nseg + 16;
        A.max_chunks = (int32_t)((zmax + o.idat_chunk_bytes - 1) / o.idat_chunk_bytes);
        cs[i] = chunks;
        chunks += (uint32_t)A.max_chunks;
        b->out_off[i] = ooff;
        ooff += (bound + 255) & ~(size_t)255;
        foff += (g.filtered + 255) & ~(size_t)255;
        segs += (size_t)nseg;
        slices += (size_t)sg.nslice;
        slice_dw += dw_i;
    }
    // band prefix in job-table order: the narrow class, the wide one, the tiled one
    uint32_t bands = 0, bands_n = 0, bands_w = 0;
    for (int k = 0; k < n; k++) {
        bs[k] = bands;
        bands += (uint32_t)((fj[k].rows + fj[k].band_rows - 1) / fj[k].band_rows);
        const int c = row_class(forder[(size_t)k]);
        if (c == 0) bands_n = bands;
        if (c <= 1) bands_w = bands;
    }
    bands_w = std::max(bands_w, bands_n);
    bs[n] = bands; ss[n] = (uint32_t)slices; cs[n] = chunks; gs[n] = (uint32_t)segs;
    cnt[0] = (uint32_t)n; cnt[1] = bands_n; cnt[2] = 0;            // filter, narrow rows
    cnt[3] = (uint32_t)n; cnt[4] = bands_w; cnt[5] = bands_n;       // filter, wide rows
    cnt[10] = (uint32_t)n; cnt[11] = bands; cnt[12] = bands_w;      // filter, tiled rows
    cnt[6] = (uint32_t)n; cnt[7] = (uint32_t)slices;
    cnt[8] = (uint32_t)n; cnt[9] = chunks;
    cnt[13] = (uint32_t)n; cnt[14] = (uint32_t)segs; cnt[15] = (uint32_t)slices;   // dynamic-Huffman tables, emit
    cnt[16] = 0;                                                    // deflate slice queue
    cnt[17] = 0; cnt[18] = 0;                                       // filter band queues (narrow, wide)
    b->n = n;
    // build the distinct tables now (outside the graph)
    for (auto& kv : b->luts) {
        LutBuild lb;
        lb.p = kv.first.p; lb.entries = kv.first.entries; lb.kind = kv.first.kind;
        R2P_HIP(launch_build_lut(lb, b->d_lut + (size_t)kv.second * kLutSlot, b->stream));
    }
    b->ran = false;
    if (b->shadow_img.data() != imgs) {
        b->shadow_img.assign(imgs, imgs + n);
        b->shadow_src.assign(d_src, d_src + n);
    }
    return R2P_OK;
}

int r2p_batch_run(r2p_batch* b, int use_graph)
{
    if (!b || b->n <= 0) { set_error("batch not set"); return R2P_ERR_STATE; }
    hipStream_t s = b->stream;
    if (!use_graph) {
        int rc = batch_enqueue(b, s, true);
        if (rc) return rc;
        b->ran = true;
        b->timed_kernels = true;
        return R2P_OK;
    }
    // graph keyed by the launch parameters that are baked into its nodes
    const int lane = b->opt.lane_bytes;
    if (!b->gexec || b->g_row_cap != b->row_cap || b->g_lut_cap != b->lut_cap ||
        b->g_band != b->band_rows || b->g_lane != lane || b->g_wg != b->opt.wg_threads ||
        b->g_row_cap_w != b->row_cap_w || b->g_band_w != b->band_rows_w || b->g_huff != b->opt.huffman) {
        if (b->gexec) { R2P_HIP(hipGraphExecDestroy(b->gexec)); b->gexec = nullptr; }
        if (b->graph) { R2P_HIP(hipGraphDestroy(b->graph)); b->graph = nullptr; }
        R2P_HIP(hipStreamSynchronize(s));
        R2P_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = batch_enqueue(b, s, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        b->graph = g;
        R2P_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
        b->g_row_cap = b->row_cap; b->g_lut_cap = b->lut_cap; b->g_band = b->band_rows; b->g_lane = lane; b->g_wg = b->opt.wg_threads;
        b->g_row_cap_w = b->row_cap_w; b->g_band_w = b->band_rows_w; b->g_huff = b->opt.huffman;
    }
    R2P_HIP(hipEventRecord(b->ev[0], s));
    R2P_HIP(hipGraphLaunch(b->gexec, s));
    R2P_HIP(hipEventRecord(b->ev[4], s));
    b->ran = true;
    b->timed_kernels = false;
    return R2P_OK;
}

int r2p_batch_sync(r2p_batch* b)
{
    if (!b) { set_error("null batch"); return R2P_ERR_ARG; }
    R2P_HIP(hipStreamSynchronize(b->stream));
    return R2P_OK;
}

int r2p_batch_result(r2p_batch* b, int i, const void** d_png, size_t* len)
{
    if (!b || i < 0 || i >= b->n || !b->ran) { set_error("bad batch result index"); return R2P_ERR_ARG; }
    if (d_png) *d_png = b->d_out + b->out_off[(size_t)i];
    if (len) *len = (size_t)b->h_res[2 * (size_t)i];
    return R2P_OK;
}

int r2p_batch_copy_result(r2p_batch* b, int i, void* h_dst, size_t cap, size_t* len)
{
    const void* d = nullptr;
    size_t n = 0;
    int rc = r2p_batch_result(b, i, &d, &n);
    if (rc) return rc;
    if (n > cap) { set_error("png of %zu bytes exceeds %zu", n, cap); return R2P_ERR_SIZE; }
    R2P_HIP(hipMemcpy(h_dst, d, n, hipMemcpyDeviceToHost));
    if (len) *len = n;
    return R2P_OK;
}

int r2p_batch_timings(r2p_batch* b, r2p_timings* t)
{
    if (!b || !t || !b->ran) { set_error("no batch timings"); return R2P_ERR_STATE; }
    R2P_HIP(hipEventSynchronize(b->ev[4]));
    *t = r2p_timings{};
    if (b->timed_kernels) {
        R2P_HIP(hipEventElapsedTime(&t->filter_ms, b->ev[0], b->ev[1]));
        R2P_HIP(hipEventElapsedTime(&t->deflate_ms, b->ev[1], b->ev[2]));
        R2P_HIP(hipEventElapsedTime(&t->scan_ms, b->ev[2], b->ev[3]));
        R2P_HIP(hipEventElapsedTime(&t->assemble_ms, b->ev[3], b->ev[4]));
    }
    R2P_HIP(hipEventElapsedTime(&t->total_ms, b->ev[0], b->ev[4]));
    return R2P_OK;
}

// Autotuner: times candidates on the declared batch -- per candidate one
// warm-up graph replay, then the best of three timed replays -- first the
// launch-wide knobs (DEFLATE workgroup size x filter band height, slices by
// image size), then, one size class at a time (filtered streams under 2 MiB,
// under 8 MiB, larger), the slice size of that class's images, keeping the
// fastest of each step and re-declaring the batch with it.
static int time_batch(r2p_batch* b, int n, const std::vector<r2p_image>& imgs, const std::vector<const void*>& srcs,
                      double* t_out)
{
    int rc = r2p_batch_set(b, n, imgs.data(), srcs.data());
    double t_min = 1e30;
    for (int rep = 0; rep < 4 && !rc; rep++) {
        rc = r2p_batch_run(b, 1);
        if (!rc) rc = r2p_batch_sync(b);
        r2p_timings t{};
        if (!rc && rep > 0 && r2p_batch_timings(b, &t) == R2P_OK) t_min = std::min(t_min, (double)t.total_ms);
    }
    *t_out = t_min;
    return rc;
}

int r2p_batch_autotune(r2p_batch* b, r2p_options* chosen, double* best_ms)
{
    if (!b || b->n <= 0) { set_error("autotune needs a declared batch"); return R2P_ERR_STATE; }
    const std::vector<r2p_image> imgs = b->shadow_img;
    const std::vector<const void*> srcs = b->shadow_src;
    const int n = b->n;
    r2p_options best = b->opt;
    best.slice_bytes = 0;
    double best_t = 1e30;
    static const int shapes[] = {256, 512};   // DEFLATE wg_threads
    static const int bands[] = {0, 2, 8};     // filter band rows (0: by row width)
    for (int sh : shapes)
        for (int bd : bands) {
            r2p_options o = best;
            o.wg_threads = sh; o.band_rows = bd;
            int rc = r2p_batch_options(b, &o);
            double t = 1e30;
            if (!rc) rc = time_batch(b, n, imgs, srcs, &t);
            if (rc) return rc;
            if (t < best_t) { best_t = t; best.wg_threads = sh; best.band_rows = bd; }
        }
    int rc = r2p_batch_options(b, &best);
    if (rc) return rc;
    // which size classes the batch has
    bool has[3] = {false, false, false};
    for (int i = 0; i < n; i++) has[size_class(r2p_image_filtered_bytes(&imgs[(size_t)i]))] = true;
    int cls[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++) {
        if (!has[c]) continue;
        for (int sl = 4096; sl <= 65536; sl <<= 1) {
            if (best.segment_bytes % sl) continue;
            for (int q = 0; q < 3; q++) b->class_slice[q] = cls[q];
            b->class_slice[c] = sl;
            double t = 1e30;
            if ((rc = time_batch(b, n, imgs, srcs, &t))) return rc;
            if (t < best_t) { best_t = t; cls[c] = sl; }
        }
    }
    for (int q = 0; q < 3; q++) b->class_slice[q] = cls[q];
    if ((rc = r2p_batch_set(b, n, imgs.data(), srcs.data()))) return rc;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    return R2P_OK;
}

int r2p_batch_class_slices(r2p_batch* b, int* slices3)
{
    if (!b || !slices3) { set_error("bad arguments"); return R2P_ERR_ARG; }
    for (int q = 0; q < 3; q++) slices3[q] = b->class_slice[q];
    return R2P_OK;
}

}  // extern "C"

// ===========================================================================
// Served form: continuous batching under an HBM budget.
//
// A scheduler thread owns `inflight_batches` lanes.  Each lane is a batch
// (its own stream, arenas and captured graph) plus a device input arena and
// a pinned output staging buffer.  Whenever a lane is free and requests are
// waiting, the scheduler admits waiting requests FIFO while their device
// footprint fits both the lane and the global HBM budget, uploads their
// pixels, replays the lane's graph and goes on; lanes run concurrently on
// the device, so requests that arrive while others are on the GPU start as
// soon as a lane frees up.  A finished lane copies its PNGs back, completes
// its requests and returns their footprint to the budget.
// ===========================================================================
namespace {

using SClock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    r2p_image img{};
    const void* h_src = nullptr;
    size_t in_bytes = 0, filtered = 0, footprint = 0;
    SClock::time_point arrival{}, done_at{};
    uint8_t* png = nullptr;            // PNG bytes, a buffer of the server's pool
    size_t png_len = 0, png_cap = 0;
    int status = R2P_OK;
    bool done = false;
};

struct Lane {
    r2p_batch* b = nullptr;
    uint8_t* d_in = nullptr;
    size_t in_cap = 0;
    hipEvent_t ev_run = nullptr, ev_copy = nullptr;
    int state = 0;  // 0 free, 1 running, 2 copying
    std::vector<Request*> reqs;
    size_t footprint = 0;
};

}  // namespace

// Recycled pinned host buffers for finished PNGs: the device copies each
// PNG straight into its request's buffer (no staging copy on the host), and
// pinning a fresh buffer costs far more than reusing one.
struct BufPool {
    std::mutex mu;
    std::multimap<size_t, uint8_t*> free;      // capacity -> buffer
    std::set<uint8_t*> pageable;               // buffers that could not be pinned
    size_t kept = 0;
    static constexpr size_t kKeep = (size_t)8 << 30;
    uint8_t* get(size_t n, size_t* cap)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = free.lower_bound(n);
            if (it != free.end() && it->first <= 2 * n + (1 << 16)) {
                uint8_t* p = it->second;
                *cap = it->first;
                kept -= it->first;
                free.erase(it);
                return p;
            }
        }
        size_t c = (size_t)1 << 16;
        while (c < n) c <<= 1;
        *cap = c;
        uint8_t* p = nullptr;
        if (hipHostMalloc((void**)&p, c, hipHostMallocDefault) == hipSuccess) return p;
        // pinned memory exhausted: a pageable buffer (a slower copy, same bytes)
        (void)hipGetLastError();
        p = new (std::nothrow) uint8_t[c];
        if (p) { std::lock_guard<std::mutex> lk(mu); pageable.insert(p); }
        return p;
    }
    void release(uint8_t* p)            // with mu held
    {
        auto it = pageable.find(p);
        if (it != pageable.end()) { pageable.erase(it); delete[] p; }
        else (void)hipHostFree(p);
    }
    void put(uint8_t* p, size_t cap)
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu);
        if (kept + cap > kKeep) { release(p); return; }
        free.emplace(cap, p);
        kept += cap;
    }
    ~BufPool() { for (auto& kv : free) release(kv.second); }
};

struct r2p_server {
    r2p_server_config cfg{};
    BufPool pool;
    std::vector<Lane> lanes;
    size_t lane_filtered = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> waiting;
    std::unordered_map<int64_t, Request*> all;
    int64_t next_id = 0;
    size_t admitted = 0;
    bool stop = false;
    std::thread th;
    r2p_server_stats st{};
    uint64_t batch_images = 0;
    int dev = 0;
};

static size_t req_footprint(size_t in_b, size_t filt)
{
    // input + filtered + segment scratch (<= 9/8) + PNG (<= 9/8) + slack
    return in_b + filt + 2 * (filt + filt / 8) + 8192;
}

static void lane_launch(r2p_server* s, Lane& L)
{
    const int n = (int)L.reqs.size();
    std::vector<r2p_image> imgs((size_t)n);
    std::vector<const void*> srcs((size_t)n);
    size_t off = 0;
    hipStream_t st = L.b->stream;
    for (int i = 0; i < n; i++) {
        Request* r = L.reqs[(size_t)i];
        imgs[(size_t)i] = r->img;
        srcs[(size_t)i] = L.d_in + off;
        off += (r->in_bytes + 255) & ~(size_t)255;
    }
    int rc = r2p_batch_set(L.b, n, imgs.data(), srcs.data());
    off = 0;
    for (int i = 0; i < n && !rc; i++) {
        Request* r = L.reqs[(size_t)i];
        if (hipMemcpyAsync(L.d_in + off, r->h_src, r->in_bytes, hipMemcpyHostToDevice, st) != hipSuccess) rc = R2P_ERR_HIP;
        off += (r->in_bytes + 255) & ~(size_t)255;
    }
    // eager launches: the batch's shape (widest row, table size) changes from
    // one admission to the next, and each change would re-capture a graph
    if (!rc) rc = r2p_batch_run(L.b, 0);
    if (!rc && hipEventRecord(L.ev_run, st) != hipSuccess) rc = R2P_ERR_HIP;
    if (rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        for (Request* r : L.reqs) { r->status = rc; r->done = true; r->done_at = SClock::now(); }
        s->admitted -= L.footprint;
        L.reqs.clear(); L.footprint = 0; L.state = 0;
        s->cv_done.notify_all();
        return;
    }
    L.state = 1;
}

// The batch is done: each PNG goes from the device straight into its
// request's pinned buffer.
static void lane_copy(r2p_server* s, Lane& L)
{
    hipStream_t st = L.b->stream;
    for (size_t i = 0; i < L.reqs.size(); i++) {
        Request* r = L.reqs[i];
        const void* d = nullptr;
        size_t len = 0;
        r2p_batch_result(L.b, (int)i, &d, &len);
        r->png_len = len;
        if (!len) continue;
        r->png = s->pool.get(len, &r->png_cap);
        if (!r->png || hipMemcpyAsync(r->png, d, len, hipMemcpyDeviceToHost, st) != hipSuccess) {
            r->status = R2P_ERR_HIP;
            r->png_len = 0;
        }
    }
    (void)hipEventRecord(L.ev_copy, st);
    L.state = 2;
}

static void lane_finish(r2p_server* s, Lane& L)
{
    const auto now = SClock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request* r : L.reqs) { r->done_at = now; r->done = true; }
    s->admitted -= L.footprint;
    s->st.completed += L.reqs.size();
    L.reqs.clear();
    L.footprint = 0;
    L.state = 0;
    s->cv_done.notify_all();
}

static void server_loop(r2p_server* s)
{
    (void)hipSetDevice(s->dev);
    for (;;) {
        bool progressed = false;
        for (Lane& L : s->lanes) {
            if (L.state == 1 && hipEventQuery(L.ev_run) == hipSuccess) { lane_copy(s, L); progressed = true; }
            if (L.state == 2 && hipEventQuery(L.ev_copy) == hipSuccess) { lane_finish(s, L); progressed = true; }
        }
        // admission into free lanes
        for (Lane& L : s->lanes) {
            if (L.state != 0) continue;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                size_t in_off = 0, filt = 0;
                while (!s->waiting.empty() && (int)L.reqs.size() < s->cfg.max_batch_images) {
                    Request* r = s->waiting.front();
                    const size_t in_need = in_off + ((r->in_bytes + 255) & ~(size_t)255);
                    if (in_need > L.in_cap || filt + r->filtered > s->lane_filtered) break;
                    if (s->admitted + r->footprint > s->cfg.hbm_budget_bytes && !L.reqs.empty()) break;
                    if (s->admitted + r->footprint > s->cfg.hbm_budget_bytes && s->admitted > 0) break;
                    s->waiting.pop_front();
                    L.reqs.push_back(r);
                    L.footprint += r->footprint;
                    s->admitted += r->footprint;
                    in_off = in_need;
                    filt += r->filtered;
                }
                if (L.reqs.empty()) continue;
                s->st.batches++;
                s->batch_images += L.reqs.size();
                s->st.peak_admitted_bytes = std::max<uint64_t>(s->st.peak_admitted_bytes, s->admitted);
            }
            lane_launch(s, L);
            progressed = true;
        }
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop) {
                bool busy = false;
                for (Lane& L : s->lanes) busy |= L.state != 0;
                if (!busy && s->waiting.empty()) return;
            }
            if (!progressed) {
                bool busy = false;
                for (Lane& L : s->lanes) busy |= L.state != 0;
                if (busy) {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(50));
                } else {
                    s->cv_work.wait_for(lk, std::chrono::milliseconds(2),
                                        [s] { return s->stop || !s->waiting.empty(); });
                }
            }
        }
    }
}

extern "C" {

int r2p_server_create(const r2p_server_config* cfg, r2p_server** out)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    int rc = r2p_init(dev);
    if (rc) return rc;
    if (!cfg || !out || cfg->hbm_budget_bytes < ((size_t)64 << 20) || cfg->max_batch_images <= 0 ||
        cfg->inflight_batches <= 0 || cfg->inflight_batches > 16) {
        set_error("bad server configuration");
        return R2P_ERR_ARG;
    }
    r2p_server* s = new r2p_server();
    s->cfg = *cfg;
    s->dev = dev;
    fill_defaults(&s->cfg.opt);
    const int nl = cfg->inflight_batches;
    const size_t lane_budget = cfg->hbm_budget_bytes / (size_t)nl;
    // per lane: input arena 1/4, filtered 1/4 of what remains after the input
    // (1/8 with dynamic-Huffman blocks, whose token slots, histograms and
    // larger slice arena take about twice the filtered size more)
    const size_t in_cap = lane_budget / 4;
    s->lane_filtered = std::max<size_t>((lane_budget - in_cap) / (s->cfg.opt.huffman ? 8 : 4), (size_t)16 << 20);
    s->lanes.resize((size_t)nl);
    for (Lane& L : s->lanes) {
        r2p_options o = s->cfg.opt;
        rc = r2p_batch_create(cfg->max_batch_images, in_cap, s->lane_filtered, &o, &L.b);
        if (rc) { r2p_server_destroy(s); return rc; }
        L.in_cap = in_cap;
        if (hipMalloc((void**)&L.d_in, in_cap) != hipSuccess ||
            hipEventCreateWithFlags(&L.ev_run, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&L.ev_copy, hipEventDisableTiming) != hipSuccess) {
            set_error("server lane allocation failed");
            r2p_server_destroy(s);
            return R2P_ERR_HIP;
        }
    }
    s->th = std::thread(server_loop, s);
    *out = s;
    return R2P_OK;
}

int r2p_server_destroy(r2p_server* s)
{
    if (!s) return R2P_OK;
    if (s->th.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->th.join();
    }
    for (Lane& L : s->lanes) {
        if (L.b) r2p_batch_destroy(L.b);
        if (L.d_in) (void)hipFree(L.d_in);
        if (L.ev_run) (void)hipEventDestroy(L.ev_run);
        if (L.ev_copy) (void)hipEventDestroy(L.ev_copy);
    }
    for (auto& kv : s->all) { s->pool.put(kv.second->png, kv.second->png_cap); delete kv.second; }
    delete s;
    return R2P_OK;
}

static int submit_at(r2p_server* s, const r2p_image* img, const void* h_src, SClock::time_point arrival,
                     int64_t* id)
{
    Geometry g;
    int rc = resolve(img, &g);
    if (rc) return rc;
    WindowPlan wp;
    if ((rc = plan_window(img, &wp))) return rc;
    if (!h_src) { set_error("null source"); return R2P_ERR_ARG; }
    auto* r = new Request();
    r->img = *img;
    r->h_src = h_src;
    r->in_bytes = (size_t)g.pitch * (img->height - 1) + g.in_row;
    r->filtered = g.filtered;
    r->footprint = req_footprint(r->in_bytes, r->filtered);
    r->arrival = arrival;
    if (r->filtered > s->lane_filtered || r->in_bytes > s->lanes[0].in_cap || r->footprint > s->cfg.hbm_budget_bytes) {
        delete r;
        set_error("request larger than a serving lane");
        return R2P_ERR_SIZE;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->all[r->id] = r;
        s->waiting.push_back(r);
        s->st.submitted++;
    }
    s->cv_work.notify_one();
    if (id) *id = r->id;
    return R2P_OK;
}

int r2p_server_submit(r2p_server* s, const r2p_image* img, const void* h_src, int64_t* id)
{
    if (!s) { set_error("null server"); return R2P_ERR_ARG; }
    return submit_at(s, img, h_src, SClock::now(), id);
}

int r2p_server_wait(r2p_server* s, int64_t id, int timeout_ms, void* h_png, size_t cap, size_t* len)
{
    if (!s) { set_error("null server"); return R2P_ERR_ARG; }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) { set_error("unknown request %lld", (long long)id); return R2P_ERR_ARG; }
    Request* r = it->second;
    auto pred = [r] { return r->done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) {
        set_error("request %lld not done", (long long)id);
        return R2P_ERR_STATE;
    }
    if (r->status) { set_error("request failed with %d", r->status); return r->status; }
    if (len) *len = r->png_len;
    if (h_png) {
        if (r->png_len > cap) { set_error("png of %zu bytes exceeds %zu", r->png_len, cap); return R2P_ERR_SIZE; }
        memcpy(h_png, r->png, r->png_len);
    }
    s->all.erase(it);
    s->pool.put(r->png, r->png_cap);
    delete r;
    return R2P_OK;
}

int r2p_server_stats_get(r2p_server* s, r2p_server_stats* st)
{
    if (!s || !st) { set_error("bad stats arguments"); return R2P_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->st;
    st->mean_batch_images = s->st.batches ? (double)s->batch_images / (double)s->st.batches : 0.0;
    return R2P_OK;
}

int r2p_server_loadtest(r2p_server* s, int n, const double* arrival_s, const int* kind, const r2p_image* imgs,
                        const void* const* h_src, double* latency_s, uint64_t* png_len, double* wall_s,
                        const void* const* expect_png, const size_t* expect_len, uint8_t* same)
{
    if (!s || n <= 0 || !arrival_s || !kind || !imgs || !h_src || !latency_s) { set_error("bad loadtest arguments"); return R2P_ERR_ARG; }
    if (expect_png && (!expect_len || !same)) { set_error("expected PNGs need their lengths and a result array"); return R2P_ERR_ARG; }
    std::vector<int64_t> ids((size_t)n, -1);
    const auto t0 = SClock::now() + std::chrono::milliseconds(5);
    int rc = R2P_OK;
    int submitted = 0;                  // guarded by s->mu
    bool client_failed = false;
    std::thread client([&] {
        for (int i = 0; i < n; i++) {
            const auto at = t0 + std::chrono::duration_cast<SClock::duration>(std::chrono::duration<double>(arrival_s[i]));
            std::this_thread::sleep_until(at);
            int64_t id = -1;
            const int k = kind[i];
            int e = submit_at(s, &imgs[k], h_src[k], at, &id);
            std::lock_guard<std::mutex> lk(s->mu);
            if (e) { rc = e; client_failed = true; s->cv_done.notify_all(); break; }
            ids[(size_t)i] = id;
            submitted = i + 1;
            s->cv_done.notify_all();
        }
    });
    // Byte checks: each finished PNG is compared with the expected bytes of
    // its image by a few verifier threads, after its latency is taken; they
    // then recycle its buffer.
    std::mutex vmu;
    std::condition_variable vcv;
    std::deque<std::pair<int, Request*>> vq;
    bool vstop = false;
    std::vector<std::thread> verifiers;
    auto release = [s](Request* r) { s->pool.put(r->png, r->png_cap); delete r; };
    if (expect_png) {
        for (int i = 0; i < n; i++) same[i] = 0;
        for (int t = 0; t < 4; t++)
            verifiers.emplace_back([&] {
                for (;;) {
                    std::pair<int, Request*> it;
                    {
                        std::unique_lock<std::mutex> lk(vmu);
                        vcv.wait(lk, [&] { return vstop || !vq.empty(); });
                        if (vq.empty()) return;
                        it = vq.front();
                        vq.pop_front();
                    }
                    Request* r = it.second;
                    const int k = kind[it.first];
                    same[it.first] = r->status == R2P_OK && r->png_len == expect_len[k] &&
                                     (r->png_len == 0 || memcmp(r->png, expect_png[k], r->png_len) == 0);
                    release(r);
                }
            });
    }
    // results are consumed in arrival order while the client keeps submitting
    SClock::time_point last = t0;
    for (int i = 0; i < n; i++) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_done.wait(lk, [&] { return submitted > i || client_failed; });
            if (submitted <= i) break;
            r = s->all[ids[(size_t)i]];
            s->cv_done.wait(lk, [r] { return r->done; });
            latency_s[i] = std::chrono::duration<double>(r->done_at - r->arrival).count();
            if (png_len) png_len[i] = r->png_len;
            if (r->done_at > last) last = r->done_at;
            if (r->status) rc = r->status;
            s->all.erase(ids[(size_t)i]);
        }
        if (expect_png) {
            {
                std::lock_guard<std::mutex> lk(vmu);
                vq.push_back({i, r});
            }
            vcv.notify_one();
        } else {
            release(r);
        }
    }
    client.join();
    {
        std::lock_guard<std::mutex> lk(vmu);
        vstop = true;
    }
    vcv.notify_all();
    for (auto& t : verifiers) t.join();
    if (wall_s) *wall_s = std::chrono::duration<double>(last - t0).count();
    return rc;
}

}  // extern "C"
