// r2p_strip.cpp -- strip-incremental DEFLATE context and single IDAT chunk
// assembly: the device pieces behind the gpu_deflate_backend.h and
// gpu_png_assemble.h facades.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "r2p_common.h"

using namespace r2p;

// A strip context keeps, in device memory, the last 32 KiB of input already
// compressed (the dictionary of the next strip's first segment), the running
// output stream and its byte length.  Each strip: [history | strip] is laid
// out contiguously in one of two window buffers, compressed into segments,
// and the segments are appended to the output by the raw-mode assembler.
struct r2p_deflate {
    hipStream_t stream = nullptr;
    size_t max_strip = 0, max_total = 0;
    int lz77 = 1, seg_bytes = 65536, slice_bytes = 16384, slice_cap_dw = 0, max_seg = 0, max_slice = 0;
    uint8_t* d_win[2] = {nullptr, nullptr};
    int cur = 0;
    size_t hist = 0;                   // history bytes at the start of d_win[cur]
    uint32_t* d_slice_out = nullptr;
    uint32_t* d_slice_bits = nullptr;
    uint32_t* d_slice_adler = nullptr;
    uint32_t* d_slice_stats = nullptr;
    uint32_t* d_slice_boff = nullptr;
    uint32_t* d_seg_tbits = nullptr;
    uint64_t* d_seg_off = nullptr;
    uint8_t* d_out = nullptr;
    uint64_t* d_state = nullptr;       // [0] running bytes, [1] base slot, [2] out_len, [3] adler
    uint64_t* h_state = nullptr;       // pinned
    uint32_t* h_stats = nullptr;       // pinned, 4 per slice
    std::vector<uint32_t> strip_adler;
    std::vector<uint64_t> strip_len;
    int pending_nslice = 0;
    bool pending = false;
    size_t out_bytes = 0;
    r2p_lz_stats stats{};
    double dist_sum = 0;
};

extern "C" {

int r2p_deflate_create(size_t max_strip_bytes, size_t max_total_bytes, int lz77, int segment_bytes,
                       r2p_deflate** out)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    int rc = r2p_init(dev);
    if (rc) return rc;
    if (!out || max_strip_bytes == 0) { set_error("bad deflate arguments"); return R2P_ERR_ARG; }
    r2p_deflate* d = new r2p_deflate();
    d->max_strip = max_strip_bytes;
    d->lz77 = lz77 < 0 ? 1 : std::min(lz77, 2);
    r2p_options o{};
    o.segment_bytes = segment_bytes;
    fill_defaults(&o);
    d->seg_bytes = o.segment_bytes;
    d->slice_bytes = resolve_slice_bytes(o, max_strip_bytes);   // one size for every strip
    d->slice_cap_dw = slice_cap_dwords(d->slice_bytes);
    d->max_seg = (int)((max_strip_bytes + d->seg_bytes - 1) / d->seg_bytes) + 1;
    d->max_slice = (int)((max_strip_bytes + d->slice_bytes - 1) / d->slice_bytes) + 1;   // (strips keep whole slices)
    // output bound: 9/8 of the input plus per-segment framing
    const size_t total_in = std::max(max_total_bytes, max_strip_bytes);
    d->max_total = total_in + total_in / 8 + 16 * (total_in / d->seg_bytes + 2) + 64;
    hipError_t e;
#define TRY(x) do { if ((e = (x)) != hipSuccess) { r2p_deflate_destroy(d); return hip_fail(e, #x); } } while (0)
    TRY(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    for (int i = 0; i < 2; i++) TRY(hipMalloc((void**)&d->d_win[i], kWindow + max_strip_bytes + kTailSlack));
    TRY(hipMalloc((void**)&d->d_slice_out, (size_t)d->max_slice * d->slice_cap_dw * 4));
    TRY(hipMalloc((void**)&d->d_slice_bits, (size_t)d->max_slice * 4));
    TRY(hipMalloc((void**)&d->d_slice_adler, (size_t)d->max_slice * 4));
    TRY(hipMalloc((void**)&d->d_slice_stats, (size_t)d->max_slice * 16));
    TRY(hipMalloc((void**)&d->d_slice_boff, (size_t)d->max_slice * 4));
    TRY(hipMalloc((void**)&d->d_seg_tbits, (size_t)d->max_seg * 4));
    TRY(hipMalloc((void**)&d->d_seg_off, ((size_t)d->max_seg + 1) * 8));
    TRY(hipMalloc((void**)&d->d_out, d->max_total));
    TRY(hipMalloc((void**)&d->d_state, 64));
    TRY(hipHostMalloc((void**)&d->h_state, 64, hipHostMallocDefault));
    TRY(hipHostMalloc((void**)&d->h_stats, (size_t)d->max_slice * 16, hipHostMallocDefault));
#undef TRY
    *out = d;
    return r2p_deflate_reset(d);
}

int r2p_deflate_destroy(r2p_deflate* d)
{
    if (!d) return R2P_OK;
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    for (auto p : d->d_win) (void)hipFree(p);
    (void)hipFree(d->d_slice_out); (void)hipFree(d->d_slice_bits); (void)hipFree(d->d_slice_adler);
    (void)hipFree(d->d_slice_stats); (void)hipFree(d->d_slice_boff); (void)hipFree(d->d_seg_tbits);
    (void)hipFree(d->d_seg_off); (void)hipFree(d->d_out);
    (void)hipFree(d->d_state);
    if (d->h_state) (void)hipHostFree(d->h_state);
    if (d->h_stats) (void)hipHostFree(d->h_stats);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
    return R2P_OK;
}

int r2p_deflate_reset(r2p_deflate* d)
{
    if (!d) { set_error("null deflate context"); return R2P_ERR_ARG; }
    R2P_HIP(hipStreamSynchronize(d->stream));
    R2P_HIP(hipMemsetAsync(d->d_state, 0, 64, d->stream));
    R2P_HIP(hipStreamSynchronize(d->stream));
    d->cur = 0; d->hist = 0; d->out_bytes = 0; d->pending = false; d->pending_nslice = 0;
    d->strip_adler.clear(); d->strip_len.clear();
    d->stats = r2p_lz_stats{};
    d->dist_sum = 0;
    return R2P_OK;
}

void* r2p_deflate_stream(r2p_deflate* d) { return d ? (void*)d->stream : nullptr; }

int r2p_deflate_sync(r2p_deflate* d);

int r2p_deflate_strip(r2p_deflate* d, const void* d_filtered, size_t len, int is_last, int bpp,
                      int row_stride)
{
    if (!d || (!d_filtered && len)) { set_error("bad deflate strip arguments"); return R2P_ERR_ARG; }
    if (len > d->max_strip) { set_error("strip of %zu bytes exceeds %zu", len, d->max_strip); return R2P_ERR_SIZE; }
    if (d->pending) { int rc = r2p_deflate_sync(d); if (rc) return rc; }
    if (len == 0 && !is_last) return R2P_OK;
    uint8_t* win = d->d_win[d->cur];
    hipStream_t s = d->stream;
    const SliceGeom sg = slice_geom((int64_t)len, d->seg_bytes, d->slice_bytes);
    const int nseg = sg.nseg;
    // the strip's worst case (9 bits per byte, every segment's framing) must
    // fit after the output delivered so far: the assembly appends unchecked
    const size_t zmax = len + len / 8 + 16 * (size_t)nseg + 16;
    if (d->out_bytes + zmax > d->max_total) {
        set_error("strip output may reach %zu bytes, the stream buffer holds %zu", d->out_bytes + zmax, d->max_total);
        return R2P_ERR_SIZE;
    }
    if (len) R2P_HIP(hipMemcpyAsync(win + d->hist, d_filtered, len, hipMemcpyDeviceToDevice, s));
    DeflateLaunch DL{};
    DeflateJob& D = DL.single;
    D.data = win; D.first = (int64_t)d->hist; D.len = (int64_t)len; D.dict_lo = 0;
    D.slice_out = d->d_slice_out; D.slice_bits = d->d_slice_bits; D.slice_adler = d->d_slice_adler;
    D.slice_stats = d->d_slice_stats; D.seg_bytes = d->seg_bytes; D.slice_bytes = d->slice_bytes;
    D.slice_cap_dw = d->slice_cap_dw; D.slice_cap_dw2 = d->slice_cap_dw;
    D.final_block = is_last ? 1 : 0; D.bpp = bpp > 0 ? bpp : 1; D.row_stride = row_stride;
    D.lz77 = d->lz77; D.nslice = sg.nslice; D.half_from = sg.half_from;
    DL.dense = d->lz77 == 2;
    DL.jobs = nullptr; DL.njobs = 1; DL.total_slices = (uint32_t)sg.nslice;
    R2P_HIP(launch_deflate(DL, 16, 256, sg.nslice, s));

    AssembleLaunch AL{};
    AssembleJob& A = AL.single;
    A.slice_out = d->d_slice_out; A.slice_bits = d->d_slice_bits; A.slice_adler = d->d_slice_adler;
    A.slice_boff = d->d_slice_boff; A.seg_tbits = d->d_seg_tbits;
    A.seg_off = d->d_seg_off; A.out = d->d_out; A.out_len = d->d_state + 2;
    A.adler_out = (uint32_t*)(d->d_state + 3); A.out_cap = (int64_t)d->max_total;
    A.total_in = (int64_t)len; A.nseg = nseg; A.nslice = sg.nslice; A.spb = sg.spb; A.half_from = sg.half_from;
    A.slice_cap_dw = d->slice_cap_dw; A.slice_cap_dw2 = d->slice_cap_dw; A.slice_bytes = d->slice_bytes;
    A.chunk_bytes = 65536; A.mode = 0; A.final_block = is_last ? 1 : 0;
    A.max_chunks = (int32_t)((zmax + 65535) / 65536);
    A.running = d->d_state; A.base_slot = d->d_state + 1;
    AL.jobs = nullptr; AL.njobs = 1; AL.total_chunks = (uint32_t)A.max_chunks;
    R2P_HIP(launch_scan(AL, 1, s));
    R2P_HIP(launch_assemble(AL, std::min(A.max_chunks, 4096), s));

    // next window: the last min(32 KiB, hist + len) bytes of [history | strip]
    const size_t have = d->hist + len;
    const size_t keep = std::min<size_t>(kWindow, have);
    if (keep) R2P_HIP(hipMemcpyAsync(d->d_win[d->cur ^ 1], win + (have - keep), keep, hipMemcpyDeviceToDevice, s));
    R2P_HIP(hipMemcpyAsync(d->h_state, d->d_state, 32, hipMemcpyDeviceToHost, s));
    R2P_HIP(hipMemcpyAsync(d->h_stats, d->d_slice_stats, (size_t)sg.nslice * 16, hipMemcpyDeviceToHost, s));
    d->cur ^= 1;
    d->hist = keep;
    d->pending = true;
    d->pending_nslice = len ? sg.nslice : 0;
    d->strip_len.push_back(len);
    return R2P_OK;
}

int r2p_deflate_sync(r2p_deflate* d)
{
    if (!d) { set_error("null deflate context"); return R2P_ERR_ARG; }
    R2P_HIP(hipStreamSynchronize(d->stream));
    if (d->pending) {
        d->out_bytes = (size_t)d->h_state[0];
        d->strip_adler.push_back((uint32_t)d->h_state[3]);
        for (int i = 0; i < d->pending_nslice; i++) {
            const uint32_t* s = d->h_stats + 4 * i;
            d->stats.matches += s[0];
            d->stats.matched_bytes += s[1];
            d->dist_sum += s[2];
            d->stats.max_len = std::max(d->stats.max_len, s[3]);
        }
        d->stats.positions += d->strip_len.back();
        d->pending = false;
    }
    return R2P_OK;
}

const void* r2p_deflate_output(const r2p_deflate* d) { return d ? d->d_out : nullptr; }

size_t r2p_deflate_output_bytes(const r2p_deflate* d) { return d ? d->out_bytes : 0; }

size_t r2p_deflate_output_capacity(const r2p_deflate* d) { return d ? d->max_total : 0; }

uint32_t r2p_deflate_adler(const r2p_deflate* d)
{
    if (!d) return 1;
    uint32_t a = 1;
    bool started = false;
    for (size_t i = 0; i < d->strip_adler.size() && i < d->strip_len.size(); i++) {
        if (d->strip_len[i] == 0) continue;
        a = started ? r2p_adler32_combine(a, d->strip_adler[i], d->strip_len[i]) : d->strip_adler[i];
        started = true;
    }
    return a;
}

int r2p_deflate_copy_to_host(const r2p_deflate* d, void* h_dst, size_t n)
{
    if (!d || (!h_dst && n)) { set_error("bad copy arguments"); return R2P_ERR_ARG; }
    if (n > d->max_total) { set_error("copy of %zu bytes exceeds output", n); return R2P_ERR_SIZE; }
    R2P_HIP(hipStreamSynchronize(d->stream));
    if (n) R2P_HIP(hipMemcpy(h_dst, d->d_out, n, hipMemcpyDeviceToHost));
    return R2P_OK;
}

int r2p_deflate_rewind(r2p_deflate* d)
{
    if (!d) { set_error("null deflate context"); return R2P_ERR_ARG; }
    if (d->pending) { int rc = r2p_deflate_sync(d); if (rc) return rc; }
    R2P_HIP(hipMemsetAsync(d->d_state, 0, 8, d->stream));
    d->out_bytes = 0;
    return R2P_OK;
}

int r2p_deflate_lz_stats(const r2p_deflate* d, r2p_lz_stats* s)
{
    if (!d || !s) { set_error("bad stats arguments"); return R2P_ERR_ARG; }
    *s = d->stats;
    s->avg_len = s->matches ? (double)s->matched_bytes / (double)s->matches : 0.0;
    s->avg_dist = s->matches ? d->dist_sum / (double)s->matches : 0.0;
    return R2P_OK;
}

// One IDAT chunk: [len BE]["IDAT"][78 FLG]?[slice][adler BE]?[crc BE].
int r2p_idat_chunk(const void* d_slice, size_t slice_len, int is_first, int is_last, uint32_t adler,
                   int zlib_level, void* d_chunk, size_t chunk_cap, void* d_work, size_t* chunk_len,
                   void* stream)
{
    if ((!d_slice && slice_len) || !d_chunk || !d_work) { set_error("bad idat arguments"); return R2P_ERR_ARG; }
    const size_t data_len = (is_first ? 2u : 0u) + slice_len + (is_last ? 4u : 0u);
    const size_t total = 12 + data_len;
    if (total > chunk_cap) { set_error("chunk of %zu bytes exceeds capacity %zu", total, chunk_cap); return R2P_ERR_SIZE; }
    hipStream_t s = (hipStream_t)stream;
    uint8_t* c = (uint8_t*)d_chunk;
    uint8_t hdr[16];
    hdr[0] = (uint8_t)(data_len >> 24); hdr[1] = (uint8_t)(data_len >> 16);
    hdr[2] = (uint8_t)(data_len >> 8); hdr[3] = (uint8_t)data_len;
    hdr[4] = 'I'; hdr[5] = 'D'; hdr[6] = 'A'; hdr[7] = 'T';
    int nh = 8;
    if (is_first) {
        const int l = zlib_level;
        hdr[8] = 0x78;
        hdr[9] = l == 1 ? 0x01 : (l >= 2 && l <= 5) ? 0x5E : (l >= 6 && l <= 7) ? 0x9C : 0xDA;
        nh = 10;
    }
    R2P_HIP(launch_put_bytes(c, hdr, nh, nullptr, s));
    if (slice_len) R2P_HIP(hipMemcpyAsync(c + nh, d_slice, slice_len, hipMemcpyDeviceToDevice, s));
    if (is_last) {
        uint8_t t[4] = {(uint8_t)(adler >> 24), (uint8_t)(adler >> 16), (uint8_t)(adler >> 8), (uint8_t)adler};
        R2P_HIP(launch_put_bytes(c + nh + slice_len, t, 4, nullptr, s));
    }
    uint32_t* crc = (uint32_t*)d_work;
    R2P_HIP(launch_crc(c + 4, (int64_t)(4 + data_len), crc + 1, crc, s));
    R2P_HIP(launch_put_bytes(c + 8 + data_len, hdr, 0, crc, s));
    if (chunk_len) *chunk_len = total;
    return R2P_OK;
}

}  // extern "C"
