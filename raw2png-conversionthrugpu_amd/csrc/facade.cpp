// facade.cpp -- the reference's C++ API (gpu_filter.h, gpu_adler32.h,
// gpu_crc32.h, gpu_deflate_backend.h, gpu_png_assemble.h, parallel_deflate.h,
// png_writer.h, compression_backend.h, pipeline.h) implemented over the
// r2p_* C-ABI.  Device failures keep the reference's fail-fast behaviour
// (message on stderr, exit(1)); pipeline entry points return false.
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "compression_backend.h"
#include "gpu_adler32.h"
#include "gpu_crc32.h"
#include "gpu_deflate_backend.h"
#include "gpu_filter.h"
#include "gpu_png_assemble.h"
#include "image_source.h"
#include "parallel_deflate.h"
#include "pipeline.h"
#include "png_writer.h"
#include "r2p.h"

namespace {

[[noreturn]] void die(const char* where)
{
    fprintf(stderr, "r2p error in %s: %s\n", where, r2p_last_error());
    exit(1);
}
void chk(int rc, const char* where)
{
    if (rc != R2P_OK) die(where);
}
void hchk(hipError_t e, const char* where)
{
    if (e != hipSuccess) {
        fprintf(stderr, "HIP error in %s: %s\n", where, hipGetErrorString(e));
        exit(1);
    }
}
using Clock = std::chrono::steady_clock;
long long us_since(Clock::time_point t0)
{
    return std::chrono::duration_cast<std::chrono::microseconds>(Clock::now() - t0).count();
}

// r2p_image for a strip of `width` pixels of `bpp` bytes.
r2p_image strip_image(int width, int bpp, int rows, const DicomPixelParams* dicom)
{
    r2p_image im{};
    im.width = (uint32_t)width;
    im.height = (uint32_t)rows;
    if (dicom) {
        const int sb = dicom->bits_allocated / 8;
        im.in_bits = dicom->bits_allocated;
        im.channels = bpp / std::max(1, sb);
        im.pix_mode = R2P_PIX_DICOM;
        static_assert(sizeof(DicomPixelParams) == sizeof(r2p_dicom_params), "layout");
        memcpy(&im.dicom, dicom, sizeof(im.dicom));
    } else if (bpp <= 4) {
        im.in_bits = 8; im.channels = bpp; im.pix_mode = R2P_PIX_READY;
    } else {
        im.in_bits = 16; im.channels = bpp / 2; im.pix_mode = R2P_PIX_READY;
    }
    return im;
}

// ---------------------------------------------------------------------------
// Cross-stream ordering of device buffers handed between contexts.
//
// Every context owns its stream, and the reference hands device buffers from
// one context to another -- the filter output to DEFLATE and Adler-32, the
// DEFLATE output to IDAT assembly -- with no event between the streams
// (pipeline.cpp:618-623, 689-698: the last strip's filter can still be
// running when the consumers start).  Here a producer registers its output
// buffer; after each write it records `written` on its stream.  A consumer
// whose input lies in a registered buffer makes its own stream wait on
// `written`, and after enqueuing its reads records its own event, which the
// producer's next write waits on.  The reference's call order, unchanged,
// is then free of both races (read-before-write and overwrite-while-read).
// ---------------------------------------------------------------------------
struct BufferOrder {
    const uint8_t* base = nullptr;
    size_t size = 0;
    hipEvent_t written = nullptr;
    std::vector<hipEvent_t> readers;   // consumers' events, recorded after their reads
};
std::mutex g_order_mu;
std::vector<BufferOrder*> g_orders;

BufferOrder* order_register(const void* base, size_t size)
{
    auto* o = new BufferOrder();
    o->base = (const uint8_t*)base;
    o->size = size;
    hchk(hipEventCreateWithFlags(&o->written, hipEventDisableTiming), "order_register");
    std::lock_guard<std::mutex> lk(g_order_mu);
    g_orders.push_back(o);
    return o;
}

void order_unregister(BufferOrder* o)
{
    if (!o) return;
    {
        std::lock_guard<std::mutex> lk(g_order_mu);
        g_orders.erase(std::remove(g_orders.begin(), g_orders.end(), o), g_orders.end());
    }
    (void)hipEventDestroy(o->written);
    delete o;
}

// A producer about to write its buffer on stream s: wait for the reads
// enqueued so far.
void order_before_write(BufferOrder* o, hipStream_t s)
{
    std::vector<hipEvent_t> r;
    {
        std::lock_guard<std::mutex> lk(g_order_mu);
        r.swap(o->readers);
    }
    for (hipEvent_t e : r) hchk(hipStreamWaitEvent(s, e, 0), "order_before_write");
}

// A producer's write is enqueued on stream s.
void order_after_write(BufferOrder* o, hipStream_t s)
{
    hchk(hipEventRecord(o->written, s), "order_after_write");
}

// A consumer on stream s about to read [p, p + n): its stream waits for the
// producer's last write.  Returns the buffer's entry (nullptr: not a
// registered buffer, e.g. the caller's own device memory).
BufferOrder* order_before_read(const void* p, size_t n, hipStream_t s)
{
    const uint8_t* q = (const uint8_t*)p;
    BufferOrder* hit = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_order_mu);
        for (BufferOrder* o : g_orders)
            if (q >= o->base && q < o->base + o->size && n <= (size_t)(o->base + o->size - q)) { hit = o; break; }
    }
    if (hit) hchk(hipStreamWaitEvent(s, hit->written, 0), "order_before_read");
    return hit;
}

// The consumer's reads are enqueued on stream s: `mine` marks their end for
// the producer's next write.
void order_after_read(BufferOrder* o, hipEvent_t mine, hipStream_t s)
{
    if (!o) return;
    hchk(hipEventRecord(mine, s), "order_after_read");
    std::lock_guard<std::mutex> lk(g_order_mu);
    if (std::find(o->readers.begin(), o->readers.end(), mine) == o->readers.end()) o->readers.push_back(mine);
}

// A consumer context goes away: its event leaves every reader list.
void order_forget_reader(hipEvent_t mine)
{
    std::lock_guard<std::mutex> lk(g_order_mu);
    for (BufferOrder* o : g_orders)
        o->readers.erase(std::remove(o->readers.begin(), o->readers.end(), mine), o->readers.end());
}

// An event for a consumer context (no timing).
hipEvent_t reader_event(const char* where)
{
    hipEvent_t e = nullptr;
    hchk(hipEventCreateWithFlags(&e, hipEventDisableTiming), where);
    return e;
}

}  // namespace

// ===========================================================================
// gpu_filter.h
// ===========================================================================
struct GpuFilterContext {
    int width = 0, bpp = 0, strip_height = 0, width_bytes = 0;
    hipStream_t stream = nullptr;
    uint8_t* d_input = nullptr;
    uint8_t* d_prior[2] = {nullptr, nullptr};  // carried row (read) / next carry (written)
    int cur = 0;
    uint8_t* d_selected = nullptr;
    uint8_t* h_output = nullptr;
    hipEvent_t ev[4] = {};
    BufferOrder* out_order = nullptr;   // d_selected, read by other contexts' streams
    hipEvent_t ev_read = nullptr;       // end of this context's reads of other contexts' buffers
};

GpuFilterContext* gpu_filter_create(int width, int bpp, int strip_height, bool device_output_only)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    chk(r2p_init(dev), "gpu_filter_create");
    auto* c = new GpuFilterContext();
    c->width = width; c->bpp = bpp; c->strip_height = strip_height; c->width_bytes = width * bpp;
    const size_t in_b = (size_t)strip_height * c->width_bytes;
    const size_t out_b = (size_t)strip_height * (c->width_bytes + 1);
    hchk(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "gpu_filter_create");
    hchk(hipMalloc((void**)&c->d_input, in_b + 16), "gpu_filter_create");
    for (auto& p : c->d_prior) hchk(hipMalloc((void**)&p, c->width_bytes + 16), "gpu_filter_create");
    hchk(hipMalloc((void**)&c->d_selected, out_b + 16), "gpu_filter_create");
    if (!device_output_only) hchk(hipHostMalloc((void**)&c->h_output, out_b + 16, hipHostMallocDefault), "gpu_filter_create");
    for (auto& e : c->ev) hchk(hipEventCreate(&e), "gpu_filter_create");
    c->out_order = order_register(c->d_selected, out_b + 16);
    c->ev_read = reader_event("gpu_filter_create");
    gpu_filter_reset(c);
    hchk(hipStreamSynchronize(c->stream), "gpu_filter_create");
    return c;
}

void gpu_filter_destroy(GpuFilterContext* c)
{
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    order_unregister(c->out_order);
    order_forget_reader(c->ev_read);
    (void)hipEventDestroy(c->ev_read);
    (void)hipFree(c->d_input); (void)hipFree(c->d_prior[0]); (void)hipFree(c->d_prior[1]);
    (void)hipFree(c->d_selected);
    if (c->h_output) (void)hipHostFree(c->h_output);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

uint8_t* gpu_filter_alloc_pinned(size_t bytes)
{
    uint8_t* p = nullptr;
    hchk(hipHostMalloc((void**)&p, bytes, hipHostMallocDefault), "gpu_filter_alloc_pinned");
    return p;
}

void gpu_filter_free_pinned(uint8_t* p)
{
    if (p) (void)hipHostFree(p);
}

void gpu_filter_reset(GpuFilterContext* c)
{
    if (!c) return;
    hchk(hipMemsetAsync(c->d_prior[c->cur], 0, c->width_bytes, c->stream), "gpu_filter_reset");
}

static const uint8_t* filter_run(GpuFilterContext* c, int rows, GpuTimings* t, const DicomPixelParams* dicom,
                                 bool want_host)
{
    r2p_image im = strip_image(c->width, c->bpp, rows, dicom);
    order_before_write(c->out_order, c->stream);        // earlier strips' readers are done
    hchk(hipEventRecord(c->ev[1], c->stream), "filter");
    chk(r2p_filter_rows(&im, c->d_input, rows, c->d_prior[c->cur], c->d_selected, c->d_prior[c->cur ^ 1], 0,
                        c->stream), "gpu_filter_process");
    c->cur ^= 1;
    hchk(hipEventRecord(c->ev[2], c->stream), "filter");
    order_after_write(c->out_order, c->stream);         // consumers on other streams wait for this
    const size_t out_b = gpu_filter_output_size(c, rows);
    if (want_host) hchk(hipMemcpyAsync(c->h_output, c->d_selected, out_b, hipMemcpyDeviceToHost, c->stream), "filter");
    hchk(hipEventRecord(c->ev[3], c->stream), "filter");
    if (want_host || t) hchk(hipEventSynchronize(c->ev[3]), "filter");
    if (t) {
        (void)hipEventElapsedTime(&t->h2d_ms, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&t->kernel_ms, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&t->d2h_ms, c->ev[2], c->ev[3]);
    }
    return want_host ? c->h_output : nullptr;
}

const uint8_t* gpu_filter_process_from_device(GpuFilterContext* c, const uint8_t* d_input,
                                              const uint8_t* d_prior_row, int rows, GpuTimings* t,
                                              const DicomPixelParams* dicom)
{
    if (rows <= 0 || rows > c->strip_height) die("gpu_filter_process_from_device: rows");
    hchk(hipEventRecord(c->ev[0], c->stream), "filter");
    BufferOrder* src = order_before_read(d_input, (size_t)rows * c->width_bytes, c->stream);
    hchk(hipMemcpyAsync(c->d_input, d_input, (size_t)rows * c->width_bytes, hipMemcpyDeviceToDevice, c->stream), "filter");
    order_after_read(src, c->ev_read, c->stream);
    if (d_prior_row) {
        BufferOrder* pr = order_before_read(d_prior_row, c->width_bytes, c->stream);
        hchk(hipMemcpyAsync(c->d_prior[c->cur], d_prior_row, c->width_bytes, hipMemcpyDeviceToDevice, c->stream), "filter");
        order_after_read(pr, c->ev_read, c->stream);
    }
    return filter_run(c, rows, t, dicom, c->h_output != nullptr);
}

const uint8_t* gpu_filter_process_from_host(GpuFilterContext* c, const uint8_t* h_input,
                                            const uint8_t* h_prior_row, int rows, GpuTimings* t,
                                            const DicomPixelParams* dicom, bool device_output_only)
{
    if (rows <= 0 || rows > c->strip_height) die("gpu_filter_process_from_host: rows");
    hchk(hipEventRecord(c->ev[0], c->stream), "filter");
    hchk(hipMemcpyAsync(c->d_input, h_input, (size_t)rows * c->width_bytes, hipMemcpyHostToDevice, c->stream), "filter");
    if (h_prior_row)
        hchk(hipMemcpyAsync(c->d_prior[c->cur], h_prior_row, c->width_bytes, hipMemcpyHostToDevice, c->stream), "filter");
    return filter_run(c, rows, t, dicom, !device_output_only && c->h_output);
}

size_t gpu_filter_output_size(const GpuFilterContext* c, int rows)
{
    return (size_t)rows * (c->width_bytes + 1);
}

const uint8_t* gpu_filter_device_output(const GpuFilterContext* c) { return c->d_selected; }

void gpu_filter_copy_prior_row_to_host(const GpuFilterContext* c, uint8_t* h_dst)
{
    hchk(hipMemcpyAsync(h_dst, c->d_prior[c->cur], c->width_bytes, hipMemcpyDeviceToHost, c->stream), "copy_prior_row");
    hchk(hipStreamSynchronize(c->stream), "copy_prior_row");
}

void* gpu_filter_stream(const GpuFilterContext* c) { return c ? (void*)c->stream : nullptr; }

// ===========================================================================
// gpu_adler32.h / gpu_crc32.h
// ===========================================================================
struct GpuChecksumCtx {
    size_t max_in = 0;
    hipStream_t stream = nullptr;
    uint8_t* d_work = nullptr;
    uint32_t* d_res = nullptr;
    uint32_t* h_res = nullptr;
    bool pending = false;
    hipEvent_t ev_read = nullptr;
};
struct GpuAdler32Context : GpuChecksumCtx {};
struct GpuCrc32Context : GpuChecksumCtx {};

static void cks_init(GpuChecksumCtx* c, size_t max_in)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    chk(r2p_init(dev), "checksum_create");
    c->max_in = max_in;
    hchk(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "checksum_create");
    hchk(hipMalloc((void**)&c->d_work, r2p_checksum_work_bytes(max_in)), "checksum_create");
    hchk(hipMalloc((void**)&c->d_res, 16), "checksum_create");
    hchk(hipHostMalloc((void**)&c->h_res, 16, hipHostMallocDefault), "checksum_create");
    c->ev_read = reader_event("checksum_create");
}
static void cks_free(GpuChecksumCtx* c)
{
    (void)hipStreamSynchronize(c->stream);
    order_forget_reader(c->ev_read);
    (void)hipEventDestroy(c->ev_read);
    (void)hipFree(c->d_work); (void)hipFree(c->d_res); (void)hipHostFree(c->h_res);
    (void)hipStreamDestroy(c->stream);
}

GpuAdler32Context* gpu_adler32_create(size_t max_input_bytes, size_t)
{
    auto* c = new GpuAdler32Context();
    cks_init(c, max_input_bytes);
    return c;
}
void gpu_adler32_destroy(GpuAdler32Context* c)
{
    if (!c) return;
    cks_free(c);
    delete c;
}
void gpu_adler32_launch(GpuAdler32Context* c, const uint8_t* d_data, size_t len)
{
    if (len > c->max_in) {
        fprintf(stderr, "gpu_adler32_launch: len %zu exceeds max_input_bytes %zu\n", len, c->max_in);
        exit(1);
    }
    BufferOrder* src = order_before_read(d_data, len, c->stream);
    chk(r2p_adler32_device(d_data, len, c->d_work, c->d_res, c->stream), "gpu_adler32_launch");
    order_after_read(src, c->ev_read, c->stream);
    hchk(hipMemcpyAsync(c->h_res, c->d_res, 4, hipMemcpyDeviceToHost, c->stream), "gpu_adler32_launch");
    c->pending = true;
}
uint32_t gpu_adler32_collect(GpuAdler32Context* c)
{
    if (!c->pending) return 1u;
    hchk(hipStreamSynchronize(c->stream), "gpu_adler32_collect");
    c->pending = false;
    return c->h_res[0];
}
uint32_t gpu_adler32_compute(GpuAdler32Context* c, const uint8_t* d, size_t len)
{
    gpu_adler32_launch(c, d, len);
    return gpu_adler32_collect(c);
}

GpuCrc32Context* gpu_crc32_create(size_t max_input_bytes, size_t)
{
    auto* c = new GpuCrc32Context();
    cks_init(c, max_input_bytes);
    return c;
}
void gpu_crc32_destroy(GpuCrc32Context* c)
{
    if (!c) return;
    cks_free(c);
    delete c;
}
uint32_t gpu_crc32_compute(GpuCrc32Context* c, const uint8_t* d, size_t len)
{
    if (len > c->max_in) {
        fprintf(stderr, "gpu_crc32_compute: len %zu exceeds max_input_bytes %zu\n", len, c->max_in);
        exit(1);
    }
    BufferOrder* src = order_before_read(d, len, c->stream);
    chk(r2p_crc32_device(d, len, c->d_work, c->d_res, c->stream), "gpu_crc32_compute");
    order_after_read(src, c->ev_read, c->stream);
    hchk(hipMemcpyAsync(c->h_res, c->d_res, 4, hipMemcpyDeviceToHost, c->stream), "gpu_crc32_compute");
    hchk(hipStreamSynchronize(c->stream), "gpu_crc32_compute");
    return c->h_res[0];
}
uint32_t gpu_crc32_combine(uint32_t c1, uint32_t c2, size_t len2) { return r2p_crc32_combine(c1, c2, len2); }

// ===========================================================================
// gpu_deflate_backend.h
// ===========================================================================
struct GpuDeflateContext {
    r2p_deflate* d = nullptr;
    GpuDeflatePhaseStats ph;
    hipEvent_t ev[2] = {};
    bool timed = false;
    BufferOrder* out_order = nullptr;   // the output stream's buffer, read by IDAT assembly
    hipEvent_t ev_read = nullptr;       // end of the reads of the filter's output
};

GpuDeflateContext* gpu_deflate_create(int max_rows, int max_row_bytes, size_t max_total_bits, bool use_lz77)
{
    auto* c = new GpuDeflateContext();
    const size_t strip = (size_t)max_rows * (size_t)max_row_bytes;
    const size_t total_in = max_total_bits / 8 + 1;
    chk(r2p_deflate_create(strip, std::max(total_in, strip), use_lz77 ? 1 : 0, 65536, &c->d), "gpu_deflate_create");
    for (auto& e : c->ev) hchk(hipEventCreate(&e), "gpu_deflate_create");
    c->out_order = order_register(r2p_deflate_output(c->d), r2p_deflate_output_capacity(c->d));
    c->ev_read = reader_event("gpu_deflate_create");
    return c;
}
void gpu_deflate_destroy(GpuDeflateContext* c)
{
    if (!c) return;
    (void)r2p_deflate_sync(c->d);
    order_unregister(c->out_order);
    order_forget_reader(c->ev_read);
    (void)hipEventDestroy(c->ev_read);
    r2p_deflate_destroy(c->d);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    delete c;
}
void gpu_deflate_reset(GpuDeflateContext* c)
{
    chk(r2p_deflate_reset(c->d), "gpu_deflate_reset");
    c->ph = GpuDeflatePhaseStats{};
}
void gpu_deflate_compress_strip(GpuDeflateContext* c, const uint8_t* d_filtered, int rows, int row_bytes,
                                bool is_last)
{
    hipStream_t s = (hipStream_t)r2p_deflate_stream(c->d);
    const size_t n = (size_t)rows * (size_t)row_bytes;
    BufferOrder* src = order_before_read(d_filtered, n, s);   // e.g. the filter's last strip
    order_before_write(c->out_order, s);                      // assembly reads of earlier output
    hchk(hipEventRecord(c->ev[0], s), "compress_strip");
    chk(r2p_deflate_strip(c->d, d_filtered, n, is_last ? 1 : 0, 1, row_bytes), "gpu_deflate_compress_strip");
    hchk(hipEventRecord(c->ev[1], s), "compress_strip");
    order_after_read(src, c->ev_read, s);
    order_after_write(c->out_order, s);
    c->timed = true;
}
void gpu_deflate_stream_sync(GpuDeflateContext* c)
{
    auto t0 = Clock::now();
    chk(r2p_deflate_sync(c->d), "gpu_deflate_stream_sync");
    c->ph.sync_us += us_since(t0);
    if (c->timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) c->ph.encode_us += (long long)(ms * 1000.f);
        c->timed = false;
    }
}
const uint8_t* gpu_deflate_output(const GpuDeflateContext* c) { return (const uint8_t*)r2p_deflate_output(c->d); }
size_t gpu_deflate_output_byte_length(const GpuDeflateContext* c) { return r2p_deflate_output_bytes(c->d); }
size_t gpu_deflate_flushable_byte_length(const GpuDeflateContext* c) { return r2p_deflate_output_bytes(c->d); }
void gpu_deflate_copy_to_host(const GpuDeflateContext* c, uint8_t* h, size_t n)
{
    chk(r2p_deflate_copy_to_host(c->d, h, n), "gpu_deflate_copy_to_host");
}
LzStats gpu_deflate_lz_stats(const GpuDeflateContext* c)
{
    r2p_lz_stats s{};
    chk(r2p_deflate_lz_stats(c->d, &s), "gpu_deflate_lz_stats");
    LzStats o{};
    o.matches_found = s.matches; o.matched_bytes = s.matched_bytes; o.max_match_len = s.max_len;
    o.avg_match_len = s.avg_len; o.avg_match_dist = s.avg_dist; o.total_positions = s.positions;
    o.coverage = s.positions ? (double)s.matched_bytes / (double)s.positions : 0.0;
    return o;
}
GpuDeflatePhaseStats gpu_deflate_get_phase_stats(const GpuDeflateContext* c) { return c->ph; }
void gpu_deflate_reset_phase_stats(GpuDeflateContext* c) { c->ph = GpuDeflatePhaseStats{}; }

// ===========================================================================
// gpu_png_assemble.h
// ===========================================================================
struct GpuPngAssembleContext {
    size_t cap = 0;
    hipStream_t stream = nullptr;
    uint8_t* d_buf = nullptr;
    uint8_t* d_work = nullptr;
    GpuPngAssembleStats st;
    hipEvent_t ev_read = nullptr;
};

GpuPngAssembleContext* gpu_png_assemble_create(size_t max_chunk_data_bytes)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    chk(r2p_init(dev), "gpu_png_assemble_create");
    auto* c = new GpuPngAssembleContext();
    c->cap = max_chunk_data_bytes + 12 + 8;
    hchk(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "gpu_png_assemble_create");
    hchk(hipMalloc((void**)&c->d_buf, c->cap), "gpu_png_assemble_create");
    hchk(hipMalloc((void**)&c->d_work, r2p_checksum_work_bytes(c->cap)), "gpu_png_assemble_create");
    c->ev_read = reader_event("gpu_png_assemble_create");
    return c;
}
void gpu_png_assemble_destroy(GpuPngAssembleContext* c)
{
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    order_forget_reader(c->ev_read);
    (void)hipEventDestroy(c->ev_read);
    (void)hipFree(c->d_buf); (void)hipFree(c->d_work); (void)hipStreamDestroy(c->stream);
    delete c;
}
size_t gpu_png_assemble_idat_chunk(GpuPngAssembleContext* c, const uint8_t* d_slice, size_t n, bool is_first,
                                   bool is_last, uint32_t adler, int level)
{
    auto t0 = Clock::now();
    size_t len = 0;
    BufferOrder* src = order_before_read(d_slice, n, c->stream);
    chk(r2p_idat_chunk(d_slice, n, is_first, is_last, adler, level, c->d_buf, c->cap, c->d_work, &len, c->stream),
        "gpu_png_assemble_idat_chunk");
    order_after_read(src, c->ev_read, c->stream);
    c->st.chunks++;
    c->st.data_d2d_us += us_since(t0);
    return len;
}
void gpu_png_assemble_copy_to_host(const GpuPngAssembleContext* c, uint8_t* h, size_t n)
{
    auto t0 = Clock::now();
    hchk(hipMemcpyAsync(h, c->d_buf, n, hipMemcpyDeviceToHost, c->stream), "gpu_png_assemble_copy_to_host");
    hchk(hipStreamSynchronize(c->stream), "gpu_png_assemble_copy_to_host");
    const_cast<GpuPngAssembleContext*>(c)->st.copy_d2h_us += us_since(t0);
}
GpuPngAssembleStats gpu_png_assemble_get_stats(const GpuPngAssembleContext* c) { return c->st; }
void gpu_png_assemble_reset_stats(GpuPngAssembleContext* c) { c->st = GpuPngAssembleStats{}; }

// ===========================================================================
// parallel_deflate.h -- CPU backend (chunked zlib, one reusable z_stream per
// worker thread)
// ===========================================================================
namespace {
struct TlsStream {
    z_stream zs{};
    int level = -1;
    bool inited = false;
    ~TlsStream()
    {
        if (inited) deflateEnd(&zs);
    }
    z_stream& get(int lvl)
    {
        if (inited && level == lvl) {
            deflateReset(&zs);
        } else {
            if (inited) deflateEnd(&zs);
            zs = z_stream{};
            deflateInit2(&zs, lvl, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY);
            level = lvl;
            inited = true;
        }
        return zs;
    }
};
thread_local TlsStream tls_zs;

struct ChunkOut {
    std::vector<uint8_t> bytes;
    unsigned long adler = 1;
    size_t n = 0;
    long long init_us = 0, comp_us = 0;
};
}  // namespace

DeflateResult deflate_strip(const uint8_t* input, size_t n, const ParallelDeflateConfig& cfg, ThreadPool& pool,
                            bool is_last)
{
    DeflateResult r;
    r.input_size = n;
    r.strip_adler = 1;
    if (n == 0) return r;
    int nt = std::max(1, cfg.num_threads);
    if ((size_t)nt > n / 4096) nt = (int)std::max<size_t>(1, n / 4096);
    const size_t chunk = (n + nt - 1) / nt;
    std::vector<std::future<ChunkOut>> fut;
    for (int i = 0; i < nt; i++) {
        const size_t a = (size_t)i * chunk;
        if (a >= n) break;
        const size_t b = std::min(n, a + chunk);
        const bool fin = is_last && b == n;
        const int lvl = cfg.zlib_level;
        fut.push_back(pool.submit([=]() {
            ChunkOut c;
            c.n = b - a;
            auto t0 = Clock::now();
            z_stream& zs = tls_zs.get(lvl);
            c.init_us = us_since(t0);
            c.bytes.resize(deflateBound(&zs, (uLong)(b - a)) + 64);
            zs.next_in = const_cast<Bytef*>(input + a);
            zs.avail_in = (uInt)(b - a);
            zs.next_out = c.bytes.data();
            zs.avail_out = (uInt)c.bytes.size();
            auto t1 = Clock::now();
            deflate(&zs, fin ? Z_FINISH : Z_FULL_FLUSH);
            c.comp_us = us_since(t1);
            c.bytes.resize(zs.total_out);
            c.adler = adler32(1L, input + a, (uInt)(b - a));
            return c;
        }));
    }
    bool started = false;
    for (auto& f : fut) {
        ChunkOut c = f.get();
        r.strip_adler = started ? adler32_combine(r.strip_adler, c.adler, (z_off_t)c.n) : c.adler;
        started = true;
        r.init_us += c.init_us;
        r.compress_us += c.comp_us;
        r.data.insert(r.data.end(), c.bytes.begin(), c.bytes.end());
    }
    return r;
}

void bench_deflate(size_t strip_bytes, int num_threads)
{
    constexpr int kWarm = 3, kIters = 15;
    std::vector<uint8_t> strip(strip_bytes);
    for (size_t i = 0; i < strip_bytes; i++) strip[i] = (uint8_t)(((uint32_t)i * 0x9E3779B9u) >> 24);
    ThreadPool pool(std::max(1, num_threads));
    const double mb = (double)strip_bytes / (1024.0 * 1024.0);
    printf("\nDeflate benchmark (zlib %s): %zu-byte strip (%.2f MB), %d chunks per strip, %d timed (+%d warm-up)\n",
           zlibVersion(), strip_bytes, mb, num_threads, kIters, kWarm);
    printf("  level   ms/strip   MB/s in   out/in   set-up sum\n");
    for (int level : {0, 1, 3}) {
        ParallelDeflateConfig cfg;
        cfg.num_threads = num_threads;
        cfg.zlib_level = level;
        for (int w = 0; w < kWarm; w++) (void)deflate_strip(strip.data(), strip_bytes, cfg, pool, false);
        double ms = 0, init = 0;
        size_t out = 0;
        for (int it = 0; it < kIters; it++) {
            const auto t0 = Clock::now();
            const DeflateResult r = deflate_strip(strip.data(), strip_bytes, cfg, pool, false);
            ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
            init += (double)r.init_us;
            out = r.data.size();
        }
        ms /= kIters;
        printf("  %5d   %8.2f   %7.1f   %6.4f   %8.0f us\n", level, ms, mb / (ms * 1e-3),
               strip_bytes ? (double)out / (double)strip_bytes : 0.0, init / kIters);
    }
    printf("  (set-up sum: z_stream acquire/reset time summed over the parallel chunks)\n");
}

void accum_adler(ParallelDeflateState& st, const DeflateResult& dr)
{
    if (dr.input_size == 0) return;
    st.running_adler = st.started ? adler32_combine(st.running_adler, dr.strip_adler, (z_off_t)dr.input_size)
                                  : dr.strip_adler;
    st.started = true;
}

void zlib_header(int level, uint8_t out[2])
{
    out[0] = 0x78;  // deflate, 32 KiB window
    out[1] = level == 1 ? 0x01 : (level >= 2 && level <= 5) ? 0x5E : (level >= 6 && level <= 7) ? 0x9C : 0xDA;
}

void zlib_trailer(const ParallelDeflateState& st, uint8_t out[4])
{
    for (int i = 0; i < 4; i++) out[i] = (uint8_t)(st.running_adler >> (24 - 8 * i));
}

unsigned long adler32_of(const uint8_t* d, size_t n) { return adler32(1L, d, (uInt)n); }
unsigned long crc32_of(const uint8_t* d, size_t n) { return crc32(0L, d, (uInt)n); }

// ===========================================================================
// png_writer.h
// ===========================================================================
static void put_u32(FILE* f, uint32_t v)
{
    const uint8_t b[4] = {(uint8_t)(v >> 24), (uint8_t)(v >> 16), (uint8_t)(v >> 8), (uint8_t)v};
    fwrite(b, 1, 4, f);
}

PngWriter::~PngWriter() { close(); }

bool PngWriter::open(const char* path, uint32_t w, uint32_t h, uint8_t depth, PngColorType ct)
{
    file_ = fopen(path, "wb");
    if (!file_) return false;
    setvbuf(file_, nullptr, _IOFBF, 512 * 1024);
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    fwrite(sig, 1, 8, file_);
    uint8_t ih[13];
    for (int i = 0; i < 4; i++) { ih[i] = (uint8_t)(w >> (24 - 8 * i)); ih[4 + i] = (uint8_t)(h >> (24 - 8 * i)); }
    ih[8] = depth; ih[9] = (uint8_t)ct; ih[10] = ih[11] = ih[12] = 0;
    write_chunk("IHDR", ih, 13);
    idat_pos_ = 0;
    return true;
}
void PngWriter::begin_idat() { idat_pos_ = 0; }
void PngWriter::write_idat_bytes(const uint8_t* d, size_t n)
{
    while (n) {
        const size_t take = std::min(n, sizeof(idat_buf_) - idat_pos_);
        memcpy(idat_buf_ + idat_pos_, d, take);
        idat_pos_ += take; d += take; n -= take;
        if (idat_pos_ == sizeof(idat_buf_)) flush_idat_buf();
    }
}
void PngWriter::end_idat()
{
    if (idat_pos_) flush_idat_buf();
}
void PngWriter::close()
{
    if (!file_) return;
    write_chunk("IEND", nullptr, 0);
    fclose(file_);
    file_ = nullptr;
}
void PngWriter::flush_idat_buf()
{
    if (!idat_pos_) return;
    write_chunk("IDAT", idat_buf_, (uint32_t)idat_pos_);
    idat_pos_ = 0;
}
void PngWriter::write_chunk(const char type[4], const uint8_t* d, uint32_t len)
{
    put_u32(file_, len);
    fwrite(type, 1, 4, file_);
    if (len) fwrite(d, 1, len, file_);
    uLong c = crc32(0L, (const Bytef*)type, 4);
    if (len) c = crc32(c, d, len);
    put_u32(file_, (uint32_t)c);
}

// ===========================================================================
// compression_backend.h -- GPU backend and factory
// ===========================================================================
GpuDeflateBackend::GpuDeflateBackend(size_t max_strip_bytes, bool lz77)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return;
    if (r2p_deflate_create(max_strip_bytes, max_strip_bytes, lz77 ? 1 : 0, 65536, &ctx_) != R2P_OK) {
        ctx_ = nullptr;
        return;
    }
    cap_ = max_strip_bytes;
    if (hipMalloc((void**)&d_in_, cap_ + 16) != hipSuccess) {
        r2p_deflate_destroy(ctx_);
        ctx_ = nullptr;
    }
}
GpuDeflateBackend::~GpuDeflateBackend()
{
    if (ctx_) r2p_deflate_destroy(ctx_);
    if (d_in_) (void)hipFree(d_in_);
}
DeflateResult GpuDeflateBackend::compress_strip(const uint8_t* input, size_t n, bool is_last)
{
    DeflateResult r;
    r.input_size = n;
    r.strip_adler = n ? adler32(1L, input, (uInt)n) : 1;
    if (!ctx_) die("GpuDeflateBackend: no device");
    if (n > cap_) {
        fprintf(stderr, "GpuDeflateBackend: strip of %zu bytes exceeds %zu\n", n, cap_);
        exit(1);
    }
    hipStream_t s = (hipStream_t)r2p_deflate_stream(ctx_);
    auto t0 = Clock::now();
    if (n) hchk(hipMemcpyAsync(d_in_, input, n, hipMemcpyHostToDevice, s), "GpuDeflateBackend");
    chk(r2p_deflate_strip(ctx_, d_in_, n, is_last ? 1 : 0, 1, 0), "GpuDeflateBackend");
    chk(r2p_deflate_sync(ctx_), "GpuDeflateBackend");
    const size_t total = r2p_deflate_output_bytes(ctx_);
    r.data.resize(total);
    if (total)
        hchk(hipMemcpy(r.data.data(), r2p_deflate_output(ctx_), total, hipMemcpyDeviceToHost), "GpuDeflateBackend");
    chk(is_last ? r2p_deflate_reset(ctx_) : r2p_deflate_rewind(ctx_), "GpuDeflateBackend");
    r.compress_us = us_since(t0);
    return r;
}

std::unique_ptr<CompressionBackend> make_compression_backend(const ParallelDeflateConfig& cfg, int num_threads)
{
    const char* force_cpu = getenv("R2P_CPU_DEFLATE");
    int n = 0;
    if (!(force_cpu && *force_cpu == '1') && hipGetDeviceCount(&n) == hipSuccess && n > 0) {
        auto g = std::make_unique<GpuDeflateBackend>((size_t)64 << 20, true);
        if (g->ok()) return g;
    }
    return std::make_unique<CPUZlibBackend>(cfg, num_threads);
}
