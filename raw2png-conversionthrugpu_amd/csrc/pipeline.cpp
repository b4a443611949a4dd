// pipeline.cpp -- pipeline.h entry points and the native input readers.
//
// encode_source_to_png (pipeline.cpp:2058-2062 of the reference) has two
// forms, chosen like run_pipeline (:1571-1667):
//   use_gpu_deflate: the whole image goes to HBM once and every stage runs on
//     the GPU (r2p_encode_device); the PNG comes back in one copy.
//   otherwise: strips are windowed+filtered on the GPU (gpu_filter.h) and
//     deflated by the CPU zlib backend into 32 KiB IDAT chunks, which is the
//     reference's default path (run_one, :298-443).
// Output stage (use_gpu_deflate): the PNG leaves the device into a pooled
// pinned host buffer and a writer thread writes the file, so the frames of
// a multi-frame DICOM encode while the previous ones are written (the
// reference's writer stage, pipeline.cpp:830-853); single-file calls wait
// for their file.
// Readers: uncompressed little-endian TIFF (image_loader.cpp:59-214 for the
// strip/byteswap/MINISWHITE semantics) and uncompressed little-endian DICOM
// Part 10 (dicom_loader.cpp:293-451 for the DicomPixelParams derivation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <filesystem>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "compression_backend.h"
#include "gpu_filter.h"
#include "image_source.h"
#include "parallel_deflate.h"
#include "pipeline.h"
#include "png_writer.h"
#include "r2p.h"

namespace {

using Clock = std::chrono::steady_clock;

PngColorType color_type(int ch)
{
    switch (ch) {
    case 1: return PngColorType::Gray;
    case 2: return PngColorType::GrayA;
    case 4: return PngColorType::RGBA;
    default: return PngColorType::RGB;
    }
}

bool read_file(const char* path, std::vector<uint8_t>& out)
{
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (n < 0) { fclose(f); return false; }
    out.resize((size_t)n);
    const size_t got = n ? fread(out.data(), 1, (size_t)n, f) : 0;
    fclose(f);
    return got == (size_t)n;
}

// In-memory source over a contiguous pixel buffer.
struct BufferSource : ImageSource {
    ImageInfo inf;
    std::vector<uint8_t> px;
    bool has_dicom = false;
    DicomPixelParams dp;
    int next = 0;
    const ImageInfo& info() const override { return inf; }
    int read_strip(uint8_t* out, int max_rows) override
    {
        const int left = (int)inf.height - next;
        if (left <= 0) return 0;
        const int n = std::min(left, max_rows);
        const size_t rb = (size_t)inf.width * inf.bpp;
        memcpy(out, px.data() + (size_t)next * rb, (size_t)n * rb);
        next += n;
        return n;
    }
    const DicomPixelParams* dicom_params() const override { return has_dicom ? &dp : nullptr; }
};

// ---------------------------------------------------------------------------
// TIFF (baseline, uncompressed, little-endian, chunky, strips)
// ---------------------------------------------------------------------------
uint32_t rd16(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
uint32_t rd32(const uint8_t* p) { return rd16(p) | (rd16(p + 2) << 16); }

bool load_tiff(const char* path, BufferSource& s)
{
    std::vector<uint8_t> f;
    if (!read_file(path, f) || f.size() < 8) { fprintf(stderr, "TIFF: cannot read %s\n", path); return false; }
    if (!(f[0] == 'I' && f[1] == 'I' && rd16(&f[2]) == 42)) { fprintf(stderr, "TIFF: only little-endian TIFF is read natively\n"); return false; }
    const uint32_t ifd = rd32(&f[4]);
    if (ifd + 2 > f.size()) return false;
    const uint32_t ne = rd16(&f[ifd]);
    uint32_t w = 0, h = 0, bps = 8, spp = 1, comp = 1, photo = 1, planar = 1, rps = 0xFFFFFFFFu;
    std::vector<uint32_t> offs, cnts;
    auto vals = [&](const uint8_t* e, std::vector<uint32_t>& v) {
        const uint32_t type = rd16(e + 2), cnt = rd32(e + 4);
        const size_t sz = type == 3 ? 2 : 4;
        const uint8_t* p = cnt * sz <= 4 ? e + 8 : f.data() + rd32(e + 8);
        if (p < f.data() || p + cnt * sz > f.data() + f.size()) return false;
        v.resize(cnt);
        for (uint32_t i = 0; i < cnt; i++) v[i] = type == 3 ? rd16(p + 2 * i) : rd32(p + 4 * i);
        return true;
    };
    for (uint32_t i = 0; i < ne; i++) {
        const size_t o = ifd + 2 + 12 * (size_t)i;
        if (o + 12 > f.size()) return false;
        const uint8_t* e = &f[o];
        std::vector<uint32_t> v;
        if (!vals(e, v) || v.empty()) continue;
        switch (rd16(e)) {
        case 256: w = v[0]; break;
        case 257: h = v[0]; break;
        case 258: bps = v[0]; break;
        case 259: comp = v[0]; break;
        case 262: photo = v[0]; break;
        case 273: offs = v; break;
        case 277: spp = v[0]; break;
        case 278: rps = v[0]; break;
        case 279: cnts = v; break;
        case 284: planar = v[0]; break;
        default: break;
        }
    }
    if (comp != 1 || planar != 1 || (bps != 8 && bps != 16) || spp < 1 || spp > 4 || !w || !h) {
        fprintf(stderr, "TIFF: unsupported layout (compression %u, planar %u, bits %u, spp %u)\n", comp, planar, bps, spp);
        return false;
    }
    s.inf.width = w; s.inf.height = h; s.inf.channels = (int)spp; s.inf.bits_per_sample = (int)bps;
    s.inf.bpp = (int)(spp * bps / 8);
    const size_t rb = (size_t)w * s.inf.bpp, total = rb * h;
    s.px.resize(total);
    size_t got = 0;
    for (size_t k = 0; k < offs.size() && got < total; k++) {
        const size_t n = std::min<size_t>(k < cnts.size() ? cnts[k] : total - got, total - got);
        if ((size_t)offs[k] + n > f.size()) return false;
        memcpy(s.px.data() + got, f.data() + offs[k], n);
        got += n;
    }
    (void)rps;
    if (got != total) { fprintf(stderr, "TIFF: short pixel data\n"); return false; }
    if (bps == 16)   // PNG wants big-endian samples (image_loader.cpp:29-36)
        for (size_t i = 0; i + 1 < total; i += 2) std::swap(s.px[i], s.px[i + 1]);
    if (photo == 0)  // MINISWHITE (image_loader.cpp:204-210)
        for (auto& b : s.px) b ^= 0xFF;
    return true;
}

// ---------------------------------------------------------------------------
// DICOM Part 10, uncompressed little-endian (explicit or implicit VR)
// ---------------------------------------------------------------------------
struct DcmParse {
    const uint8_t* p;
    size_t n;
    bool explicit_vr = true;
};

bool long_vr(const char* vr)
{
    static const char* L[] = {"OB", "OD", "OF", "OL", "OW", "SQ", "UC", "UN", "UR", "UT", "OV", "SV", "UV"};
    for (auto v : L) if (vr[0] == v[0] && vr[1] == v[1]) return true;
    return false;
}

struct DcmElem {
    uint32_t tag;
    char vr[3];
    uint32_t len;     // 0xFFFFFFFF = undefined
    size_t off;       // value offset
};

bool dcm_next(const uint8_t* p, size_t n, size_t& pos, bool explicit_vr, DcmElem& e)
{
    if (pos + 8 > n) return false;
    const uint32_t g = rd16(p + pos), el = rd16(p + pos + 2);
    e.tag = (g << 16) | el;
    if (g == 0xFFFE) {                                   // item / delimiters: no VR
        e.vr[0] = 0; e.len = rd32(p + pos + 4); e.off = pos + 8;
        pos += 8;
        return true;
    }
    if (explicit_vr) {
        e.vr[0] = (char)p[pos + 4]; e.vr[1] = (char)p[pos + 5]; e.vr[2] = 0;
        if (long_vr(e.vr)) {
            if (pos + 12 > n) return false;
            e.len = rd32(p + pos + 8); e.off = pos + 12;
        } else {
            e.len = rd16(p + pos + 6); e.off = pos + 8;
        }
    } else {
        e.vr[0] = 0; e.len = rd32(p + pos + 4); e.off = pos + 8;
        if (e.tag == 0x7FE00010u) { e.vr[0] = 'O'; e.vr[1] = 'W'; }
    }
    pos = e.off;
    return true;
}

// Skips an undefined-length sequence starting at pos (just after its header).
bool dcm_skip_undefined(const uint8_t* p, size_t n, size_t& pos, bool explicit_vr, int depth)
{
    if (depth > 32) return false;
    DcmElem e;
    while (dcm_next(p, n, pos, explicit_vr, e)) {
        if (e.tag == 0xFFFEE0DDu) return true;            // sequence delimitation
        if (e.tag == 0xFFFEE00Du) continue;               // item delimitation
        if (e.tag == 0xFFFEE000u) {                       // item
            if (e.len != 0xFFFFFFFFu) { pos = e.off + e.len; continue; }
            continue;                                     // undefined item: parse its elements inline
        }
        if (e.len == 0xFFFFFFFFu) { if (!dcm_skip_undefined(p, n, pos, explicit_vr, depth + 1)) return false; continue; }
        pos = e.off + e.len;
        if (pos > n) return false;
    }
    return false;
}

double ds_first(const uint8_t* p, uint32_t len, bool* ok)
{
    std::string s((const char*)p, len);
    const size_t bs = s.find('\\');
    if (bs != std::string::npos) s.resize(bs);
    char* end = nullptr;
    const double v = strtod(s.c_str(), &end);
    *ok = end != s.c_str();
    return v;
}

// A parsed Part 10 file: its bytes, the pixel data's place and the image
// attributes; frames are cut from it without reading the file again.
struct DicomFile {
    std::vector<uint8_t> bytes;
    size_t pix_off = 0, pix_len = 0;
    ImageInfo inf;
    DicomPixelParams dp;
    int nframes = 1;
};

bool parse_dicom(const char* path, DicomFile& df)
{
    std::vector<uint8_t>& f = df.bytes;
    if (!read_file(path, f) || f.size() < 132 || memcmp(&f[128], "DICM", 4) != 0) {
        fprintf(stderr, "DICOM: cannot load '%s' (not Part 10)\n", path);
        return false;
    }
    const uint8_t* p = f.data();
    const size_t n = f.size();
    size_t pos = 132;
    std::string ts;
    DcmElem e;
    // file meta group (always explicit VR little endian)
    while (pos + 8 <= n && rd16(p + pos) == 0x0002) {
        if (!dcm_next(p, n, pos, true, e)) return false;
        if (e.len == 0xFFFFFFFFu || e.off + e.len > n) { fprintf(stderr, "DICOM: malformed file meta\n"); return false; }
        if (e.tag == 0x00020010u) ts.assign((const char*)p + e.off, e.len);
        pos = e.off + e.len;
    }
    while (!ts.empty() && (ts.back() == 0 || ts.back() == ' ')) ts.pop_back();
    bool explicit_vr;
    if (ts == "1.2.840.10008.1.2.1") explicit_vr = true;
    else if (ts == "1.2.840.10008.1.2") explicit_vr = false;
    else { fprintf(stderr, "DICOM: transfer syntax %s is not read natively\n", ts.c_str()); return false; }
    uint32_t rows = 0, cols = 0, ba = 16, bs = 16, hb = 0, rep = 0, spp = 1;
    bool has_hb = false, has_slope = false, has_icpt = false, has_wc = false, has_ww = false;
    double slope = 1.0, icpt = 0.0, wc = 0.0, ww = 0.0;
    int nframes = 1;
    size_t pix_off = 0, pix_len = 0;
    while (pos < n) {
        if (!dcm_next(p, n, pos, explicit_vr, e)) break;
        if (e.len == 0xFFFFFFFFu) {
            if (e.tag == 0x7FE00010u) { fprintf(stderr, "DICOM: encapsulated pixel data\n"); return false; }
            if (!dcm_skip_undefined(p, n, pos, explicit_vr, 0)) return false;
            continue;
        }
        if (e.off + e.len > n) return false;
        const uint8_t* v = p + e.off;
        bool ok = false;
        // the US attributes below need their two value bytes
        const bool us_tag = e.tag == 0x00280002u || e.tag == 0x00280010u || e.tag == 0x00280011u ||
                            (e.tag >= 0x00280100u && e.tag <= 0x00280103u);
        if (us_tag && e.len < 2) { fprintf(stderr, "DICOM: short value of (0028,%04X)\n", e.tag & 0xFFFFu); return false; }
        switch (e.tag) {
        case 0x00280010u: rows = rd16(v); break;
        case 0x00280011u: cols = rd16(v); break;
        case 0x00280100u: ba = rd16(v); break;
        case 0x00280101u: bs = rd16(v); break;
        case 0x00280102u: hb = rd16(v); has_hb = true; break;
        case 0x00280103u: rep = rd16(v); break;
        case 0x00280002u: spp = rd16(v); break;
        case 0x00280008u: { const double d = ds_first(v, e.len, &ok); if (ok && d > 0) nframes = (int)d; break; }
        case 0x00281050u: wc = ds_first(v, e.len, &ok); has_wc = ok; break;
        case 0x00281051u: ww = ds_first(v, e.len, &ok); has_ww = ok; break;
        case 0x00281052u: icpt = ds_first(v, e.len, &ok); has_icpt = ok; break;
        case 0x00281053u: slope = ds_first(v, e.len, &ok); has_slope = ok; break;
        case 0x7FE00010u: pix_off = e.off; pix_len = e.len; break;
        default: break;
        }
        pos = e.off + e.len;
    }
    if (!rows || !cols || (ba != 8 && ba != 16) || spp < 1 || spp > 4 || !pix_len) {
        fprintf(stderr, "DICOM: unsupported or incomplete image attributes\n");
        return false;
    }
    if (!has_hb) hb = bs - 1;                                  // dicom_loader.cpp:329-330
    df.nframes = nframes;
    df.pix_off = pix_off;
    df.pix_len = pix_len;
    df.inf.width = cols; df.inf.height = rows; df.inf.channels = (int)spp; df.inf.bits_per_sample = (int)ba;
    df.inf.bpp = (int)(spp * ba / 8);
    DicomPixelParams& d = df.dp;
    d = DicomPixelParams{};
    d.bits_allocated = (int)ba; d.bits_stored = (int)bs; d.high_bit = (int)hb; d.pixel_rep = (int)rep;
    if ((has_slope || has_icpt) && (std::fabs(slope - 1.0) > 1e-9 || std::fabs(icpt) > 1e-9)) {   // :380-388
        d.apply_rescale = 1; d.rescale_slope = (float)slope; d.rescale_intercept = (float)icpt;
    }
    if (has_wc && has_ww && ww > 0.0) {                                                         // :390-397
        d.apply_window = 1; d.window_center = (float)wc; d.window_width = (float)ww;
    }
    return true;
}

// Frame `frame` of a parsed file as an in-memory source.
bool dicom_frame(const DicomFile& df, int frame, BufferSource& s)
{
    if (frame < 0 || frame >= df.nframes) { fprintf(stderr, "DICOM: frame %d out of range\n", frame); return false; }
    s.inf = df.inf;
    const size_t fbytes = (size_t)df.inf.height * df.inf.width * df.inf.bpp;
    if ((size_t)(frame + 1) * fbytes > df.pix_len) { fprintf(stderr, "DICOM: short pixel data\n"); return false; }
    const uint8_t* base = df.bytes.data() + df.pix_off + (size_t)frame * fbytes;
    s.px.assign(base, base + fbytes);
    s.dp = df.dp;
    s.has_dicom = true;
    s.next = 0;
    return true;
}

// ---------------------------------------------------------------------------
// device-resident whole-image encode (use_gpu_deflate)
// ---------------------------------------------------------------------------
struct DeviceEncodeCache {
    r2p_encoder* enc = nullptr;
    size_t cap = 0;
    int lz77 = -1, level = -1;
    uint8_t* d_src = nullptr; size_t src_cap = 0;
    uint8_t* d_png = nullptr; size_t png_cap = 0;
    ~DeviceEncodeCache()
    {
        if (enc) r2p_encoder_destroy(enc);
        if (d_src) (void)hipFree(d_src);
        if (d_png) (void)hipFree(d_png);
    }
};
thread_local DeviceEncodeCache tl_dev;

bool grow(uint8_t** p, size_t* cap, size_t need)
{
    if (need <= *cap) return true;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    if (hipMalloc((void**)p, need) != hipSuccess) { *cap = 0; return false; }
    *cap = need;
    return true;
}

// One writer thread per encoding thread: jobs of (path, pinned buffer),
// at most kMaxInFlight of them, buffers reused once written.
class FileWriteQueue {
public:
    static constexpr int kMaxInFlight = 3;

    ~FileWriteQueue()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        // (pinned buffers are left to process exit: the runtime may be gone)
    }

    // A host buffer of at least `need` bytes (pinned when possible), once
    // fewer than kMaxInFlight writes are pending.
    uint8_t* acquire(size_t need, size_t* cap)
    {
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ < kMaxInFlight; });
        for (size_t i = 0; i < free_.size(); i++)
            if (free_[i].second >= need) {
                auto b = free_[i];
                free_.erase(free_.begin() + (long)i);
                *cap = b.second;
                return b.first;
            }
        lk.unlock();
        uint8_t* p = nullptr;
        const size_t c = need + need / 8 + 4096;
        if (hipHostMalloc((void**)&p, c, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p = (uint8_t*)malloc(c);                // pageable fallback
            if (!p) return nullptr;
            std::lock_guard<std::mutex> g(mu_);
            pageable_.push_back(p);
        }
        *cap = c;
        return p;
    }

    void submit(const char* path, uint8_t* buf, size_t len, size_t cap)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back({path, buf, len, cap});
            pending_++;
            if (!th_.joinable()) th_ = std::thread([this] { loop(); });
        }
        cv_.notify_all();
    }

    // Waits for every submitted write; false if any failed since the last wait.
    bool wait()
    {
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
        const bool ok = !failed_;
        failed_ = false;
        return ok;
    }

private:
    struct Job { std::string path; uint8_t* buf; size_t len, cap; };

    void loop()
    {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                j = std::move(q_.front());
                q_.pop_front();
            }
            FILE* f = fopen(j.path.c_str(), "wb");
            bool ok = f != nullptr;
            if (!f) fprintf(stderr, "Cannot open output: %s\n", j.path.c_str());
            else {
                ok = fwrite(j.buf, 1, j.len, f) == j.len;
                ok = (fclose(f) == 0) && ok;
                if (!ok) fprintf(stderr, "write failed: %s\n", j.path.c_str());
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                free_.push_back({j.buf, j.cap});
                pending_--;
                if (!ok) failed_ = true;
            }
            done_.notify_all();
        }
    }

    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::deque<Job> q_;
    std::vector<std::pair<uint8_t*, size_t>> free_;
    std::vector<uint8_t*> pageable_;
    int pending_ = 0;
    bool failed_ = false, stop_ = false;
    std::thread th_;
};
thread_local FileWriteQueue tl_writer;

std::atomic<long long> g_files{0}, g_dicom_ms{0}, g_total_ms{0};

// The PNG of one image; with `async` the file is written by the writer
// thread after this returns (tl_writer.wait() collects the result).
bool encode_gpu_whole(const ImageInfo& info, const uint8_t* px, const DicomPixelParams* dicom,
                      const PipelineConfig& cfg, const char* out_path, bool async)
{
    r2p_image im{};
    im.width = info.width; im.height = info.height; im.channels = info.channels;
    im.in_bits = info.bits_per_sample;
    im.pix_mode = dicom ? R2P_PIX_DICOM : R2P_PIX_READY;
    if (dicom) memcpy(&im.dicom, dicom, sizeof(im.dicom));
    const size_t fb = r2p_image_filtered_bytes(&im);
    if (!fb) { fprintf(stderr, "encode: %s\n", r2p_last_error()); return false; }
    DeviceEncodeCache& c = tl_dev;
    const int lz = cfg.use_gpu_lz77 ? 1 : 0;
    if (!c.enc || c.cap < fb || c.lz77 != lz || c.level != cfg.deflate_level) {
        if (c.enc) r2p_encoder_destroy(c.enc);
        c.enc = nullptr;
        r2p_options o;
        r2p_options_default(&o);
        o.lz77 = lz;
        o.zlib_level = cfg.deflate_level;
        if (r2p_encoder_create(fb, &o, &c.enc) != R2P_OK) { fprintf(stderr, "encode: %s\n", r2p_last_error()); return false; }
        c.cap = fb; c.lz77 = lz; c.level = cfg.deflate_level;
    }
    const size_t in_b = r2p_image_input_bytes(&im);
    const size_t bound = r2p_png_bound(&im, 65536, 65536);
    if (!grow(&c.d_src, &c.src_cap, in_b) || !grow(&c.d_png, &c.png_cap, bound)) { fprintf(stderr, "encode: out of device memory\n"); return false; }
    hipStream_t s = (hipStream_t)r2p_encoder_stream(c.enc);
    if (hipMemcpyAsync(c.d_src, px, in_b, hipMemcpyHostToDevice, s) != hipSuccess) return false;
    if (r2p_encode_device(c.enc, &im, c.d_src, c.d_png, c.png_cap, nullptr) != R2P_OK) {
        fprintf(stderr, "encode: %s\n", r2p_last_error());
        return false;
    }
    size_t len = 0;
    if (r2p_encoder_sync(c.enc, &len, nullptr, nullptr) != R2P_OK) return false;
    size_t cap = 0;
    uint8_t* hb = tl_writer.acquire(len, &cap);
    if (!hb) { fprintf(stderr, "encode: out of host memory\n"); return false; }
    if (hipMemcpy(hb, c.d_png, len, hipMemcpyDeviceToHost) != hipSuccess) return false;
    tl_writer.submit(out_path, hb, len, cap);
    return async ? true : tl_writer.wait();
}

// strips: GPU filter + CPU zlib (the reference's default path)
bool encode_gpu_filter_cpu_deflate(ImageSource& src, const PipelineConfig& cfg, const char* out_path)
{
    const ImageInfo& info = src.info();
    const int sh = cfg.strip_height > 0 ? cfg.strip_height : 256;
    const size_t rb = (size_t)info.width * info.bpp;
    GpuFilterContext* g = gpu_filter_create((int)info.width, info.bpp, sh);
    ParallelDeflateConfig dc;
    dc.num_threads = cfg.deflate_threads;
    dc.zlib_level = cfg.deflate_level;
    CPUZlibBackend be(dc, cfg.deflate_threads);
    PngWriter w;
    const int bits = info.bits_per_sample;
    if (!w.open(out_path, info.width, info.height, (uint8_t)bits, color_type(info.channels))) {
        gpu_filter_destroy(g);
        fprintf(stderr, "Cannot open output: %s\n", out_path);
        return false;
    }
    w.begin_idat();
    uint8_t zh[2];
    zlib_header(cfg.deflate_level, zh);
    w.write_idat_bytes(zh, 2);
    ParallelDeflateState st;
    uint8_t* strip = gpu_filter_alloc_pinned((size_t)sh * rb + 16);
    const int nstrips = ((int)info.height + sh - 1) / sh;
    gpu_filter_reset(g);
    bool ok = true;
    for (int s = 0; s < nstrips && ok; s++) {
        const int rows = src.read_strip(strip, sh);
        if (rows <= 0) { ok = false; break; }
        const uint8_t* filt = gpu_filter_process_from_host(g, strip, nullptr, rows, nullptr, src.dicom_params());
        DeflateResult dr = be.compress_strip(filt, gpu_filter_output_size(g, rows), s == nstrips - 1);
        accum_adler(st, dr);
        w.write_idat_bytes(dr.data.data(), dr.data.size());
    }
    uint8_t tr[4];
    zlib_trailer(st, tr);
    w.write_idat_bytes(tr, 4);
    w.end_idat();
    w.close();
    gpu_filter_free_pinned(strip);
    gpu_filter_destroy(g);
    return ok;
}

bool run(ImageSource& src, const char* out_path, const PipelineConfig& cfg, bool async = false)
{
    if (cfg.format != ExportFormat::PNG) { fprintf(stderr, "only PNG export is implemented\n"); return false; }
    const ImageInfo& info = src.info();
    if (!cfg.use_gpu_deflate) return encode_gpu_filter_cpu_deflate(src, cfg, out_path);
    std::vector<uint8_t> px((size_t)info.width * info.bpp * info.height);
    const size_t rb = (size_t)info.width * info.bpp;
    int got = 0;
    while (got < (int)info.height) {
        const int r = src.read_strip(px.data() + (size_t)got * rb, (int)info.height - got);
        if (r <= 0) return false;
        got += r;
    }
    return encode_gpu_whole(info, px.data(), src.dicom_params(), cfg, out_path, async);
}

}  // namespace

bool encode_source_to_png(ImageSource& src, const char* output_path, const PipelineConfig& cfg)
{
    return run(src, output_path, cfg);
}

bool encode_tiff_to_png(const char* in, const char* out, const PipelineConfig& cfg)
{
    auto t0 = Clock::now();
    BufferSource s;
    if (!load_tiff(in, s)) return false;
    const bool ok = run(s, out, cfg);
    if (cfg.use_gpu_deflate)
        pipeline_record_file_times(0, std::chrono::duration_cast<std::chrono::milliseconds>(Clock::now() - t0).count());
    return ok;
}

bool encode_raw_to_png(const char*, const char*, const PipelineConfig&)
{
    fprintf(stderr, "camera RAW decoding (LibRaw) is not part of this encoder\n");
    return false;
}

bool encode_dicom_to_png(const char* in, const char* out, const PipelineConfig& cfg)
{
    auto t0 = Clock::now();
    DicomFile df;
    BufferSource s;
    if (!parse_dicom(in, df) || !dicom_frame(df, cfg.frame >= 0 ? cfg.frame : 0, s)) return false;
    auto t1 = Clock::now();
    const bool ok = run(s, out, cfg);
    if (cfg.use_gpu_deflate)
        pipeline_record_file_times(std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count(),
                                   std::chrono::duration_cast<std::chrono::milliseconds>(Clock::now() - t0).count());
    return ok;
}

// encode_dicom_all_frames_to_png (pipeline.cpp:1760-1879 of the reference):
// the file is read and parsed once; K frame workers (cfg.gpu_streams, else
// 4, at most the frame count) take frames in turn, each with its own device
// encoder and stream (and its share of the deflate threads on the CPU-deflate
// path), writing <out_dir>/slice_%04d.png.  The output folder is created.
bool encode_dicom_all_frames_to_png(const char* in, const char* out_dir, const PipelineConfig& cfg)
{
    DicomFile df;
    if (!parse_dicom(in, df)) return false;
    const int n = df.nframes;
    std::error_code ec;
    std::filesystem::create_directories(out_dir, ec);
    if (ec) {
        fprintf(stderr, "Cannot create output folder '%s': %s\n", out_dir, ec.message().c_str());
        return false;
    }
    fprintf(stdout, "Detected %d frame%s\n", n, n == 1 ? "" : "s");
    const int workers = std::max(1, std::min(cfg.gpu_streams > 0 ? cfg.gpu_streams : 4, n));
    PipelineConfig wcfg = cfg;
    wcfg.deflate_threads = std::max(1, cfg.deflate_threads / workers);
    std::atomic<int> next{0};
    std::atomic<bool> ok{true};
    std::mutex print_mu;
    auto worker = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n || !ok.load()) break;
            BufferSource s;
            if (!dicom_frame(df, i, s)) { ok = false; break; }
            char path[4096];
            snprintf(path, sizeof(path), "%s/slice_%04d.png", out_dir, i);
            {
                std::lock_guard<std::mutex> lk(print_mu);
                fprintf(stdout, "[%d/%d] slice_%04d.png\n", i + 1, n, i);
            }
            if (!run(s, path, wcfg, true)) { ok = false; break; }
        }
        if (!tl_writer.wait()) ok = false;           // this worker's files written
    };
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::vector<std::thread> th;
    for (int w = 0; w < workers; w++)
        th.emplace_back([&, dev] {
            (void)hipSetDevice(dev);
            worker();
        });
    for (auto& t : th) t.join();
    if (!ok.load()) fprintf(stderr, "Export aborted after an error.\n");
    return ok.load();
}

void pipeline_reset_gpu_batch_stats() { g_files = 0; g_dicom_ms = 0; g_total_ms = 0; }

void pipeline_record_file_times(long long dicom_ms, long long total_ms)
{
    g_files++;
    g_dicom_ms += dicom_ms;
    g_total_ms += total_ms;
}

void pipeline_print_gpu_batch_summary(int total_files, int succeeded, double total_wall_s, long long sum_ms,
                                      int workers)
{
    const long long n = g_files.load();
    fprintf(stdout, "GPU batch: %d/%d files, %.2f s wall, %d workers, %.2f files/s\n", succeeded, total_files,
            total_wall_s, workers, total_wall_s > 0 ? succeeded / total_wall_s : 0.0);
    if (n)
        fprintf(stdout, "  per file: decode %.1f ms, total %.1f ms (batch clock %lld ms)\n",
                (double)g_dicom_ms.load() / n, (double)g_total_ms.load() / n, sum_ms);
}
