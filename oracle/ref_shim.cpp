// ref_shim.cpp -- extern "C" entry points over the reference's own zlib
// framing sources (src/parallel_deflate.cpp, src/png_writer.cpp), compiled
// where they lie by oracle/Makefile into oracle/_ref/libref.so.
// TEST INFRASTRUCTURE ONLY: used to pin the oracle's framing restatement.
#include "parallel_deflate.h"
#include "png_writer.h"
#include "thread_pool.h"

#include <cstring>
#include <memory>

extern "C" {

void ref_zlib_header(int level, uint8_t out[2]) { zlib_header(level, out); }

void ref_zlib_trailer(uint32_t adler, uint8_t out[4])
{
    ParallelDeflateState st;
    st.running_adler = adler;
    st.started = true;
    zlib_trailer(st, out);
}

// deflate_strip over one strip (src/parallel_deflate.cpp:117-192).
// Returns the raw-deflate byte count (0 with *overflow = 1 if out_cap is short).
size_t ref_deflate_strip(const uint8_t* in, size_t n, int level, int threads,
                         int is_last, uint8_t* out, size_t out_cap,
                         uint32_t* strip_adler, int* overflow)
{
    static std::unique_ptr<ThreadPool> pool;
    static int pool_threads = 0;
    if (!pool || pool_threads != threads) { pool.reset(new ThreadPool(threads)); pool_threads = threads; }
    ParallelDeflateConfig cfg;
    cfg.num_threads = threads;
    cfg.zlib_level = level;
    DeflateResult r = deflate_strip(in, n, cfg, *pool, is_last != 0);
    *strip_adler = (uint32_t)r.strip_adler;
    if (r.data.size() > out_cap) { *overflow = 1; return 0; }
    *overflow = 0;
    std::memcpy(out, r.data.data(), r.data.size());
    return r.data.size();
}

// accum_adler (src/parallel_deflate.cpp:197-209) over a list of strips.
uint32_t ref_accum_adler(const uint32_t* adlers, const uint64_t* sizes, int n)
{
    ParallelDeflateState st;
    for (int i = 0; i < n; i++) {
        DeflateResult dr;
        dr.strip_adler = adlers[i];
        dr.input_size = (size_t)sizes[i];
        accum_adler(st, dr);
    }
    return (uint32_t)st.running_adler;
}

// PngWriter (src/png_writer.cpp:33-125): writes a PNG holding `zlib_stream`.
int ref_png_write(const char* path, uint32_t w, uint32_t h, int bit_depth,
                  int color_type, const uint8_t* zlib_stream, size_t n)
{
    PngWriter wr;
    if (!wr.open(path, w, h, (uint8_t)bit_depth, (PngColorType)color_type)) return 0;
    wr.begin_idat();
    wr.write_idat_bytes(zlib_stream, n);
    wr.end_idat();
    wr.close();
    return 1;
}

}  // extern "C"
