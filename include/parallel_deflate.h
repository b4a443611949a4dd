// parallel_deflate.h -- chunked multi-threaded CPU DEFLATE with zlib.
// Drop-in for include/parallel_deflate.h:17-80 of the reference: strips are
// cut into chunks deflated in parallel (raw DEFLATE, Z_FULL_FLUSH, the very
// last chunk Z_FINISH) so their concatenation is one valid stream.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "thread_pool.h"

struct ParallelDeflateConfig {
    int num_threads = 6;
    int zlib_level  = 3;
};

struct DeflateResult {
    std::vector<uint8_t> data;        // raw DEFLATE bytes of the strip
    unsigned long        strip_adler; // adler32 of the strip's input
    size_t               input_size;
    long long            init_us     = 0;
    long long            compress_us = 0;
};

struct ParallelDeflateState {
    unsigned long running_adler = 1;
    bool          started       = false;
};

DeflateResult deflate_strip(const uint8_t* input, size_t input_size, const ParallelDeflateConfig& cfg,
                            ThreadPool& pool, bool is_last);
void accum_adler(ParallelDeflateState& state, const DeflateResult& dr);
void zlib_header(int level, uint8_t out[2]);
void zlib_trailer(const ParallelDeflateState& state, uint8_t out[4]);
unsigned long adler32_of(const uint8_t* data, size_t len);
unsigned long crc32_of(const uint8_t* data, size_t len);
