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

// Standalone CPU DEFLATE throughput table (replaces include/parallel_deflate.h:82
// and src/parallel_deflate.cpp:261-335 of the reference, the CLI's
// --bench-deflate): deflate_strip over strip_bytes of a high-entropy synthetic
// pattern (byte i = the top byte of i * 0x9E3779B9, as the reference's) with
// num_threads chunks at levels 0, 1 and 3; 3 warm-ups and 15 timed strips per
// level; prints ms per strip, input MB/s, compressed size and the summed
// z_stream set-up time to stdout.
void bench_deflate(size_t strip_bytes, int num_threads);
