// compression_backend.h -- "compress one filtered strip" behind one interface.
// Drop-in for include/compression_backend.h:33-78 of the reference, plus the
// GPU implementation the reference left for later: GpuDeflateBackend runs the
// segment-parallel HIP encoder and, because its strips end byte-aligned with a
// sync flush, its output concatenates exactly like CPUZlibBackend's.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>

#include "parallel_deflate.h"
#include "thread_pool.h"

class CompressionBackend {
public:
    virtual ~CompressionBackend() = default;
    // input: host bytes of one filtered strip; is_last only for the final strip.
    virtual DeflateResult compress_strip(const uint8_t* input, size_t input_size, bool is_last) = 0;
    virtual const char* name() const = 0;
};

class CPUZlibBackend : public CompressionBackend {
public:
    CPUZlibBackend(const ParallelDeflateConfig& cfg, int num_threads) : cfg_(cfg), pool_(num_threads) {}
    DeflateResult compress_strip(const uint8_t* input, size_t input_size, bool is_last) override
    {
        return deflate_strip(input, input_size, cfg_, pool_, is_last);
    }
    const char* name() const override { return "zlib (CPU)"; }

private:
    ParallelDeflateConfig cfg_;
    ThreadPool pool_;
};

struct r2p_deflate;

class GpuDeflateBackend : public CompressionBackend {
public:
    // max_strip_bytes bounds one strip; lz77 enables matching.
    explicit GpuDeflateBackend(size_t max_strip_bytes, bool lz77 = true);
    ~GpuDeflateBackend() override;
    DeflateResult compress_strip(const uint8_t* input, size_t input_size, bool is_last) override;
    const char* name() const override { return "r2p segment DEFLATE (MI355X)"; }
    bool ok() const { return ctx_ != nullptr; }

private:
    r2p_deflate* ctx_ = nullptr;
    uint8_t* d_in_ = nullptr;
    size_t cap_ = 0;
};

// Returns the GPU backend when a gfx9 device is usable, else the CPU one.
std::unique_ptr<CompressionBackend> make_compression_backend(const ParallelDeflateConfig& cfg,
                                                             int num_threads);
