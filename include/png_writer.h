// png_writer.h -- PNG container writer taking caller-made zlib bytes.
// Drop-in for include/png_writer.h:7-57 of the reference (IDAT chunks of at
// most 32 KiB, CRC-32 per chunk, IEND on close).
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstdio>

enum class PngColorType : uint8_t { Gray = 0, RGB = 2, GrayA = 4, RGBA = 6 };

class PngWriter {
public:
    PngWriter() = default;
    ~PngWriter();
    bool open(const char* path, uint32_t width, uint32_t height, uint8_t bit_depth,
              PngColorType color_type);
    void begin_idat();
    void write_idat_bytes(const uint8_t* data, size_t size);
    void end_idat();
    void close();
    bool is_open() const { return file_ != nullptr; }

private:
    void write_chunk(const char type[4], const uint8_t* data, uint32_t len);
    void flush_idat_buf();
    FILE*   file_ = nullptr;
    uint8_t idat_buf_[32768];
    size_t  idat_pos_ = 0;
};
