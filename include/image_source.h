// image_source.h -- strip-wise image source consumed by encode_source_to_png.
// Drop-in for include/image_loader.h:6-12 (ImageInfo) and
// include/image_source.h:7-17 (ImageSource) of the reference.
#pragma once
#include <cstddef>
#include <cstdint>

#include "dicom_params.h"

struct ImageInfo {
    uint32_t width           = 0;
    uint32_t height          = 0;
    int      channels        = 0;   // 1, 3 (or 2, 4)
    int      bits_per_sample = 0;   // 8 or 16
    int      bpp             = 0;   // channels * bits_per_sample / 8
};

struct ImageSource {
    virtual ~ImageSource() = default;
    virtual const ImageInfo& info() const = 0;
    // Copies up to max_rows rows into out; returns rows copied, 0 at the end,
    // a negative value on error.
    virtual int read_strip(uint8_t* out, int max_rows) = 0;
    // Non-null only for DICOM sources: the samples are raw little-endian
    // stored values that the encoder windows on the device.
    virtual const DicomPixelParams* dicom_params() const { return nullptr; }
    virtual int num_frames() const { return 1; }
};
