// dicom_params.h -- pixel transform parameters of a DICOM source.
// Drop-in for include/dicom_params.h:6-21 of the reference: same fields,
// order, types and defaults (the struct is passed by value to the device).
#pragma once

struct DicomPixelParams {
    int   bits_allocated    = 16;        // (0028,0100)
    int   bits_stored       = 16;        // (0028,0101)
    int   high_bit          = 15;        // (0028,0102); right shift = high_bit - bits_stored + 1
    int   pixel_rep         = 0;         // (0028,0103) 0 unsigned, 1 two's complement
    int   apply_rescale     = 0;         // slope/intercept are non-trivial
    float rescale_slope     = 1.f;       // (0028,1053)
    float rescale_intercept = 0.f;       // (0028,1052)
    int   apply_window      = 0;         // window centre/width present
    float window_center     = 32768.f;   // (0028,1050)
    float window_width      = 65536.f;   // (0028,1051)
};
