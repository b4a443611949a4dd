/*
 * r2p_oracle.h -- CPU oracle for the PNG encode hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links this library; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and
 * only as the checker (or as the timed CPU baseline).  Every function is a
 * scalar restatement of the reference algorithm it cites; reference paths are
 * relative to the RAW2PNG-ConversionThruGPU source tree.
 *
 * Float semantics: IEEE binary32, operations evaluated in source order, no
 * contraction (built with -ffp-contract=off), correctly rounded division.
 */
#ifndef R2P_ORACLE_H
#define R2P_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and types as DicomPixelParams (include/dicom_params.h:6-21). */
typedef struct r2po_dicom_params {
    int   bits_allocated;
    int   bits_stored;
    int   high_bit;
    int   pixel_rep;
    int   apply_rescale;
    float rescale_slope;
    float rescale_intercept;
    int   apply_window;
    float window_center;
    float window_width;
} r2po_dicom_params;

/* a2: dicom_preprocess_kernel (src/gpu_filter.cu:71-162).  Converts `samples`
 * little-endian stored samples to PNG-ready samples (big-endian when
 * bits_allocated == 16).  in and out may not alias. */
void r2po_dicom_window(const uint8_t* in, uint8_t* out, size_t samples,
                       const r2po_dicom_params* p);

/* a3: try_apply_linear_window coefficient derivation
 * (src/batch_processor.cpp:202-248).  Returns 1 and fills A/B when the Q15
 * path applies, 0 when the reference would fall back to the LUT. */
int r2po_q15_coeffs(const r2po_dicom_params* p, int16_t* a_fp, int32_t* b_fp);

/* a3: the Q15 pixel loop (src/batch_processor.cpp:250-264). */
void r2po_q15_apply(const uint8_t* in_le, uint8_t* out8, size_t pixels,
                    int16_t a_fp, int32_t b_fp, int pixel_rep);

/* a4: build_dicom16_lut (src/batch_processor.cpp:100-147); lut_size = 1<<bits_stored. */
void r2po_lut8_build(uint8_t* lut, const r2po_dicom_params* p, int lut_size);

/* a4: apply_dicom16_lut (src/batch_processor.cpp:157-185). */
void r2po_lut8_apply(const uint8_t* in_le, uint8_t* out8, size_t pixels,
                     const uint8_t* lut, int shift, uint16_t mask16);

/* a5: encode_predecoded_dicom's 16->8 bit dispatch
 * (src/batch_processor.cpp:397-409): Q15 when it applies, else the LUT. */
void r2po_window8(const uint8_t* in_le, uint8_t* out8, size_t pixels,
                  const r2po_dicom_params* p);

/* a7: filter_select_kernel (src/gpu_filter.cu:183-309) over `rows` rows of
 * row_bytes bytes.  prior = the row above row 0 (NULL = zeros).  out receives
 * rows * (row_bytes + 1) bytes. */
void r2po_filter_rows(const uint8_t* in, const uint8_t* prior, uint8_t* out,
                      int rows, int row_bytes, int bpp);

/* Per-row filter scores (5 per row) as the reference computes them. */
void r2po_filter_scores(const uint8_t* in, const uint8_t* prior, uint32_t* scores,
                        int rows, int row_bytes, int bpp);

/* PNG unfilter (PNG spec section 9), the inverse of any filter choice. */
int r2po_unfilter_rows(const uint8_t* filtered, const uint8_t* prior, uint8_t* out,
                       int rows, int row_bytes, int bpp);

/* a12 / a13: checksums from RFC 1950 and PNG Annex D, written from the math. */
uint32_t r2po_adler32(uint32_t adler, const uint8_t* p, size_t n);
uint32_t r2po_adler32_combine(uint32_t a1, uint32_t a2, uint64_t len2);
uint32_t r2po_crc32(uint32_t crc, const uint8_t* p, size_t n);
uint32_t r2po_crc32_combine(uint32_t c1, uint32_t c2, uint64_t len2);

/* a15: zlib_header FLG byte by level (src/parallel_deflate.cpp:214-228). */
uint8_t r2po_zlib_flg(int level);

/* The CPU baseline: window (mode) + filter + chunked zlib deflate following
 * deflate_strip (src/parallel_deflate.cpp:117-192), zlib framing and
 * PngWriter's 32 KiB IDAT chunks (src/png_writer.cpp:33-125).
 *   mode 0: input already PNG-ready (8-bit or big-endian 16-bit)
 *   mode 1: 16-bit little-endian input, byte-swapped to big-endian
 *   mode 2: DICOM float window (a2), output bits = bits_allocated
 *   mode 3: DICOM 16->8 bit (a5: Q15 else LUT)
 * Returns the PNG size, or 0 if out_cap is too small / arguments invalid.
 * filtered_out (optional) receives the filtered stream. */
size_t r2po_encode_png(const uint8_t* in, int width, int height, int channels,
                       int in_bits, int mode, const r2po_dicom_params* p,
                       int strip_height, int level, int threads,
                       uint8_t* out, size_t out_cap, uint8_t* filtered_out);

/* Bytes of the filtered stream for an image: height * (width*bpp_out + 1). */
size_t r2po_filtered_size(int width, int height, int channels, int out_bits);

const char* r2po_zlib_version(void);

#ifdef __cplusplus
}
#endif
#endif
