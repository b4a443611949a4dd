/*
 * r2p.h -- C-ABI of the MI355X (gfx950) PNG encoder.
 *
 * Plain C: opaque handles, raw pointers, sizes and integer status codes; no
 * C++ or HIP types cross this boundary.  Every entry point names the
 * reference interface it replaces (paths relative to the
 * RAW2PNG-ConversionThruGPU tree).  The C++ drop-in headers next to this file
 * (gpu_filter.h, gpu_deflate_backend.h, gpu_adler32.h, gpu_crc32.h,
 * gpu_png_assemble.h, compression_backend.h, pipeline.h) are implemented on
 * top of these functions.
 *
 * Pointers documented as "device" must address memory on the current HIP
 * device (hipMalloc'd, or a torch CUDA tensor's data_ptr()).  Streams are
 * passed as void* (a hipStream_t; NULL = the context's own stream).
 *
 * Thread safety: a context is used by one host thread at a time; distinct
 * contexts may be used concurrently.
 */
#ifndef R2P_H
#define R2P_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum {
    R2P_OK              = 0,
    R2P_ERR_ARG         = 1,  /* invalid argument / parameter combination */
    R2P_ERR_HIP         = 2,  /* a HIP runtime call failed (see r2p_last_error) */
    R2P_ERR_SIZE        = 3,  /* input exceeds the capacity fixed at create time */
    R2P_ERR_NODEV       = 4,  /* no usable gfx9 device */
    R2P_ERR_UNSUPPORTED = 5,  /* valid input this build does not handle */
    R2P_ERR_STATE       = 6   /* call out of order (e.g. collect without launch) */
};

/* Human-readable text of the last error on the calling thread. */
const char* r2p_last_error(void);
const char* r2p_version(void);

/* Binds the calling thread to `device`, checks it is a gfx9 GPU and uploads
 * the constant tables.  Called implicitly by every *_create. */
int r2p_init(int device);

/* ---- pixel transform parameters ----------------------------------------- */
/* Same field order and types as DicomPixelParams (include/dicom_params.h:6-21). */
typedef struct r2p_dicom_params {
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
} r2p_dicom_params;

/* How raw input samples become PNG samples. */
enum {
    R2P_PIX_READY  = 0, /* already PNG-ready: 8-bit, or big-endian 16-bit
                           (TIFF/RAW after the loader's byteswap, image_loader.cpp:29-36) */
    R2P_PIX_SWAP16 = 1, /* little-endian 16-bit, byte-swapped to big-endian */
    R2P_PIX_DICOM  = 2, /* dicom_preprocess_kernel semantics (gpu_filter.cu:71-162):
                           shift/mask/sign/rescale/PS3.3 window, output bits = bits_allocated */
    R2P_PIX_DICOM8 = 3  /* 16-bit DICOM -> 8-bit as encode_predecoded_dicom does
                           (batch_processor.cpp:397-409): Q15 (:202-266) else LUT (:100-185) */
};

/* One image as the encoder sees it. */
typedef struct r2p_image {
    uint32_t width, height;
    int      channels;        /* 1, 2, 3 or 4 */
    int      in_bits;         /* 8 or 16: bits per stored input sample */
    int      pix_mode;        /* R2P_PIX_* */
    int64_t  src_pitch;       /* bytes between input rows; 0 = packed */
    r2p_dicom_params dicom;   /* used by R2P_PIX_DICOM / R2P_PIX_DICOM8 */
} r2p_image;

/* Output bits per sample and filtered row size of an image (no device work). */
int    r2p_image_out_bits(const r2p_image* img);
size_t r2p_image_row_bytes(const r2p_image* img);       /* width*bpp_out */
size_t r2p_image_filtered_bytes(const r2p_image* img);  /* height*(row_bytes+1) */
size_t r2p_image_input_bytes(const r2p_image* img);
/* Upper bound of the PNG produced by r2p_encode_* for this image and options. */
size_t r2p_png_bound(const r2p_image* img, int segment_bytes, int idat_chunk_bytes);

/* ---- encoder options ---------------------------------------------------- */
typedef struct r2p_options {
    int segment_bytes;     /* DEFLATE segment (independent fixed-Huffman block, closed by a
                              sync flush), a multiple of 4096; default 65536 */
    int lz77;              /* 0: literal-only fixed Huffman (the reference default,
                              pipeline.h:40); 1: LZ77 over a 32 KiB primed window, hashed
                              candidates searched at every 4th position (default); 2: at
                              every 2nd position (slower; with fixed-Huffman blocks about
                              1-4 % smaller output) */
    int idat_chunk_bytes;  /* IDAT chunk payload size; default 65536 */
    int zlib_level;        /* only sets the zlib header FLEVEL (parallel_deflate.cpp:214-228) */
    int band_rows;         /* rows per filter workgroup; 0 = auto */
    int lane_bytes;        /* LZ77 parse span per lane; 16 is the one supported value */
    int wg_threads;        /* DEFLATE workgroup size (256 or 512); 0 = auto (256) */
    int slice_bytes;       /* DEFLATE work unit: one workgroup per slice of a segment (a power
                              of two >= 4096 dividing segment_bytes); 0 = auto by the image's
                              filtered size (16 KiB from 8 MiB, 8 KiB from 4 MiB, else 4 KiB).
                              Changes speed only, never the decoded output */
    int huffman;           /* 0: fixed-Huffman blocks (the reference's, gpu_deflate_backend.h:3);
                              1: each segment a dynamic-Huffman block where that is smaller
                              (plain, batched and served forms; not the no-allocation entry
                              point) */
} r2p_options;

void r2p_options_default(r2p_options* o);

/* ======================================================================== */
/* Plain per-image encoder (the device-resident form of run_pipeline /
 * encode_source_to_png, src/pipeline.cpp:1571-1667, 2058-2062).            */
/* ======================================================================== */
typedef struct r2p_encoder r2p_encoder;

/* Sizes device scratch for images up to max_filtered_bytes filtered bytes. */
int r2p_encoder_create(size_t max_filtered_bytes, const r2p_options* opt, r2p_encoder** out);
int r2p_encoder_destroy(r2p_encoder* enc);
void* r2p_encoder_stream(r2p_encoder* enc);

/* Encodes one image whose raw pixels are already in device memory into a
 * complete PNG file image in device memory (d_png, capacity png_cap).
 * Asynchronous on the encoder's stream; *png_len (host) is valid after
 * r2p_encoder_sync.  d_filtered_out (optional, device) receives the filtered
 * scanlines. */
int r2p_encode_device(r2p_encoder* enc, const r2p_image* img, const void* d_src,
                      void* d_png, size_t png_cap, void* d_filtered_out);
int r2p_encoder_sync(r2p_encoder* enc, size_t* png_len, uint32_t* adler, uint32_t* nseg);

/* Host-to-host convenience: H2D, encode, D2H.  Returns the PNG size in *png_len. */
int r2p_encode_host(r2p_encoder* enc, const r2p_image* img, const void* h_src,
                    void* h_png, size_t png_cap, size_t* png_len);

/* Per-kernel device timings of the last encode (ms, from HIP events).  The
 * events are recorded only while timing is on (r2p_encoder_set_timing; off
 * by default): each is one more packet on the stream, about 3 us of every
 * encode. */
typedef struct r2p_timings {
    float filter_ms, deflate_ms, scan_ms, assemble_ms, total_ms;
} r2p_timings;
int r2p_encoder_set_timing(r2p_encoder* enc, int on);
int r2p_encoder_timings(r2p_encoder* enc, r2p_timings* t);

/* No-allocation variant of r2p_encode_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It allocates nothing: every buffer of the encode lives in one
 * static device array of R2P_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends), e.g. images of up to about 11 MiB of filtered bytes with
 * the default options (2048x2048 16-bit); a larger image, or dynamic-Huffman
 * blocks, is refused (R2P_ERR_SIZE / R2P_ERR_UNSUPPORTED).  The caller passes
 * the input, the PNG buffer (png_cap >= r2p_png_bound) and d_result (device,
 * 16 bytes: uint64 PNG length, then the uint32 Adler-32 at byte 8), all
 * written asynchronously on `stream` (a hipStream_t, NULL = the null stream).
 * opt: NULL = r2p_options_default.  Every call on one GPU shares that GPU's
 * array: calls that overlap on a GPU (other streams, other threads) race. */
#define R2P_NOALLOC_SCRATCH_BYTES (24u << 20)
int r2p_encode_device_noalloc(const r2p_image* img, const void* d_src, void* d_png, size_t png_cap,
                              void* d_result, const r2p_options* opt, void* stream);

/* ======================================================================== */
/* Kernel-level entry points (used by the drop-in C++ facade).               */
/* ======================================================================== */

/* Window + adaptive filter of `rows` rows (filter_select_kernel and
 * dicom_preprocess_kernel, gpu_filter.cu:71-309; run_kernels :404-463).
 * d_prior: PNG-ready row above row 0 (device) or NULL for zeros.
 * d_last_row (optional, device): receives the PNG-ready last row, the carry
 * gpu_filter.cu:428-434 keeps for the next strip. */
int r2p_filter_rows(const r2p_image* img, const void* d_src, int rows,
                    const void* d_prior, void* d_filtered, void* d_last_row,
                    int band_rows, void* stream);

/* Adler-32 of device bytes (gpu_adler32.h:24-44).  Writes the final value to
 * d_result (device uint32).  d_work: r2p_checksum_work_bytes(len) bytes. */
size_t r2p_checksum_work_bytes(size_t len);
int r2p_adler32_device(const void* d_data, size_t len, void* d_work, uint32_t* d_result,
                       void* stream);
/* CRC-32 of device bytes (gpu_crc32.h:19-38). */
int r2p_crc32_device(const void* d_data, size_t len, void* d_work, uint32_t* d_result,
                     void* stream);
/* Host-side combine steps (adler32_combine / crc32_combine semantics). */
uint32_t r2p_adler32_combine(uint32_t a1, uint32_t a2, uint64_t len2);
uint32_t r2p_crc32_combine(uint32_t c1, uint32_t c2, uint64_t len2);

/* Strip-incremental DEFLATE (gpu_deflate_backend.h:40-127). */
typedef struct r2p_deflate r2p_deflate;
int r2p_deflate_create(size_t max_strip_bytes, size_t max_total_bytes, int lz77,
                       int segment_bytes, r2p_deflate** out);
int r2p_deflate_destroy(r2p_deflate* d);
int r2p_deflate_reset(r2p_deflate* d);
/* Compresses one strip of filtered bytes (device).  Every strip ends byte
 * aligned (sync flush) except the one with is_last, which ends the stream. */
int r2p_deflate_strip(r2p_deflate* d, const void* d_filtered, size_t len, int is_last,
                      int bpp, int row_stride);
int r2p_deflate_sync(r2p_deflate* d);
const void* r2p_deflate_output(const r2p_deflate* d);
size_t r2p_deflate_output_bytes(const r2p_deflate* d);
/* Bytes the output buffer holds (the bound fixed at create time).  No
 * reference counterpart: lets the facade register the buffer for its
 * cross-stream ordering. */
size_t r2p_deflate_output_capacity(const r2p_deflate* d);
uint32_t r2p_deflate_adler(const r2p_deflate* d);     /* adler32 of all input so far */
int r2p_deflate_copy_to_host(const r2p_deflate* d, void* h_dst, size_t n);
/* Drops the output delivered so far: the next strip's bytes start at offset 0
 * of the output buffer; the 32 KiB history and the Adler-32 state are kept.
 * Lets a caller that copies each strip's bytes out (a CompressionBackend)
 * size the output buffer for one strip instead of a whole image.  No
 * reference counterpart (the reference keeps the whole image on the device,
 * gpu_deflate_backend.h:40-62). */
int r2p_deflate_rewind(r2p_deflate* d);
void* r2p_deflate_stream(r2p_deflate* d);
typedef struct r2p_lz_stats {
    uint64_t matches, matched_bytes, positions;
    uint32_t max_len;
    double   avg_len, avg_dist;
} r2p_lz_stats;
int r2p_deflate_lz_stats(const r2p_deflate* d, r2p_lz_stats* s);

/* One IDAT chunk from a device slice (gpu_png_assemble.cu:64-154): writes
 * [len][IDAT][(78 FLG)][slice][(adler BE)][crc] to d_chunk, CRC computed on
 * the device.  Returns the chunk size in *chunk_len. */
int r2p_idat_chunk(const void* d_slice, size_t slice_len, int is_first, int is_last,
                   uint32_t adler, int zlib_level, void* d_chunk, size_t chunk_cap,
                   void* d_work, size_t* chunk_len, void* stream);

/* ======================================================================== */
/* Batched form: many images in one persistent launch sequence, captured in
 * a HIP graph (the device-side analogue of encode_folder,
 * src/batch_processor.cpp:482-732).                                         */
/* ======================================================================== */
typedef struct r2p_batch r2p_batch;
int r2p_batch_create(int max_images, size_t max_input_bytes, size_t max_filtered_bytes,
                     const r2p_options* opt, r2p_batch** out);
int r2p_batch_destroy(r2p_batch* b);
/* Declares the images of the next run.  d_src[i]: device pointers to each
 * image's pixels.  Output PNGs land in the batch's device arena. */
int r2p_batch_set(r2p_batch* b, int n, const r2p_image* imgs, const void* const* d_src);
/* Runs the batch (graph replay when use_graph).  Asynchronous. */
int r2p_batch_run(r2p_batch* b, int use_graph);
int r2p_batch_sync(r2p_batch* b);
/* After sync: device pointer and size of image i's PNG. */
int r2p_batch_result(r2p_batch* b, int i, const void** d_png, size_t* len);
int r2p_batch_copy_result(r2p_batch* b, int i, void* h_dst, size_t cap, size_t* len);
int r2p_batch_timings(r2p_batch* b, r2p_timings* t);
/* Autotuner: times candidate settings on the declared batch and keeps the
 * fastest -- DEFLATE workgroup size and filter band height for the launch,
 * then the slice size per size class of the images (filtered streams under
 * 2 MiB, under 8 MiB, larger).  Writes the launch settings to *chosen (its
 * slice_bytes 0: per image, see r2p_batch_job_options). */
int r2p_batch_autotune(r2p_batch* b, r2p_options* chosen, double* best_ms);
/* Sets the options (and clears the autotuner's per-class slice sizes). */
int r2p_batch_options(r2p_batch* b, const r2p_options* opt);
/* The options image i of the declared batch is encoded with, its slice size
 * resolved: the plain encoder with these options produces the same PNG. */
int r2p_batch_job_options(r2p_batch* b, int i, r2p_options* opt);
/* The autotuner's slice size per size class (0: by image size). */
int r2p_batch_class_slices(r2p_batch* b, int* slices3);

/* ======================================================================== */
/* Served form: continuous batching under an HBM budget.                      */
/* ======================================================================== */
typedef struct r2p_server r2p_server;
typedef struct r2p_server_config {
    size_t hbm_budget_bytes;   /* admission budget for device-resident request state */
    int    max_batch_images;   /* images per device batch */
    int    inflight_batches;   /* batches overlapping on the device (streams) */
    r2p_options opt;
} r2p_server_config;
int r2p_server_create(const r2p_server_config* cfg, r2p_server** out);
int r2p_server_destroy(r2p_server* s);
/* Submits a request (host pixels, copied at admission).  Returns an id >= 0
 * in *id.  Non-blocking. */
int r2p_server_submit(r2p_server* s, const r2p_image* img, const void* h_src, int64_t* id);
/* Waits for request `id` (timeout_ms < 0: forever).  Copies the PNG into
 * h_png (cap bytes).  R2P_ERR_STATE when it is not done by the timeout. */
int r2p_server_wait(r2p_server* s, int64_t id, int timeout_ms, void* h_png, size_t cap,
                    size_t* len);
typedef struct r2p_server_stats {
    uint64_t submitted, completed, batches;
    double   mean_batch_images;
    uint64_t peak_admitted_bytes;
} r2p_server_stats;
int r2p_server_stats_get(r2p_server* s, r2p_server_stats* st);

/* Open-loop Poisson load test driven from native code: request i (of n)
 * arrives at arrival_s[i] seconds after start and is image imgs[kind[i]]
 * with pixels h_src[kind[i]].  Fills latency_s[i] (arrival -> PNG in host
 * memory) and png_len[i]; *wall_s receives the makespan.  With expect_png
 * (optional: expected PNG bytes and lengths per image kind), same[i] = 1
 * when request i's PNG equals its kind's expected bytes, compared by
 * verifier threads after the latency is taken. */
int r2p_server_loadtest(r2p_server* s, int n, const double* arrival_s, const int* kind,
                        const r2p_image* imgs, const void* const* h_src, double* latency_s,
                        uint64_t* png_len, double* wall_s, const void* const* expect_png,
                        const size_t* expect_len, uint8_t* same);

#ifdef __cplusplus
}
#endif
#endif
