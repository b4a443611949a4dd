/*
 * r2p_oracle.c -- CPU oracle for the PNG encode hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see r2p_oracle.h).  Scalar restatements of the
 * reference algorithms, each citing the reference file:line it follows.
 * Build: gcc -O2 -std=gnu11 -ffp-contract=off -fno-fast-math -fPIC -shared
 *        -pthread r2p_oracle.c -lz
 */
#include "r2p_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#define ADLER_MOD 65521u

/* ------------------------------------------------------------------------- */
/* a2: DICOM float window, src/gpu_filter.cu:71-162                          */
/* ------------------------------------------------------------------------- */
static float dicom_value(const uint8_t* in, size_t i, const r2po_dicom_params* p)
{
    float val;
    if (p->bits_allocated == 16) {
        uint16_t raw = (uint16_t)(in[2 * i] | (in[2 * i + 1] << 8)); /* :87-88 */
        int shift = p->high_bit - p->bits_stored + 1;                  /* :93 */
        if (shift > 0) raw = (uint16_t)(raw >> shift);                 /* :94 */
        if (p->bits_stored < 16)                                       /* :97-100 */
            raw = (uint16_t)(raw & (uint16_t)((1u << p->bits_stored) - 1u));
        if (p->pixel_rep == 1) {                                       /* :102-116 */
            int32_t sv;
            if (p->bits_stored == 16) {
                sv = (int32_t)(int16_t)raw;
            } else {
                int half = 1 << (p->bits_stored - 1);
                sv = (int32_t)raw;
                if (sv >= half) sv -= (1 << p->bits_stored);
            }
            if (p->apply_rescale) {
                float m = (float)sv * p->rescale_slope;
                val = m + p->rescale_intercept;
            } else {
                val = (float)sv;
            }
        } else {                                                       /* :117-121 */
            val = (float)raw;
            if (p->apply_rescale) {
                float m = val * p->rescale_slope;
                val = m + p->rescale_intercept;
            }
        }
    } else {                                                           /* :122-130 */
        if (p->pixel_rep == 1) val = (float)(int8_t)in[i];
        else                   val = (float)in[i];
        if (p->apply_rescale) {
            float m = val * p->rescale_slope;
            val = m + p->rescale_intercept;
        }
    }
    return val;
}

static uint32_t dicom_window_value(float val, const r2po_dicom_params* p)
{
    const int out_max = (p->bits_allocated == 16) ? 65535 : 255;       /* :82 */
    if (p->apply_window) {                                             /* :136-151 */
        float wm1 = p->window_width - 1.0f;
        float cen = p->window_center;
        if (wm1 <= 0.0f) {
            val = (val >= cen) ? (float)out_max : 0.0f;
        } else {
            float half = wm1 * 0.5f;
            float lo = (cen - half) - 0.5f;
            float hi = (cen + half) - 0.5f;
            if (val <= lo) {
                val = 0.0f;
            } else if (val > hi) {
                val = (float)out_max;
            } else {
                float t = val - cen;
                t = t + 0.5f;
                t = t / wm1;
                t = t + 0.5f;
                val = t * (float)out_max;
            }
        }
    }
    val = fmaxf(0.f, fminf((float)out_max, val));                     /* :153-154 */
    return (uint32_t)val;
}

void r2po_dicom_window(const uint8_t* in, uint8_t* out, size_t samples,
                       const r2po_dicom_params* p)
{
    for (size_t i = 0; i < samples; i++) {
        uint32_t ov = dicom_window_value(dicom_value(in, i, p), p);
        if (p->bits_allocated == 16) {                                 /* :156-158 */
            out[2 * i]     = (uint8_t)(ov >> 8);
            out[2 * i + 1] = (uint8_t)(ov & 0xFF);
        } else {
            out[i] = (uint8_t)ov;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* a3: Q15 linear window, src/batch_processor.cpp:202-266                    */
/* ------------------------------------------------------------------------- */
int r2po_q15_coeffs(const r2po_dicom_params* p, int16_t* a_fp, int32_t* b_fp)
{
    const int shift = p->high_bit - p->bits_stored + 1;               /* :206-207 */
    if (shift != 0) return 0;
    float A_f, B_f;
    if (p->apply_window && p->window_width > 0.f) {                    /* :213-220 */
        const float lo = p->window_center - p->window_width * 0.5f;
        const float slope = p->apply_rescale ? p->rescale_slope : 1.f;
        const float intercept = p->apply_rescale ? p->rescale_intercept : 0.f;
        A_f = slope * 255.f / p->window_width;
        B_f = (intercept - lo) * 255.f / p->window_width;
    } else if (!p->apply_window) {                                     /* :221-235 */
        const float slope = p->apply_rescale ? p->rescale_slope : 1.f;
        const float intercept = p->apply_rescale ? p->rescale_intercept : 0.f;
        const float range = (float)(1u << p->bits_stored);
        if (p->pixel_rep == 1) {
            const float half = range * 0.5f;
            A_f = slope * 255.f / range;
            B_f = (intercept + half) * 255.f / range;
        } else {
            A_f = slope * 255.f / (range - 1.f);
            B_f = intercept * 255.f / (range - 1.f);
        }
    } else {
        return 0;                                                      /* :236-238 */
    }
    const float A_q15 = A_f * 32768.f;                                 /* :241-242 */
    if (A_q15 < -32767.f || A_q15 > 32767.f) return 0;
    *a_fp = (int16_t)(int)A_q15;                                       /* :244 */
    float bq;                                                          /* :245-248 */
    if (A_f >= 0.f) {
        if (A_q15 == (float)(int)A_q15) bq = B_f * 32768.f;
        else                            bq = B_f * 32768.f + 0.5f;
    } else {
        bq = B_f * 32768.f - 0.5f;
    }
    *b_fp = (int32_t)bq;
    return 1;
}

void r2po_q15_apply(const uint8_t* in_le, uint8_t* out8, size_t pixels,
                    int16_t a_fp, int32_t b_fp, int pixel_rep)
{
    for (size_t i = 0; i < pixels; i++) {
        const uint16_t raw = (uint16_t)(in_le[2 * i] | (in_le[2 * i + 1] << 8));
        const int32_t x = pixel_rep == 1 ? (int32_t)(int16_t)raw       /* :251-256 */
                                         : (int32_t)raw;               /* :258-263 */
        /* int32 multiply-add with two's-complement wrap, arithmetic shift */
        const uint32_t acc = (uint32_t)x * (uint32_t)(int32_t)a_fp + (uint32_t)b_fp;
        const int32_t v = ((int32_t)acc) >> 15;
        out8[i] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
}

/* ------------------------------------------------------------------------- */
/* a4: LUT fallback, src/batch_processor.cpp:100-185                         */
/* ------------------------------------------------------------------------- */
void r2po_lut8_build(uint8_t* lut, const r2po_dicom_params* p, int lut_size)
{
    const int mask = lut_size - 1;
    float norm_scale = 1.f, norm_offset = 0.f;                         /* :104-115 */
    if (!p->apply_window) {
        if (p->pixel_rep == 1) {
            const float half = (float)(lut_size >> 1);
            const float range = (float)lut_size;
            norm_scale = 1.f / range;
            norm_offset = half / range;
        } else {
            norm_scale = 1.f / (float)(lut_size - 1);
        }
    }
    float lo = 0.f, inv_range = 1.f;                                   /* :117-122 */
    if (p->apply_window) {
        lo = p->window_center - p->window_width * 0.5f;
        const float hi = p->window_center + p->window_width * 0.5f;
        inv_range = (hi > lo) ? 1.f / (hi - lo) : 0.f;
    }
    for (int i = 0; i < lut_size; i++) {                               /* :125-146 */
        float val;
        if (p->pixel_rep == 1 && p->bits_stored < 16) {
            const int sign_bit = lut_size >> 1;
            int32_t s = i;
            if (s & sign_bit) s |= ~mask;
            val = (float)s;
        } else if (p->pixel_rep == 1) {
            val = (float)(int16_t)(uint16_t)(unsigned)i;
        } else {
            val = (float)(unsigned)i;
        }
        if (p->apply_rescale) {
            float m = val * p->rescale_slope;
            val = m + p->rescale_intercept;
        }
        float norm;
        if (p->apply_window) {
            norm = (val - lo) * inv_range;
        } else {
            float m = val * norm_scale;
            norm = m + norm_offset;
        }
        float s255 = norm * 255.f;
        const int v = (int)(s255 + 0.5f);
        lut[i] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
}

void r2po_lut8_apply(const uint8_t* in_le, uint8_t* out8, size_t pixels,
                     const uint8_t* lut, int shift, uint16_t mask16)
{
    for (size_t i = 0; i < pixels; i++) {                              /* :176-184 */
        const uint16_t raw = (uint16_t)(in_le[2 * i] | (in_le[2 * i + 1] << 8));
        out8[i] = lut[(raw >> shift) & mask16];
    }
}

void r2po_window8(const uint8_t* in_le, uint8_t* out8, size_t pixels,
                  const r2po_dicom_params* p)
{
    int16_t a; int32_t b;
    if (r2po_q15_coeffs(p, &a, &b)) {                                  /* :402 */
        r2po_q15_apply(in_le, out8, pixels, a, b, p->pixel_rep);
        return;
    }
    const int lut_size = 1 << p->bits_stored;                          /* :406-408, :306-323 */
    uint8_t* lut = (uint8_t*)malloc((size_t)lut_size);
    r2po_lut8_build(lut, p, lut_size);
    r2po_lut8_apply(in_le, out8, pixels, lut, p->high_bit - p->bits_stored + 1,
                    (uint16_t)(lut_size - 1));
    free(lut);
}

/* ------------------------------------------------------------------------- */
/* a7: adaptive filter, src/gpu_filter.cu:183-309                            */
/* ------------------------------------------------------------------------- */
static inline uint8_t paeth(uint8_t a, uint8_t b, uint8_t c)
{
    const int pv = (int)a + (int)b - (int)c;                           /* :212-216 */
    const int pa = abs(pv - (int)a);
    const int pb = abs(pv - (int)b);
    const int pc = abs(pv - (int)c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

static inline uint32_t cost(uint8_t v) { return v <= 128u ? v : 256u - v; } /* :226 */

static inline uint8_t filt(int f, uint8_t x, uint8_t a, uint8_t b, uint8_t c)
{
    switch (f) {
    case 0: return x;
    case 1: return (uint8_t)(x - a);
    case 2: return (uint8_t)(x - b);
    case 3: return (uint8_t)(x - (uint8_t)(((unsigned)a + (unsigned)b) >> 1));
    default: return (uint8_t)(x - paeth(a, b, c));
    }
}

static void row_scores(const uint8_t* cur, const uint8_t* prev, int row_bytes,
                       int bpp, uint32_t s[5])
{
    s[0] = s[1] = s[2] = s[3] = s[4] = 0;
    for (int i = 0; i < row_bytes; i++) {                              /* :206-233 */
        const uint8_t x = cur[i];
        const uint8_t a = i >= bpp ? cur[i - bpp] : 0;
        const uint8_t b = prev[i];
        const uint8_t c = i >= bpp ? prev[i - bpp] : 0;
        for (int f = 0; f < 5; f++) s[f] += cost(filt(f, x, a, b, c));
    }
}

static void filter_one_row(const uint8_t* cur, const uint8_t* prev, uint8_t* dst,
                           int row_bytes, int bpp)
{
    uint32_t s[5];
    row_scores(cur, prev, row_bytes, bpp, s);
    int best = 0; uint32_t bsc = s[0];                                 /* :267-273 */
    for (int f = 1; f < 5; f++) if (s[f] < bsc) { bsc = s[f]; best = f; }
    dst[0] = (uint8_t)best;                                            /* :275 */
    for (int i = 0; i < row_bytes; i++) {                              /* :285-308 */
        const uint8_t a = i >= bpp ? cur[i - bpp] : 0;
        const uint8_t c = i >= bpp ? prev[i - bpp] : 0;
        dst[1 + i] = filt(best, cur[i], a, prev[i], c);
    }
}

void r2po_filter_rows(const uint8_t* in, const uint8_t* prior, uint8_t* out,
                      int rows, int row_bytes, int bpp)
{
    uint8_t* zeros = NULL;
    if (!prior) { zeros = (uint8_t*)calloc((size_t)row_bytes + 1, 1); prior = zeros; }
    for (int r = 0; r < rows; r++) {
        const uint8_t* cur = in + (size_t)r * row_bytes;
        const uint8_t* prev = r ? in + (size_t)(r - 1) * row_bytes : prior; /* :197-199 */
        filter_one_row(cur, prev, out + (size_t)r * (row_bytes + 1), row_bytes, bpp);
    }
    free(zeros);
}

void r2po_filter_scores(const uint8_t* in, const uint8_t* prior, uint32_t* scores,
                        int rows, int row_bytes, int bpp)
{
    uint8_t* zeros = NULL;
    if (!prior) { zeros = (uint8_t*)calloc((size_t)row_bytes + 1, 1); prior = zeros; }
    for (int r = 0; r < rows; r++) {
        const uint8_t* cur = in + (size_t)r * row_bytes;
        const uint8_t* prev = r ? in + (size_t)(r - 1) * row_bytes : prior;
        row_scores(cur, prev, row_bytes, bpp, scores + 5 * (size_t)r);
    }
    free(zeros);
}

int r2po_unfilter_rows(const uint8_t* filtered, const uint8_t* prior, uint8_t* out,
                       int rows, int row_bytes, int bpp)
{
    uint8_t* zeros = NULL;
    if (!prior) { zeros = (uint8_t*)calloc((size_t)row_bytes + 1, 1); prior = zeros; }
    int rc = 0;
    for (int r = 0; r < rows && !rc; r++) {
        const uint8_t* src = filtered + (size_t)r * (row_bytes + 1);
        uint8_t* cur = out + (size_t)r * row_bytes;
        const uint8_t* prev = r ? out + (size_t)(r - 1) * row_bytes : prior;
        const int f = src[0];
        if (f > 4) { rc = -1; break; }
        for (int i = 0; i < row_bytes; i++) {
            const uint8_t a = i >= bpp ? cur[i - bpp] : 0;
            const uint8_t b = prev[i];
            const uint8_t c = i >= bpp ? prev[i - bpp] : 0;
            uint8_t pred = 0;
            switch (f) {
            case 1: pred = a; break;
            case 2: pred = b; break;
            case 3: pred = (uint8_t)(((unsigned)a + b) >> 1); break;
            case 4: pred = paeth(a, b, c); break;
            default: break;
            }
            cur[i] = (uint8_t)(src[1 + i] + pred);
        }
    }
    free(zeros);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* a12: Adler-32 (RFC 1950 section 9).                                        */
/* A = 1 + sum(x_i), B = n + sum((n - i + 1) x_i), both mod 65521.            */
/* ------------------------------------------------------------------------- */
uint32_t r2po_adler32(uint32_t adler, const uint8_t* p, size_t n)
{
    uint32_t a = adler & 0xFFFF, b = adler >> 16;
    while (n) {
        size_t blk = n < 5552 ? n : 5552;  /* largest run with no 32-bit overflow */
        n -= blk;
        while (blk--) { a += *p++; b += a; }
        a %= ADLER_MOD; b %= ADLER_MOD;
    }
    return (b << 16) | a;
}

/* adler(M1 || M2) from adler(M1), adler(M2), |M2|:
 *   A = A1 + A2 - 1,  B = B1 + B2 + |M2| (A1 - 1)   (mod 65521) */
uint32_t r2po_adler32_combine(uint32_t a1, uint32_t a2, uint64_t len2)
{
    const uint64_t A1 = a1 & 0xFFFF, B1 = a1 >> 16, A2 = a2 & 0xFFFF, B2 = a2 >> 16;
    const uint64_t n = len2 % ADLER_MOD;
    const uint64_t A = (A1 + A2 + ADLER_MOD - 1) % ADLER_MOD;
    const uint64_t B = (B1 + B2 + n * ((A1 + ADLER_MOD - 1) % ADLER_MOD)) % ADLER_MOD;
    return (uint32_t)((B << 16) | A);
}

/* ------------------------------------------------------------------------- */
/* a13: CRC-32, reflected polynomial 0xEDB88320, init/xorout 0xFFFFFFFF       */
/* (PNG specification Annex D).                                               */
/* ------------------------------------------------------------------------- */
static uint32_t crc_table[256];
static pthread_once_t crc_once = PTHREAD_ONCE_INIT;

static void crc_init(void)
{
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
}

uint32_t r2po_crc32(uint32_t crc, const uint8_t* p, size_t n)
{
    pthread_once(&crc_once, crc_init);
    uint32_t c = crc ^ 0xFFFFFFFFu;
    while (n--) c = crc_table[(c ^ *p++) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

/* Polynomial product a*b mod P in the reflected representation
 * (bit 31 holds the x^0 coefficient). */
static uint32_t gf2_mulmod(uint32_t a, uint32_t b)
{
    uint32_t prod = 0;
    for (int k = 0; k < 32; k++) {
        if (a & (0x80000000u >> k)) prod ^= b;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;   /* b *= x */
    }
    return prod;
}

/* x^(8 n) mod P by square-and-multiply. */
static uint32_t gf2_x8n(uint64_t n)
{
    uint32_t result = 0x80000000u;          /* 1 */
    uint32_t sq = 0x00800000u;              /* x^8 */
    while (n) {
        if (n & 1) result = gf2_mulmod(result, sq);
        sq = gf2_mulmod(sq, sq);
        n >>= 1;
    }
    return result;
}

/* crc(M1 || M2) = crc(M1) * x^(8|M2|) xor crc(M2)  (the init/xorout terms
 * cancel because both checksums carry them). */
uint32_t r2po_crc32_combine(uint32_t c1, uint32_t c2, uint64_t len2)
{
    return gf2_mulmod(gf2_x8n(len2), c1) ^ c2;
}

/* ------------------------------------------------------------------------- */
/* a15: zlib header FLG by level, src/parallel_deflate.cpp:214-228           */
/* ------------------------------------------------------------------------- */
uint8_t r2po_zlib_flg(int level)
{
    if (level == 1) return 0x01;
    if (level >= 2 && level <= 5) return 0x5E;
    if (level >= 6 && level <= 7) return 0x9C;
    return 0xDA;
}

const char* r2po_zlib_version(void) { return zlibVersion(); }

/* ------------------------------------------------------------------------- */
/* CPU baseline: persistent worker pool, one reusable z_stream per worker     */
/* (ZStreamTLS, src/parallel_deflate.cpp:40-62).                              */
/* ------------------------------------------------------------------------- */
typedef void (*task_fn)(void* arg, int idx);

typedef struct pool_t {
    pthread_mutex_t mu;
    pthread_cond_t  cv_work, cv_done;
    int             nthreads;
    pthread_t*      th;
    task_fn         fn;
    void*           arg;
    int             ntasks, next, pending;
} pool_t;

static pool_t g_pool = { PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER,
                         PTHREAD_COND_INITIALIZER, 0, NULL, NULL, NULL, 0, 0, 0 };

typedef struct zs_tls { z_stream zs; int level; int inited; } zs_tls;
static __thread zs_tls tl_zs;

static z_stream* zs_acquire(int level)
{
    if (tl_zs.inited && tl_zs.level == level) {
        deflateReset(&tl_zs.zs);
    } else {
        if (tl_zs.inited) deflateEnd(&tl_zs.zs);
        memset(&tl_zs.zs, 0, sizeof(tl_zs.zs));
        deflateInit2(&tl_zs.zs, level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY);
        tl_zs.level = level;
        tl_zs.inited = 1;
    }
    return &tl_zs.zs;
}

static void* pool_worker(void* unused)
{
    (void)unused;
    pthread_mutex_lock(&g_pool.mu);
    for (;;) {
        while (g_pool.next >= g_pool.ntasks) pthread_cond_wait(&g_pool.cv_work, &g_pool.mu);
        int idx = g_pool.next++;
        task_fn fn = g_pool.fn; void* arg = g_pool.arg;
        pthread_mutex_unlock(&g_pool.mu);
        fn(arg, idx);
        pthread_mutex_lock(&g_pool.mu);
        if (--g_pool.pending == 0) pthread_cond_broadcast(&g_pool.cv_done);
    }
    return NULL;
}

static void pool_ensure(int n)
{
    pthread_mutex_lock(&g_pool.mu);
    if (g_pool.nthreads < n) {
        g_pool.th = (pthread_t*)realloc(g_pool.th, sizeof(pthread_t) * (size_t)n);
        for (int i = g_pool.nthreads; i < n; i++)
            pthread_create(&g_pool.th[i], NULL, pool_worker, NULL);
        g_pool.nthreads = n;
    }
    pthread_mutex_unlock(&g_pool.mu);
}

/* Runs fn(arg, 0..ntasks-1) on the pool; the caller waits. */
static void pool_run(task_fn fn, void* arg, int ntasks)
{
    if (ntasks <= 0) return;
    pthread_mutex_lock(&g_pool.mu);
    g_pool.fn = fn; g_pool.arg = arg; g_pool.ntasks = ntasks;
    g_pool.next = 0; g_pool.pending = ntasks;
    pthread_cond_broadcast(&g_pool.cv_work);
    while (g_pool.pending) pthread_cond_wait(&g_pool.cv_done, &g_pool.mu);
    pthread_mutex_unlock(&g_pool.mu);
}

typedef struct strip_job {
    /* window stage */
    const uint8_t* in; int width, channels, in_bits, out_bits, mode;
    const r2po_dicom_params* p; int16_t qa; int32_t qb; int use_q15; const uint8_t* lut;
    int lut_shift; uint16_t lut_mask;
    uint8_t* ready; size_t in_row_bytes, row_bytes; int row0, nrows, nparts;
    /* filter stage */
    const uint8_t* prior; uint8_t* filtered; int bpp;
    /* deflate stage */
    int level, is_last_strip, nchunks, last_active; size_t chunk_size, fsize;
    uint8_t** cbuf; size_t* clen; uint32_t* cadler;
} strip_job;

static void window_part(void* arg, int part)
{
    strip_job* j = (strip_job*)arg;
    const int r0 = (int)((long)j->nrows * part / j->nparts);
    const int r1 = (int)((long)j->nrows * (part + 1) / j->nparts);
    for (int r = r0; r < r1; r++) {
        const uint8_t* src = j->in + (size_t)(j->row0 + r) * j->in_row_bytes;
        uint8_t* dst = j->ready + (size_t)r * j->row_bytes;
        const size_t samples = (size_t)j->width * j->channels;
        switch (j->mode) {
        case 0: memcpy(dst, src, j->row_bytes); break;
        case 1:
            for (size_t k = 0; k < samples; k++) { dst[2*k] = src[2*k+1]; dst[2*k+1] = src[2*k]; }
            break;
        case 2: r2po_dicom_window(src, dst, samples, j->p); break;
        default:
            if (j->use_q15) r2po_q15_apply(src, dst, samples, j->qa, j->qb, j->p->pixel_rep);
            else r2po_lut8_apply(src, dst, samples, j->lut, j->lut_shift, j->lut_mask);
            break;
        }
    }
}

static void filter_part(void* arg, int part)
{
    strip_job* j = (strip_job*)arg;
    const int r0 = (int)((long)j->nrows * part / j->nparts);
    const int r1 = (int)((long)j->nrows * (part + 1) / j->nparts);
    for (int r = r0; r < r1; r++) {
        const uint8_t* cur = j->ready + (size_t)r * j->row_bytes;
        const uint8_t* prev = r ? cur - j->row_bytes : j->prior;
        filter_one_row(cur, prev, j->filtered + (size_t)r * (j->row_bytes + 1),
                       (int)j->row_bytes, j->bpp);
    }
}

/* compress_chunk, src/parallel_deflate.cpp:75-112 */
static void deflate_part(void* arg, int i)
{
    strip_job* j = (strip_job*)arg;
    const size_t start = (size_t)i * j->chunk_size;
    if (start >= j->fsize) { j->clen[i] = 0; return; }
    const size_t end = start + j->chunk_size < j->fsize ? start + j->chunk_size : j->fsize;
    const int mode = (j->is_last_strip && i == j->last_active) ? Z_FINISH : Z_FULL_FLUSH;
    z_stream* zs = zs_acquire(j->level);
    const uLong bound = deflateBound(zs, (uLong)(end - start)) + 64;
    zs->next_in = (Bytef*)(j->filtered + start);
    zs->avail_in = (uInt)(end - start);
    zs->next_out = j->cbuf[i];
    zs->avail_out = (uInt)bound;
    deflate(zs, mode);
    j->clen[i] = zs->total_out;
    j->cadler[i] = (uint32_t)adler32(1L, j->filtered + start, (uInt)(end - start));
}

static void put_be32(uint8_t* d, uint32_t v)
{
    d[0] = (uint8_t)(v >> 24); d[1] = (uint8_t)(v >> 16); d[2] = (uint8_t)(v >> 8); d[3] = (uint8_t)v;
}

typedef struct png_out { uint8_t* buf; size_t cap, len; int overflow;
                         uint8_t idat[32768]; size_t idat_pos; } png_out;

static void out_bytes(png_out* o, const void* d, size_t n)
{
    if (o->len + n > o->cap) { o->overflow = 1; return; }
    memcpy(o->buf + o->len, d, n); o->len += n;
}

/* PngWriter::write_chunk, src/png_writer.cpp:113-125 */
static void out_chunk(png_out* o, const char type[4], const uint8_t* data, uint32_t len)
{
    uint8_t b[4];
    put_be32(b, len); out_bytes(o, b, 4);
    out_bytes(o, type, 4);
    if (len) out_bytes(o, data, len);
    uint32_t crc = r2po_crc32(0, (const uint8_t*)type, 4);
    if (len) crc = r2po_crc32(crc, data, len);
    put_be32(b, crc); out_bytes(o, b, 4);
}

/* PngWriter::write_idat_bytes, src/png_writer.cpp:74-86 (32 KiB IDAT chunks) */
static void out_idat(png_out* o, const uint8_t* d, size_t n)
{
    while (n) {
        size_t room = sizeof(o->idat) - o->idat_pos, take = n < room ? n : room;
        memcpy(o->idat + o->idat_pos, d, take);
        o->idat_pos += take; d += take; n -= take;
        if (o->idat_pos == sizeof(o->idat)) { out_chunk(o, "IDAT", o->idat, (uint32_t)o->idat_pos); o->idat_pos = 0; }
    }
}

size_t r2po_filtered_size(int width, int height, int channels, int out_bits)
{
    return (size_t)height * ((size_t)width * channels * (out_bits / 8) + 1);
}

size_t r2po_encode_png(const uint8_t* in, int width, int height, int channels,
                       int in_bits, int mode, const r2po_dicom_params* p,
                       int strip_height, int level, int threads,
                       uint8_t* out, size_t out_cap, uint8_t* filtered_out)
{
    if (width <= 0 || height <= 0 || channels <= 0 || strip_height <= 0 || threads <= 0) return 0;
    if ((mode >= 2) && !p) return 0;
    int out_bits = in_bits;
    if (mode == 2) out_bits = p->bits_allocated;
    if (mode == 3) out_bits = 8;
    const int bpp = channels * (out_bits / 8);
    const size_t in_row = (size_t)width * channels * (in_bits / 8);
    const size_t row_bytes = (size_t)width * bpp;
    pool_ensure(threads);

    strip_job j;
    memset(&j, 0, sizeof(j));
    j.in = in; j.width = width; j.channels = channels; j.in_bits = in_bits;
    j.out_bits = out_bits; j.mode = mode; j.p = p; j.in_row_bytes = in_row;
    j.row_bytes = row_bytes; j.bpp = bpp; j.level = level;
    uint8_t* lut = NULL;
    if (mode == 3) {
        j.use_q15 = r2po_q15_coeffs(p, &j.qa, &j.qb);
        if (!j.use_q15) {
            const int lut_size = 1 << p->bits_stored;
            lut = (uint8_t*)malloc((size_t)lut_size);
            r2po_lut8_build(lut, p, lut_size);
            j.lut = lut; j.lut_shift = p->high_bit - p->bits_stored + 1;
            j.lut_mask = (uint16_t)(lut_size - 1);
        }
    }
    uint8_t* ready = (uint8_t*)malloc(row_bytes * (size_t)strip_height);
    uint8_t* prior = (uint8_t*)calloc(row_bytes + 1, 1);
    uint8_t* filt = (uint8_t*)malloc((row_bytes + 1) * (size_t)strip_height);
    const int maxch = threads;
    uint8_t** cbuf = (uint8_t**)calloc((size_t)maxch, sizeof(uint8_t*));
    size_t* clen = (size_t*)calloc((size_t)maxch, sizeof(size_t));
    uint32_t* cadl = (uint32_t*)calloc((size_t)maxch, sizeof(uint32_t));
    const size_t max_chunk = ((row_bytes + 1) * (size_t)strip_height);
    for (int i = 0; i < maxch; i++) cbuf[i] = (uint8_t*)malloc(compressBound((uLong)max_chunk) + 128);
    j.ready = ready; j.filtered = filt; j.prior = prior;
    j.cbuf = cbuf; j.clen = clen; j.cadler = cadl;

    png_out* o = (png_out*)calloc(1, sizeof(png_out));
    o->buf = out; o->cap = out_cap;
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    out_bytes(o, sig, 8);
    uint8_t ihdr[13];
    put_be32(ihdr, (uint32_t)width); put_be32(ihdr + 4, (uint32_t)height);
    ihdr[8] = (uint8_t)out_bits;
    ihdr[9] = channels == 1 ? 0 : channels == 3 ? 2 : channels == 4 ? 6 : 2; /* pipeline.cpp:80-88 */
    ihdr[10] = ihdr[11] = ihdr[12] = 0;
    out_chunk(o, "IHDR", ihdr, 13);
    const uint8_t zh[2] = {0x78, r2po_zlib_flg(level)};
    out_idat(o, zh, 2);

    uint32_t running = 1; int started = 0;
    const int nstrips = (height + strip_height - 1) / strip_height;
    for (int s = 0; s < nstrips; s++) {
        const int row0 = s * strip_height;
        const int nrows = height - row0 < strip_height ? height - row0 : strip_height;
        j.row0 = row0; j.nrows = nrows;
        j.nparts = threads < nrows ? threads : nrows;
        pool_run(window_part, &j, j.nparts);
        pool_run(filter_part, &j, j.nparts);
        const size_t fsize = (size_t)nrows * (row_bytes + 1);
        if (filtered_out) memcpy(filtered_out + (size_t)row0 * (row_bytes + 1), filt, fsize);
        memcpy(prior, ready + (size_t)(nrows - 1) * row_bytes, row_bytes);
        /* deflate_strip chunking, src/parallel_deflate.cpp:133-144 */
        int nt = threads;
        if ((size_t)nt > fsize / 4096) nt = (int)(fsize / 4096 > 1 ? fsize / 4096 : 1);
        if (nt < 1) nt = 1;
        j.nchunks = nt; j.fsize = fsize; j.is_last_strip = (s == nstrips - 1);
        j.chunk_size = (fsize + (size_t)nt - 1) / (size_t)nt;
        j.last_active = 0;
        for (int i = nt - 1; i >= 0; i--) if ((size_t)i * j.chunk_size < fsize) { j.last_active = i; break; }
        pool_run(deflate_part, &j, nt);
        for (int i = 0; i < nt; i++) {                       /* :172-188 */
            if ((size_t)i * j.chunk_size >= fsize) continue;
            const size_t ilen = ((size_t)(i + 1) * j.chunk_size < fsize ? j.chunk_size : fsize - (size_t)i * j.chunk_size);
            if (!started) { running = cadl[i]; started = 1; }
            else running = r2po_adler32_combine(running, cadl[i], ilen);
            out_idat(o, cbuf[i], clen[i]);
        }
    }
    uint8_t tr[4]; put_be32(tr, running);                    /* zlib_trailer :230-237 */
    out_idat(o, tr, 4);
    if (o->idat_pos) out_chunk(o, "IDAT", o->idat, (uint32_t)o->idat_pos);
    out_chunk(o, "IEND", NULL, 0);
    size_t total = o->overflow ? 0 : o->len;
    for (int i = 0; i < maxch; i++) free(cbuf[i]);
    free(cbuf); free(clen); free(cadl); free(ready); free(prior); free(filt); free(lut); free(o);
    return total;
}
