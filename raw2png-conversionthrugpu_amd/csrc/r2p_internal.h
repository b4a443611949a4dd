// r2p_internal.h -- descriptors shared by the HIP kernels and their host-side
// launchers.  Not part of the public C-ABI.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <algorithm>

#include "r2p.h"

namespace r2p {

// ---------------------------------------------------------------------------
// Window stage modes (how one input row becomes a PNG-ready row).
// ---------------------------------------------------------------------------
enum WinMode : int32_t {
    WM_COPY        = 0,  // bytes already PNG-ready
    WM_SWAP16      = 1,  // LE 16-bit -> BE 16-bit
    WM_LUT16       = 2,  // (raw16 >> shift) & mask -> u16 table -> BE 16-bit
    WM_LUT8_FROM8  = 3,  // raw8 -> u8 table                 (8-bit DICOM window)
    WM_Q15         = 4,  // (raw16 * A + B) >> 15, clamp     (Q15 16 -> 8)
    WM_LUT8_FROM16 = 5,  // (raw16 >> shift) & mask -> u8 table (LUT 16 -> 8)
    WM_LUT16G      = 6,  // as WM_LUT16, the table read from global memory (16 bits
                         // stored: 65536 entries, 128 KiB, more than the LDS budget)
};

// One image (or strip) for the window+filter kernel.
struct FilterJob {
    const uint8_t* src;        // raw input rows
    uint8_t*       dst;        // filtered rows: rows * (row_bytes + 1)
    const uint8_t* prior;      // PNG-ready row above row 0, or nullptr (= zeros)
    uint8_t*       last_row;   // receives the PNG-ready last row, or nullptr
    const void*    lut;        // device table for the LUT modes
    int64_t        src_pitch;  // bytes between input rows
    int32_t        row_bytes;  // PNG-ready bytes per row
    int32_t        rows;
    int32_t        bpp;        // PNG filter distance (bytes per complete pixel, >= 1)
    int32_t        wmode;
    int32_t        shift;      // right shift of the raw 16-bit sample
    uint32_t       mask;       // mask after the shift (LUT index range - 1)
    int32_t        q15_a, q15_b, pixel_rep;
    int32_t        lut_bytes;  // bytes of the table staged in LDS (0 = none; WM_LUT16G reads
                               // its table from global memory and stages nothing)
    int32_t        band_rows;  // rows per band (workgroup work item) of this job
};

// Parameters of the DICOM table builder (one table per parameter set).
struct LutBuild {
    r2p_dicom_params p;
    int32_t entries;
    int32_t kind;   // 0: a2 float window -> u16 (bits_allocated 16) or u8 (8)
                    // 1: a4 8-bit LUT -> u8
};

// One stream of filtered bytes to compress.  The stream is cut into
// segments of seg_bytes -- the unit of the output format: one fixed-Huffman
// block each, closed by EOB and a sync flush (or BFINAL on the stream's last
// segment) -- and every segment into slices of slice_bytes, the unit of
// work: one workgroup, with the 32 KiB before the slice as history.  A
// slice's tokens go to its own buffer as a bit string (the first slice of a
// segment starts with the 3-bit block header); k_scan and k_assemble join a
// segment's slices at bit granularity and append the segment's trailer.
struct DeflateJob {
    const uint8_t* data;       // base of the addressable bytes
    int64_t        first;      // offset of the first byte to compress
    int64_t        len;        // bytes to compress
    int64_t        dict_lo;    // lowest offset a match may reference
    uint32_t*      slice_out;  // nslice * slice_cap_dw dwords of scratch
    uint32_t*      slice_bits; // token bits per slice
    uint32_t*      slice_adler;// adler32 of each slice's input
    uint32_t*      slice_stats;// optional: 4 u32 per slice (matches, matched bytes, dist sum lo, max len)
    int32_t        seg_bytes;
    int32_t        slice_bytes;// divides seg_bytes
    int32_t        slice_cap_dw;
    int32_t        slice_cap_dw2; // a half slice's buffer (from segment half_from on)
    int32_t        final_block; // the job's last segment sets BFINAL (else sync flush)
    int32_t        bpp;
    int32_t        row_stride; // filtered row size (match candidate one row up), 0 = none
    int32_t        lz77;
    int32_t        nslice;     // slice_geom's count
    int32_t        half_from;  // first segment in half-size slices (slice_geom)
    // dynamic-Huffman blocks (k_deflate<.., true>): instead of bits, each
    // slice writes a 16-bit slot per position (slice_bytes u16 per slice: a
    // literal, 0x8000 | length - 3 where a match starts and its distance - 1
    // in the next slot, kTokNone elsewhere; in units of 4096 positions, the
    // slot pair at positions 16l + 8q + 2j, + 1 in dword 4(l + 256q) + j),
    // the slice's length in tok_count, and adds its symbol counts to its
    // segment's histogram of kHuffSyms u32 (literal/length symbols 0..285,
    // then distance codes 0..29; zero before the launch, and k_huff_tables
    // leaves it zero again)
    uint32_t*      tokens;
    uint32_t*      tok_count;
    uint32_t*      hist;
};

// Where a job's segments go once their sizes are known.
struct AssembleJob {
    const uint32_t* slice_out;
    const uint32_t* slice_bits;
    const uint32_t* slice_adler;
    uint32_t*       slice_boff;  // scan output: bit offset of each slice inside its segment
    uint32_t*       seg_tbits;   // scan output: token bits of each segment (header included)
    uint64_t*       seg_off;     // scan output: nseg + 1 byte offsets in the deflate payload
    uint8_t*        out;         // PNG (or raw) destination
    uint64_t*       out_len;     // total bytes written (device)
    uint32_t*       adler_out;   // final adler32 (device), optional
    uint64_t*       host_res;    // optional: the same two results into mapped host memory ([0] length, [1] adler)
    int64_t         out_cap;
    int64_t         total_in;    // uncompressed bytes (for the adler combine)
    int32_t         nseg;
    int32_t         nslice;
    int32_t         spb;         // slices per segment (seg_bytes / slice_bytes), twice as many from half_from on
    int32_t         half_from;   // first segment in half-size slices (slice_geom)
    int32_t         slice_cap_dw;
    int32_t         slice_cap_dw2;
    int32_t         slice_bytes;
    int32_t         chunk_bytes; // IDAT payload bytes per chunk
    int32_t         mode;        // 0 = raw deflate concatenation, 1 = complete PNG
    int32_t         final_block; // the last segment ends the stream (no sync flush)
    uint32_t        width, height;
    uint8_t         bit_depth, color_type, zflg, pad_;
    int32_t         max_chunks;  // chunks this job may use (grid sizing bound)
    uint32_t        ihdr_crc;    // CRC-32 of the IHDR chunk (type and data), written with chunk 0
    int32_t         eob_in_bits; // the segments' bit strings end with their end-of-block code (dynamic blocks)
    uint64_t*       running;     // optional: running output offset (raw mode, strip appends)
    uint64_t*       base_slot;   // optional: where k_scan records this job's base offset
};

// Kernel launch descriptors: a job table for the batched form, or one job
// passed by value (jobs == nullptr) for the plain form.
struct FilterLaunch {
    const FilterJob*  jobs;
    const uint32_t*   band_start;  // prefix of band counts per job (njobs + 1)
    int32_t           njobs;
    int32_t           band_rows;
    uint32_t          total_bands;
    int32_t           row_cap;     // LDS row capacity (bytes, multiple of 16)
    int32_t           lut_cap;     // LDS table capacity (bytes, multiple of 16)
    const uint32_t*   counts;      // optional device [njobs, band_end, band_begin]: overrides njobs and
                                   // total_bands, and the launch takes the bands [band_begin, band_end)
    uint32_t*         queue;       // k_filter over a job table: a counter, zero at launch, that hands out
                                   // the bands after each workgroup's first; null: every grid-th band
    FilterJob         single;
};

struct DeflateLaunch {
    const DeflateJob* jobs;
    const uint32_t*   slice_start; // prefix of slice counts per job (njobs + 1)
    int32_t           njobs;
    uint32_t          total_slices;
    const uint32_t*   counts;      // optional device [njobs, total]
    DeflateJob        single;
    int32_t           dyn;         // the jobs' dynamic-Huffman form (k_deflate<.., true, ..>)
    int32_t           dense;       // hashed search at every 2nd position (r2p_options.lz77 = 2), else every 4th
    uint32_t*         queue;       // persistent workgroups: a counter, zero at launch, that hands out
                                   // the slices after each workgroup's first; null: every grid-th slice
};

// Dynamic-Huffman blocks of one job (after k_deflate<.., true>): per
// segment the code tables and the block header, then per slice the bits.
struct HuffJob {
    const uint32_t* tokens;      // as DeflateJob
    const uint32_t* tok_count;
    uint32_t*       hist;        // per segment (read, then cleared)
    uint32_t*       codes;       // nseg * kHuffTab: bit-reversed code (extra bits in place) | bit count << 24
    uint32_t*       hdr;         // nseg * kHuffHdrWords: the block header, BFINAL/BTYPE first
    uint32_t*       hdr_bits;    // nseg
    uint32_t*       slice_out;   // as DeflateJob: the slices' bit strings
    uint32_t*       slice_bits;
    int32_t         slice_cap_dw;
    int32_t         slice_cap_dw2;
    int32_t         slice_bytes;
    int32_t         spb;         // slices per segment (twice as many from half_from on)
    int32_t         half_from;
    int32_t         nslice;
    int32_t         nseg;
    int32_t         final_block;
};

// The dynamic-Huffman launches over one job or a table of them: workgroups
// take global segments (k_huff_tables) or slices (k_huff_emit) in order.
struct HuffLaunch {
    const HuffJob*  jobs;        // null: the single job
    const uint32_t* seg_start;   // prefix of segment counts per job (njobs + 1)
    const uint32_t* slice_start; // prefix of slice counts per job (njobs + 1)
    const uint32_t* counts;      // optional device [njobs, total segments, total slices]
    int32_t         njobs;
    uint32_t        total_segs, total_slices;
    HuffJob         single;
};

struct AssembleLaunch {
    const AssembleJob* jobs;
    const uint32_t*    chunk_start; // prefix of max_chunks per job (njobs + 1)
    int32_t            njobs;
    uint32_t           total_chunks;
    const uint32_t*    counts;      // optional device [njobs, total]
    AssembleJob        single;
};

constexpr int kFilterThreads = 256;      // k_filter (one row per wave) and k_filter_tiled
constexpr int kFilterBandRowBytes = 6144; // rows up to this go to k_filter (at most 6 items per lane)
constexpr int kFilterRowThreads = 512;   // k_filter_split: 8 waves; band rows x waves per row = 8
constexpr int kFilterMaxK = 8;           // k_filter_split: most 16-byte items of a row per lane
constexpr int kFilterPrefetch = 2;       // k_filter_split: raw 16-byte vectors of the next band per thread
constexpr int kFilterSplitWgPerCu = 2;   // k_filter_split: persistent workgroups per CU
// Wider rows than k_filter takes (one row per wave, K <= 6): k_filter_split.
inline bool filter_split(int row_cap) { return row_cap > kFilterBandRowBytes; }
// k_filter_split's items per lane for rows of `items` 16-byte items in bands
// of band_rows rows (each row spread over 8 / band_rows waves).
inline int filter_items_per_lane(int items, int band_rows)
{
    const int lanes = 64 * (kFilterRowThreads / 64 / band_rows);
    return (items + lanes - 1) / lanes;
}
constexpr int kDeflateThreads = 256;
constexpr int kXcdRun = 16;             // consecutive DEFLATE slices dealt to one XCD
constexpr int kAssembleThreads = 1024;
constexpr int kAsmScanSlices = 2048;  // a single job of up to this many slices and
constexpr int kAsmScanSegs = 1024;    // segments is scanned by k_assemble itself
constexpr int kCrcShiftLevels = 29;       // CRC shift tables for 64 * 2^k zero bytes, k < 29
constexpr int kMaxRowBytesLds = 32768;   // widest PNG-ready row k_filter stages in LDS whole;
                                         // wider rows go to k_filter_tiled
constexpr int kTileBytes = 16384;        // k_filter_tiled: PNG-ready bytes of a row per tile
constexpr int kFilterLdsMax = 160 * 1024 - 256;  // LDS one filter workgroup may use
constexpr int kWindow = 32768;           // DEFLATE window
constexpr int kHuffSyms = 286 + 30;      // literal/length symbols, then distance codes
constexpr int kHuffTab = kHuffSyms + 256; // the emit table: the symbols, then match lengths 3..258
constexpr int kHuffHdrWords = 96;        // a dynamic block header: at most 3 + 14 + 57 + 316 * 7 bits
// Every buffer k_deflate reads holds at least this many readable bytes after
// the last data byte, so its 16-byte loads need no tail guard.
constexpr int kTailSlack = 64;
constexpr int kChecksumBlock = 16384;    // bytes per Adler-32 partial
constexpr int kCrcBlock = 131072;        // bytes per CRC-32 partial

// Worst-case token dwords of one slice of `slice_bytes` input bytes (9 bits
// per literal, the 3-bit block header), plus the slack k_assemble's 5-dword
// reads may touch past the last written word.
inline int slice_cap_dwords(int slice_bytes) { return (int)(((int64_t)slice_bytes * 9 + 3 + 31) / 32) + 8; }

// Slice layout of a DEFLATE stream: segments of seg_bytes, each cut into
// spb slices of slice_bytes, except the segments from half_from on, each cut
// into 2 spb slices of slice_bytes / 2.  One workgroup deflates a slice;
// slices stored back to back, every one at a full slice's buffer stride.
__host__ __device__ inline int lay_first(int s, int spb, int half_from)
{
    return s <= half_from ? s * spb : half_from * spb + (s - half_from) * 2 * spb;
}
__host__ __device__ inline int lay_seg(int k, int spb, int half_from)
{
    const int k0 = half_from * spb;
    return k < k0 ? k / spb : half_from + (k - k0) / (2 * spb);
}
__host__ __device__ inline int64_t lay_start(int k, int spb, int half_from, int slice_bytes)
{
    const int k0 = half_from * spb;
    return k < k0 ? (int64_t)k * slice_bytes : (int64_t)k0 * slice_bytes + (int64_t)(k - k0) * (slice_bytes / 2);
}
__host__ __device__ inline int lay_bytes(int k, int spb, int half_from, int slice_bytes)
{
    return k < half_from * spb ? slice_bytes : slice_bytes / 2;
}
// Dword offset of slice k's bit buffer: cap dwords per full slice, cap2 per
// half slice.
__host__ __device__ inline int64_t lay_buf(int k, int spb, int half_from, int cap, int cap2)
{
    const int k0 = half_from * spb;
    return k < k0 ? (int64_t)k * cap : (int64_t)k0 * cap + (int64_t)(k - k0) * cap2;
}

// Segment / slice counts of a stream of `len` bytes (an empty stream still
// has one segment and one slice: the block header and its end).  With
// halve_tail (fixed-Huffman streams whose slice size is left to the encoder), slices of
// 8 KiB or more fill whole rounds of kSliceRound workgroups (MI355X's
// resident DEFLATE workgroups, 4 on each of 256 CUs) at full size, and the
// segments past those rounds are cut into half-size slices: the last,
// partial round is made of shorter work items that spread over the CUs as
// they free up.  (A function of the stream's size and the options only.)
constexpr int kSliceRound = 1024;
struct SliceGeom { int nseg, nslice, spb, half_from; };
inline SliceGeom slice_geom(int64_t len, int seg_bytes, int slice_bytes, bool halve_tail = false)
{
    SliceGeom g;
    g.nseg = len > 0 ? (int)((len + seg_bytes - 1) / seg_bytes) : 1;
    g.spb = seg_bytes / slice_bytes;
    const int64_t nfull = len > 0 ? (len + slice_bytes - 1) / slice_bytes : 1;
    const int64_t rounds = nfull / kSliceRound;
    g.half_from = g.nseg;
    if (halve_tail && slice_bytes >= 8192 && rounds >= 1 && nfull > rounds * kSliceRound)
        g.half_from = std::min<int64_t>(g.nseg, rounds * kSliceRound / g.spb);
    const int64_t full_len = std::min<int64_t>(len, (int64_t)g.half_from * seg_bytes);
    const int64_t rest = len - full_len;
    g.nslice = len > 0 ? (int)((full_len + slice_bytes - 1) / slice_bytes + (rest + slice_bytes / 2 - 1) / (slice_bytes / 2))
                       : 1;
    return g;
}
// Most slices slice_geom gives a stream of len bytes, and a bound on the
// bit buffers' dwords with full slices of cap dwords (two half slices of a
// full one take at most cap + 20 dwords).
inline size_t slice_count_bound(size_t len, int slice_bytes) { return 2 * ((len + slice_bytes - 1) / slice_bytes) + 1; }
inline size_t slice_dword_bound(size_t len, int slice_bytes, int cap)
{
    return ((len + slice_bytes - 1) / slice_bytes + 1) * ((size_t)cap + 20);
}

}  // namespace r2p
