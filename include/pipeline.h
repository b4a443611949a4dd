// pipeline.h -- image -> PNG entry points.
// Drop-in for include/pipeline.h:8-122 of the reference for the PNG paths.
// With use_gpu_deflate every stage (window, filter, DEFLATE, checksums, IDAT
// assembly) runs on the GPU with the image resident in HBM; otherwise strips
// are filtered on the GPU and deflated by the CPU zlib backend.
#pragma once
#include <cstdint>

struct ImageSource;

enum class ExportFormat { PNG, BMP };

struct PipelineConfig {
    int  strip_height    = 0;      // 0 = auto: 1024 for GPU deflate, else 256 (main.cpp:165-170)
    int  deflate_threads = 6;
    int  deflate_level   = 3;      // zlib level; only the header FLEVEL on the GPU path
    int  frame           = -1;
    bool verbose         = false;
    int  gpu_streams     = 0;
    bool use_gpu_deflate = false;
    bool use_gpu_lz77    = false;
    bool gpu_lz77_debug  = false;
    ExportFormat format  = ExportFormat::PNG;
};

// Uncompressed little-endian TIFF (strip layout) and Explicit/Implicit VR
// little-endian DICOM Part 10 are read natively; other encodings return false.
bool encode_tiff_to_png(const char* input_path, const char* output_path, const PipelineConfig& cfg);
bool encode_raw_to_png(const char* input_path, const char* output_path, const PipelineConfig& cfg);
bool encode_dicom_to_png(const char* input_path, const char* output_path, const PipelineConfig& cfg);
bool encode_dicom_all_frames_to_png(const char* input_path, const char* output_dir,
                                    const PipelineConfig& cfg);
bool encode_source_to_png(ImageSource& src, const char* output_path, const PipelineConfig& cfg);

void pipeline_reset_gpu_batch_stats();
void pipeline_record_file_times(long long dicom_ms, long long total_ms);
void pipeline_print_gpu_batch_summary(int total_files, int succeeded, double total_wall_s,
                                      long long sum_file_batch_ms, int num_workers);
