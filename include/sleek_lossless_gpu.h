/*
 * SLEEK for MI355X: drop-in lossless GPU API.
 *
 * Same overloads, parameters and synchronous behaviour as the reference's
 * include/sleek_lossless_gpu.h:47-100: device pointers in, device int64 size
 * out, results complete on return.  `blocks` is a persistent grid size
 * (sleek_thread_blocks()); the bytes written never depend on it.
 * Failures (e.g. a malformed stream, where the reference traps the device)
 * are printed and recorded in sleek_last_status().
 */
#pragma once

#include "sleek_common_gpu.h"

static inline void sleek_compress_gpu(const int blocks, const float* const d_input, const long long num_elements,
                                      byte* const d_encoded, long long* const d_encsize)
{
  sleek_report(sleek_gpu_compress_f32(blocks, d_input, num_elements, d_encoded, d_encsize), "sleek_compress_gpu");
}

static inline void sleek_decompress_gpu(const int blocks, const byte* const d_encoded, float* const d_decoded,
                                        long long* const d_decsize)
{
  sleek_report(sleek_gpu_decompress_f32(blocks, d_encoded, d_decoded, d_decsize), "sleek_decompress_gpu");
}

static inline void sleek_compress_gpu(const int blocks, const double* const d_input, const long long num_elements,
                                      byte* const d_encoded, long long* const d_encsize)
{
  sleek_report(sleek_gpu_compress_f64(blocks, d_input, num_elements, d_encoded, d_encsize), "sleek_compress_gpu");
}

static inline void sleek_decompress_gpu(const int blocks, const byte* const d_encoded, double* const d_decoded,
                                        long long* const d_decsize)
{
  sleek_report(sleek_gpu_decompress_f64(blocks, d_encoded, d_decoded, d_decsize), "sleek_decompress_gpu");
}
