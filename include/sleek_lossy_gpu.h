/*
 * SLEEK for MI355X: drop-in error-bounded GPU API.
 *
 * Same types and overloads as the reference's include/sleek_lossy_gpu.h:
 * sleek_error_bound_info<T> (:47-53), sleek_error_bound(eb) (:56-71, which
 * exits the process for a bound that is too large, as the reference does),
 * and the synchronous compress/decompress overloads (:74-127).  The decoder
 * needs the same eb_info as the encoder: the stream does not record it.
 */
#pragma once

#include "sleek_common_gpu.h"

template <typename T>
struct sleek_error_bound_info
{
  const int eb_e;    // biased exponent of the bound
  const int thr_e;   // biased exponent from which values are kept exactly
  const T offs_e;    // offset of the exactly kept codes
};

template <typename F>
static inline sleek_error_bound_info<typename std::conditional<sizeof(F) == 4, unsigned int, unsigned long long>::type>
sleek_error_bound(const F eb)
{
  static_assert(sizeof(F) == 4 || sizeof(F) == 8, "float or double");
  int e = 0, t = 0;
  int rc;
  if constexpr (sizeof(F) == 4) {
    unsigned int o = 0;
    rc = sleek_gpu_error_bound_f32((float)eb, &e, &t, &o);
    if (rc != SLEEK_OK) {fprintf(stderr, "ERROR: error_bound is too large\n"); exit(-1);}
    return {e, t, o};
  } else {
    unsigned long long o = 0;
    rc = sleek_gpu_error_bound_f64((double)eb, &e, &t, &o);
    if (rc != SLEEK_OK) {fprintf(stderr, "ERROR: error_bound is too large\n"); exit(-1);}
    return {e, t, o};
  }
}

static inline void sleek_compress_gpu(const int blocks, const sleek_error_bound_info<unsigned int> eb_info,
                                      const float* const d_input, const long long num_elements,
                                      byte* const d_encoded, long long* const d_encsize)
{
  sleek_report(sleek_gpu_compress_lossy_f32(blocks, eb_info.eb_e, eb_info.thr_e, eb_info.offs_e, d_input,
                                            num_elements, d_encoded, d_encsize), "sleek_compress_gpu");
}

static inline void sleek_decompress_gpu(const int blocks, const sleek_error_bound_info<unsigned int> eb_info,
                                        const byte* const d_encoded, float* const d_decoded,
                                        long long* const d_decsize)
{
  sleek_report(sleek_gpu_decompress_lossy_f32(blocks, eb_info.eb_e, eb_info.thr_e, eb_info.offs_e, d_encoded,
                                              d_decoded, d_decsize), "sleek_decompress_gpu");
}

static inline void sleek_compress_gpu(const int blocks, const sleek_error_bound_info<unsigned long long> eb_info,
                                      const double* const d_input, const long long num_elements,
                                      byte* const d_encoded, long long* const d_encsize)
{
  sleek_report(sleek_gpu_compress_lossy_f64(blocks, eb_info.eb_e, eb_info.thr_e, eb_info.offs_e, d_input,
                                            num_elements, d_encoded, d_encsize), "sleek_compress_gpu");
}

static inline void sleek_decompress_gpu(const int blocks, const sleek_error_bound_info<unsigned long long> eb_info,
                                        const byte* const d_encoded, double* const d_decoded,
                                        long long* const d_decsize)
{
  sleek_report(sleek_gpu_decompress_lossy_f64(blocks, eb_info.eb_e, eb_info.thr_e, eb_info.offs_e, d_encoded,
                                              d_decoded, d_decsize), "sleek_decompress_gpu");
}
