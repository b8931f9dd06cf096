/*
 * SLEEK CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU codec (include/sleek_common_cpu.h,
 * include/sleek_lossy_cpu.h) used as the checker for the HIP codec.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it;
 * the product path never calls into it.
 *
 * All functions work on byte buffers; the element type is selected by the
 * _u32 (float32) / _u64 (float64) suffix.
 *
 * Return conventions: encoders return the stream length in bytes; decoders
 * return 0 on success, a negative SLEEK_ORACLE_E* code otherwise.
 */
#ifndef SLEEK_ORACLE_H
#define SLEEK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLEEK_ORACLE_CS 16384          /* chunk size in bytes (format constant) */
#define SLEEK_ORACLE_SC 32             /* subchunks per chunk (format constant) */

#define SLEEK_ORACLE_OK 0
#define SLEEK_ORACLE_EARG (-1)         /* bad argument (NULL, negative size, ...) */
#define SLEEK_ORACLE_ECORRUPT (-3)     /* stream structure or trailer mismatch */
#define SLEEK_ORACLE_EEB (-4)          /* error bound too large */
#define SLEEK_ORACLE_ECAP (-5)         /* decoded size exceeds the output capacity */

long long sleek_oracle_upper_bound_size(long long insize);

int sleek_oracle_error_bound_f32(float eb, int* eb_e, int* thr_e, uint32_t* offs_e);
int sleek_oracle_error_bound_f64(double eb, int* eb_e, int* thr_e, uint64_t* offs_e);

/* per-value transforms (exposed for known-answer and bijection tests) */
uint32_t sleek_oracle_fwd_u32(uint32_t v);
uint32_t sleek_oracle_inv_u32(uint32_t v);
uint64_t sleek_oracle_fwd_u64(uint64_t v);
uint64_t sleek_oracle_inv_u64(uint64_t v);
uint32_t sleek_oracle_quantize_u32(uint32_t v, int eb_e, int thr_e, uint32_t offs);
uint32_t sleek_oracle_dequantize_u32(uint32_t c, int thr_e, uint32_t offs);
uint64_t sleek_oracle_quantize_u64(uint64_t v, int eb_e, int thr_e, uint64_t offs);
uint64_t sleek_oracle_dequantize_u64(uint64_t c, int thr_e, uint64_t offs);

/* whole-array codec.  lossy==0 ignores eb_e/thr_e/offs. */
long long sleek_oracle_encode_u32(const uint8_t* in, long long insize, uint8_t* out,
                                  int lossy, int eb_e, int thr_e, uint32_t offs);
long long sleek_oracle_encode_u64(const uint8_t* in, long long insize, uint8_t* out,
                                  int lossy, int eb_e, int thr_e, uint64_t offs);
int sleek_oracle_decode_u32(const uint8_t* in, long long inlen, uint8_t* out, long long outcap,
                            long long* outsize, int lossy, int thr_e, uint32_t offs);
int sleek_oracle_decode_u64(const uint8_t* in, long long inlen, uint8_t* out, long long outcap,
                            long long* outsize, int lossy, int thr_e, uint64_t offs);

/* exhaustive check of the float32 lossless transform over all 2^32 bit
 * patterns: returns the number of patterns x with inv(fwd(x)) != x, plus the
 * number of collisions detected by an injectivity bitmap (0 means bijective). */
long long sleek_oracle_check_bijection_u32(void);

#ifdef __cplusplus
}
#endif

#endif
