/*
 * SLEEK CPU oracle -- TEST INFRASTRUCTURE ONLY (see sleek_oracle.h).
 *
 * Restates the reference CPU codec, include/sleek_common_cpu.h and
 * include/sleek_lossy_cpu.h, in plain C.  Built by oracle/Makefile into
 * oracle/libsleek_oracle.so, loaded by the tests and the CPU-baseline leg of
 * bench.py only.
 */
#include "sleek_oracle.h"

#include <stdlib.h>
#include <string.h>

/* sleek_common_cpu.h:67-71: 12 + 2 B per chunk + a full chunk per chunk */
long long sleek_oracle_upper_bound_size(long long insize)
{
  const long long chunks = (insize + SLEEK_ORACLE_CS - 1) / SLEEK_ORACLE_CS;
  return 3 * 4 + chunks * 2 + chunks * SLEEK_ORACLE_CS;
}

/* sleek_lossy_cpu.h:56-71.  The reference calls exit(-1) when the bound is
 * too large; the oracle returns SLEEK_ORACLE_EEB instead. */
int sleek_oracle_error_bound_f32(float eb, int* eb_e, int* thr_e, uint32_t* offs_e)
{
  uint32_t bits;
  memcpy(&bits, &eb, 4);
  const int e = (int)((bits >> 23) & 0xff);
  const int t = e + 24;
  if (t >= 255) return SLEEK_ORACLE_EEB;
  *eb_e = e;
  *thr_e = t;
  *offs_e = ((uint32_t)t << 23) - (1u << 23);
  return SLEEK_ORACLE_OK;
}

int sleek_oracle_error_bound_f64(double eb, int* eb_e, int* thr_e, uint64_t* offs_e)
{
  uint64_t bits;
  memcpy(&bits, &eb, 8);
  const int e = (int)((bits >> 52) & 0x7ff);
  const int t = e + 53;
  if (t >= 2047) return SLEEK_ORACLE_EEB;
  *eb_e = e;
  *thr_e = t;
  *offs_e = ((uint64_t)t << 52) - (1ull << 52);
  return SLEEK_ORACLE_OK;
}

#define W uint32_t
#define WB 32
#define EXPB 8
#define MANB 23
#define SFX(x) x##_u32
#include "sleek_oracle_impl.h"
#undef W
#undef WB
#undef EXPB
#undef MANB
#undef SFX

#define W uint64_t
#define WB 64
#define EXPB 11
#define MANB 52
#define SFX(x) x##_u64
#include "sleek_oracle_impl.h"
#undef W
#undef WB
#undef EXPB
#undef MANB
#undef SFX

long long sleek_oracle_check_bijection_u32(void)
{
  /* one bit per 32-bit code: 512 MiB bitmap */
  uint64_t* seen = (uint64_t*)calloc((size_t)1 << 26, sizeof(uint64_t));
  if (!seen) return -1;
  long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (long long hi = 0; hi < 65536; hi++) {
    for (uint32_t lo = 0; lo < 65536; lo++) {
      const uint32_t x = ((uint32_t)hi << 16) | lo;
      const uint32_t c = sleek_oracle_fwd_u32(x);
      if (sleek_oracle_inv_u32(c) != x) bad++;
      const uint64_t m = 1ull << (c & 63);
      const uint64_t old = __atomic_fetch_or(&seen[c >> 6], m, __ATOMIC_RELAXED);
      if (old & m) bad++;
    }
  }
  free(seen);
  return bad;
}
