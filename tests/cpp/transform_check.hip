// Host-side check of the codec's value transforms (sleek_amd/csrc/sleek_device.h,
// the very functions the kernels inline) against the oracle: lossless code
// and inverse, and error-bounded code and inverse (both forms of each,
// integer and float; the inverse also on every 32-bit code) at bounds of
// every class (denormal, tiny, ordinary, the largest allowed), over all 2^32
// float32 bit patterns and over random float64 patterns covering every
// exponent.
// Exit status 0 = no mismatch.
#include <stdint.h>
#include <stdio.h>

#include <atomic>
#include <thread>
#include <vector>

#include "../../oracle/sleek_oracle.h"
#include "../../sleek_amd/csrc/sleek_device.h"

struct Bound32 { int e, t; uint32_t o; };
struct Bound64 { int e, t; uint64_t o; };

int main()
{
  std::vector<Bound32> b32;
  for (float eb : {1e-40f, 1.5e-38f, 1e-30f, 1e-3f, 1.0f, 9.0e30f}) {
    Bound32 b;
    if (sleek_oracle_error_bound_f32(eb, &b.e, &b.t, &b.o) != 0) return 2;
    b32.push_back(b);
  }
  std::vector<Bound64> b64;
  for (double eb : {1e-310, 1e-200, 1e-9, 1.0, 1e290}) {
    Bound64 b;
    if (sleek_oracle_error_bound_f64(eb, &b.e, &b.t, &b.o) != 0) return 2;
    b64.push_back(b);
  }
  std::atomic<unsigned long long> bad{0};
  const unsigned nt = 8;
  std::vector<std::thread> th;
  for (unsigned w = 0; w < nt; w++) {
    th.emplace_back([w, nt, &bad, &b32, &b64] {
      unsigned long long b = 0;
      const uint64_t lo = (1ull << 32) / nt * w, hi = (1ull << 32) / nt * (w + 1);
      for (uint64_t i = lo; i < hi; i++) {
        const uint32_t x = (uint32_t)i;
        const uint32_t c = sleek::lossless_code<uint32_t>(x);
        b += (c != sleek_oracle_fwd_u32(x)) + (sleek::lossless_value<uint32_t>(c) != x);
        for (const Bound32& q : b32) {
          const uint32_t qc = sleek::lossy_code<uint32_t>(x, q.e, q.t, q.o);
          b += (qc != sleek_oracle_quantize_u32(x, q.e, q.t, q.o));
          b += (sleek::lossy_code_f32(x, sleek::lossy_f32_params(q.t, q.o)) != qc);   // the float form
          b += (sleek::lossy_value<uint32_t>(qc, q.t, q.o) != sleek_oracle_dequantize_u32(qc, q.t, q.o));
          // x taken as a code: every code, valid or not, both inverse forms against the oracle
          const uint32_t dx = sleek_oracle_dequantize_u32(x, q.t, q.o);
          b += (sleek::lossy_value<uint32_t>(x, q.t, q.o) != dx);
          b += (sleek::lossy_value_f32(x, sleek::lossy_f32_inv_params(q.t, q.o)) != dx);   // the float form
        }
      }
      uint64_t s = 0x9E3779B97F4A7C15ull * (w + 1);
      for (long k = 0; k < 4000000L; k++) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        uint64_t x = s;
        if (k & 1) x = (x & 0x800FFFFFFFFFFFFFull) | ((uint64_t)((k >> 1) & 0x7ff) << 52);   // every exponent
        const uint64_t c = sleek::lossless_code<uint64_t>(x);
        b += (c != sleek_oracle_fwd_u64(x)) + (sleek::lossless_value<uint64_t>(c) != x);
        for (const Bound64& q : b64) {
          const uint64_t qc = sleek::lossy_code<uint64_t>(x, q.e, q.t, q.o);
          b += (qc != sleek_oracle_quantize_u64(x, q.e, q.t, q.o));
          b += (sleek::lossy_value<uint64_t>(qc, q.t, q.o) != sleek_oracle_dequantize_u64(qc, q.t, q.o));
          if (x != ~0ull)   // any code but 2^64 - 1, whose decoding deviates (see lossy_value)
            b += (sleek::lossy_value<uint64_t>(x, q.t, q.o) != sleek_oracle_dequantize_u64(x, q.t, q.o));
        }
      }
      bad += b;
    });
  }
  for (auto& t : th) t.join();
  printf("transform mismatches: %llu\n", (unsigned long long)bad.load());
  return bad.load() != 0;
}
