// Host-side check of the codec's value transforms (sleek_amd/csrc/sleek_device.h,
// the very functions the kernels inline) against the oracle: lossless code
// and inverse over all 2^32 float32 bit patterns, and over random float64
// patterns covering every exponent.  Exit status 0 = no mismatch.
#include <stdint.h>
#include <stdio.h>

#include <atomic>
#include <thread>
#include <vector>

#include "../../oracle/sleek_oracle.h"
#include "../../sleek_amd/csrc/sleek_device.h"

int main()
{
  std::atomic<unsigned long long> bad{0};
  const unsigned nt = 8;
  std::vector<std::thread> th;
  for (unsigned w = 0; w < nt; w++) {
    th.emplace_back([w, nt, &bad] {
      unsigned long long b = 0;
      const uint64_t lo = (1ull << 32) / nt * w, hi = (1ull << 32) / nt * (w + 1);
      for (uint64_t i = lo; i < hi; i++) {
        const uint32_t x = (uint32_t)i;
        const uint32_t c = sleek::lossless_code<uint32_t>(x);
        b += (c != sleek_oracle_fwd_u32(x)) + (sleek::lossless_value<uint32_t>(c) != x);
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
      }
      bad += b;
    });
  }
  for (auto& t : th) t.join();
  printf("transform mismatches: %llu\n", (unsigned long long)bad.load());
  return bad.load() != 0;
}
