// On the GPU: the float forms of the float32 error-bounded code and of its
// inverse (lossy_code_f32 and lossy_value_f32, the ones the kernels run)
// against the integer forms (lossy_code<uint32_t> and lossy_value<uint32_t>,
// equal to the oracle's quantize and dequantize on every pattern and code,
// checked on the host by transform_check) over all 2^32 float32 bit patterns
// (each also taken as a code), at bounds of every class.  This pins the float
// forms on the hardware itself (its denormal handling and its conversions).
// Prints one line per bound; exit status 0 = no mismatch.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../sleek_amd/csrc/sleek_device.h"

// Each thread checks kPer consecutive patterns; the grid covers 2^32 exactly.
constexpr int kThreadsPerBlock = 256;
constexpr uint32_t kPer = 256;
constexpr uint32_t kBlocks = (uint32_t)((1ull << 32) / (kThreadsPerBlock * (uint64_t)kPer));   // 65,536

__global__ __launch_bounds__(kThreadsPerBlock) void check_kernel(int eb_e, int thr_e, uint32_t offs,
                                                                 unsigned long long* bad, unsigned* first)
{
  const sleek::LossyF32 p = sleek::lossy_f32_params(thr_e, offs);
  const sleek::LossyF32Inv pi = sleek::lossy_f32_inv_params(thr_e, offs);
  const uint64_t t = (uint64_t)blockIdx.x * kThreadsPerBlock + threadIdx.x;
  const uint32_t x0 = (uint32_t)(t * kPer);
  unsigned long long n = 0;
  for (uint32_t i = 0; i < kPer; i++) {
    const uint32_t x = x0 + i;
    const uint32_t a = sleek::lossy_code_f32(x, p), b = sleek::lossy_code<uint32_t>(x, eb_e, thr_e, offs);
    // x also as a code: the float inverse against the integer one
    const uint32_t c = sleek::lossy_value_f32(x, pi), d = sleek::lossy_value<uint32_t>(x, thr_e, offs);
    if (a != b || c != d) {
      n++;
      atomicMin(first, x);
    }
  }
  if (n) atomicAdd(bad, n);
}

static int check(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return 1;
  }
  return 0;
}

int main()
{
  // error-bound parameters (eb_e, thr_e = eb_e + 24, offs = (thr_e << 23) - (1 << 23)) of bounds of
  // every class: denormal (eb_e 0), the smallest normal, tiny, ordinary, large, the largest allowed
  const int ebe[] = {0, 1, 20, 117, 127, 160, 230};
  unsigned long long* d_bad = nullptr;
  unsigned* d_first = nullptr;
  if (check(hipMalloc(&d_bad, sizeof(unsigned long long)), "hipMalloc") ||
      check(hipMalloc(&d_first, sizeof(unsigned)), "hipMalloc"))
    return 2;
  unsigned long long total = 0;
  for (int e : ebe) {
    const int thr = e + 24;
    const uint32_t offs = ((uint32_t)thr << 23) - (1u << 23);
    const unsigned long long z = 0;
    const unsigned ff = 0xffffffffu;
    if (check(hipMemcpy(d_bad, &z, sizeof(z), hipMemcpyHostToDevice), "copy") ||
        check(hipMemcpy(d_first, &ff, sizeof(ff), hipMemcpyHostToDevice), "copy"))
      return 2;
    hipLaunchKernelGGL(check_kernel, dim3(kBlocks), dim3(kThreadsPerBlock), 0, 0, e, thr, offs, d_bad, d_first);
    if (check(hipGetLastError(), "launch") || check(hipDeviceSynchronize(), "sync")) return 2;
    unsigned long long bad = 0;
    unsigned first = 0;
    if (check(hipMemcpy(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost), "copy") ||
        check(hipMemcpy(&first, d_first, sizeof(first), hipMemcpyDeviceToHost), "copy"))
      return 2;
    printf("eb_e %d thr_e %d: %llu mismatches (first 0x%08x)\n", e, thr, bad, bad ? first : 0u);
    total += bad;
  }
  printf("float-form mismatches: %llu\n", total);
  return total != 0;
}
