// Drop-in check: a caller written against the reference's GPU API
// (sleek_lossless_gpu.h / sleek_lossy_gpu.h overloads, sleek_thread_blocks,
// sleek_upper_bound_size, sleek_error_bound) compiled with plain g++ and
// linked against libsleek_amd.so.  Allocation goes through the HIP runtime,
// as a reference caller's cudaMalloc would.
//
// usage: dropin_check <f32|f64> <n> <eb or 0 for lossless> <stream_out> <decoded_out>
// The input is i * 0.001 + sin-free polynomial noise, so the test can rebuild it.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sleek_lossless_gpu.h"
#include "sleek_lossy_gpu.h"

template <typename F>
static int run(long long n, double eb, const char* sout, const char* dout)
{
  std::vector<F> h(n);
  for (long long i = 0; i < n; i++) h[i] = (F)(i * 0.001 + (double)((i * 2654435761ull) % 1000) * 1e-6);
  if (sleek_system_checks() != 0) return 2;
  const long long insize = n * (long long)sizeof(F);
  const long long maxsize = sleek_upper_bound_size(insize);
  F* d_in = nullptr;
  F* d_out = nullptr;
  byte* d_enc = nullptr;
  long long* d_sz = nullptr;
  if (hipMalloc((void**)&d_in, insize > 0 ? insize : 16) != hipSuccess) return 3;
  if (hipMalloc((void**)&d_out, insize > 0 ? insize : 16) != hipSuccess) return 3;
  if (hipMalloc((void**)&d_enc, maxsize > 8 ? maxsize : 8) != hipSuccess) return 3;
  if (hipMalloc((void**)&d_sz, 8) != hipSuccess) return 3;
  hipMemcpy(d_in, h.data(), insize, hipMemcpyHostToDevice);
  const int blocks = sleek_thread_blocks();
  if (eb > 0) {
    const auto info = sleek_error_bound((F)eb);
    sleek_compress_gpu(blocks, info, d_in, n, d_enc, d_sz);
  } else {
    sleek_compress_gpu(blocks, d_in, n, d_enc, d_sz);
  }
  sleek_check(__LINE__);
  long long encsize = 0;
  hipMemcpy(&encsize, d_sz, 8, hipMemcpyDeviceToHost);
  std::vector<unsigned char> s(encsize);
  hipMemcpy(s.data(), d_enc, encsize, hipMemcpyDeviceToHost);
  if (eb > 0) {
    const auto info = sleek_error_bound((F)eb);
    sleek_decompress_gpu(blocks, info, d_enc, d_out, d_sz);
  } else {
    sleek_decompress_gpu(blocks, d_enc, d_out, d_sz);
  }
  sleek_check(__LINE__);
  long long decsize = 0;
  hipMemcpy(&decsize, d_sz, 8, hipMemcpyDeviceToHost);
  std::vector<F> d(n);
  hipMemcpy(d.data(), d_out, insize, hipMemcpyDeviceToHost);
  FILE* f = fopen(sout, "wb");
  fwrite(s.data(), 1, encsize, f);
  fclose(f);
  f = fopen(dout, "wb");
  fwrite(d.data(), sizeof(F), n, f);
  fclose(f);
  printf("original size: %lld bytes\nencoded size: %lld bytes\ndecoded size: %lld bytes\n", insize, encsize, decsize);
  hipFree(d_in);
  hipFree(d_out);
  hipFree(d_enc);
  hipFree(d_sz);
  return decsize == insize ? 0 : 4;
}

int main(int argc, char* argv[])
{
  if (argc != 6) {
    printf("USAGE: %s f32|f64 num_elements error_bound(0=lossless) stream_file decoded_file\n", argv[0]);
    return 1;
  }
  const long long n = atoll(argv[2]);
  const double eb = atof(argv[3]);
  if (!strcmp(argv[1], "f32")) return run<float>(n, eb, argv[4], argv[5]);
  return run<double>(n, eb, argv[4], argv[5]);
}
