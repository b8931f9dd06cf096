// TEST INFRASTRUCTURE ONLY.
//
// Exposes the reference CPU codec (include/sleek_lossless_cpu.h:46-73 and
// include/sleek_lossy_cpu.h:56-101 of the reference tree, compiled where it
// lies) through a C ABI so the tests can hold the oracle against it and the
// benchmark can time it as the CPU baseline.  Built by oracle/Makefile into
// oracle/_ref/libsleek_ref.so, which is never committed.
#include "sleek_lossless_cpu.h"
#include "sleek_lossy_cpu.h"

extern "C" {

long long sleek_ref_compress_f32(const float* in, long long n, unsigned char* out)
{
  long long sz = 0;
  sleek_compress_cpu(in, n, out, sz);
  return sz;
}

long long sleek_ref_compress_f64(const double* in, long long n, unsigned char* out)
{
  long long sz = 0;
  sleek_compress_cpu(in, n, out, sz);
  return sz;
}

long long sleek_ref_decompress_f32(const unsigned char* in, float* out)
{
  long long sz = 0;
  sleek_decompress_cpu(in, out, sz);
  return sz;
}

long long sleek_ref_decompress_f64(const unsigned char* in, double* out)
{
  long long sz = 0;
  sleek_decompress_cpu(in, out, sz);
  return sz;
}

// The reference exits the process for too-large bounds (sleek_lossy_cpu.h:68);
// callers must only pass bounds the oracle accepts.
long long sleek_ref_compress_lossy_f32(float eb, const float* in, long long n, unsigned char* out)
{
  long long sz = 0;
  sleek_compress_cpu(sleek_error_bound(eb), in, n, out, sz);
  return sz;
}

long long sleek_ref_compress_lossy_f64(double eb, const double* in, long long n, unsigned char* out)
{
  long long sz = 0;
  sleek_compress_cpu(sleek_error_bound(eb), in, n, out, sz);
  return sz;
}

long long sleek_ref_decompress_lossy_f32(float eb, const unsigned char* in, float* out)
{
  long long sz = 0;
  sleek_decompress_cpu(sleek_error_bound(eb), in, out, sz);
  return sz;
}

long long sleek_ref_decompress_lossy_f64(double eb, const unsigned char* in, double* out)
{
  long long sz = 0;
  sleek_decompress_cpu(sleek_error_bound(eb), in, out, sz);
  return sz;
}

void sleek_ref_error_bound_f32(float eb, int* eb_e, int* thr_e, unsigned int* offs)
{
  const auto i = sleek_error_bound(eb);
  *eb_e = i.eb_e; *thr_e = i.thr_e; *offs = i.offs_e;
}

void sleek_ref_error_bound_f64(double eb, int* eb_e, int* thr_e, unsigned long long* offs)
{
  const auto i = sleek_error_bound(eb);
  *eb_e = i.eb_e; *thr_e = i.thr_e; *offs = i.offs_e;
}

}
