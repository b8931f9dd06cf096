// Command-line drivers with the reference CLIs' arguments, stdout lines and
// files (lossless_compressors_gpu/sleek_single_compressor.cu:59-150,
// sleek_single_decompressor.cu, lossy_compressors_gpu/
// sleek_single_compressor_lossy.cu:73-171, sleek_single_decompressor_lossy.cu
// and their double-precision twins), built on the drop-in headers so that the
// reference's harnesses (compress_decompress_inputs_*.py, verify_cpu_gpu_*.py)
// can run them unchanged.  One source, eight programs:
//   SLEEK_F64       0: float, 1: double
//   SLEEK_LOSSY     0: lossless, 1: error-bounded
//   SLEEK_DECODE    0: compressor, 1: decompressor
//   ARTIFACT        defined: the timed mode (warm-up call, then one call
//                   timed with HIP events; lossy bounds are relative to the
//                   input's value range, and the lossy decompressor checks
//                   the point-wise bound against the original file)
// Usage (as the reference's):
//   compressor   input_file compressed_file [error_bound]
//   decompressor compressed_file decompressed_file [error_bound [original_file]]
// The harnesses read the last three lines of a compressor's output ("ratio:",
// "encoding time:", "encoding throughput: X Gbytes/s") and the last line of a
// decompressor's ("decoding throughput: X Gbytes/s").
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sleek_lossless_gpu.h"
#include "sleek_lossy_gpu.h"

#ifndef SLEEK_F64
#define SLEEK_F64 0
#endif
#ifndef SLEEK_LOSSY
#define SLEEK_LOSSY 0
#endif
#ifndef SLEEK_DECODE
#define SLEEK_DECODE 0
#endif

using T = std::conditional<SLEEK_F64, double, float>::type;

namespace {

constexpr const char* kPrecision = SLEEK_F64 ? "double" : "single";
constexpr const char* kTag = SLEEK_F64 ? "d64" : "f32";

// A HIP-event timer on the null stream (the drop-in calls run there).
struct Timer {
  hipEvent_t beg{}, end{};
  Timer()
  {
    if (hipEventCreate(&beg) != hipSuccess || hipEventCreate(&end) != hipSuccess) {
      fprintf(stderr, "ERROR: hipEventCreate failed\n");
      exit(-1);
    }
  }
  ~Timer()
  {
    (void)hipEventDestroy(beg);
    (void)hipEventDestroy(end);
  }
  void start() { (void)hipEventRecord(beg, 0); }
  double stop()
  {
    float ms = 0;
    if (hipEventRecord(end, 0) != hipSuccess || hipEventSynchronize(end) != hipSuccess ||
        hipEventElapsedTime(&ms, beg, end) != hipSuccess) {
      fprintf(stderr, "ERROR: event timing failed\n");
      exit(-1);
    }
    return 0.001 * ms;
  }
};

void hip_ok(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    fprintf(stderr, "ERROR: %s: %s\n", what, hipGetErrorString(e));
    exit(-1);
  }
}

// The whole file, or an empty vector with *ok = false.
std::vector<unsigned char> read_file(const char* name, bool* ok)
{
  std::vector<unsigned char> buf;
  *ok = false;
  FILE* const f = fopen(name, "rb");
  if (f == nullptr) return buf;
  if (fseek(f, 0, SEEK_END) == 0) {
    const long long n = ftell(f);
    if (n >= 0 && fseek(f, 0, SEEK_SET) == 0) {
      buf.resize((size_t)n);
      *ok = (n == 0) || fread(buf.data(), 1, (size_t)n, f) == (size_t)n;
    }
  }
  fclose(f);
  return buf;
}

void write_file(const char* name, const void* data, long long n)
{
  FILE* const f = fopen(name, "wb");
  if (f == nullptr || (n > 0 && fwrite(data, 1, (size_t)n, f) != (size_t)n)) {
    fprintf(stderr, "ERROR: cannot write %s\n", name);
    exit(-1);
  }
  fclose(f);
}

#if SLEEK_LOSSY && defined(ARTIFACT)
// An absolute bound from a bound relative to the value range (max - min), in
// the data's precision, as the reference's computeNOAeb prints it.  NaNs are
// skipped (the reference's std::min/std::max fold depends on their position).
T range_bound(const T* x, long long n, T rel)
{
  T lo = INFINITY, hi = -INFINITY;
  for (long long i = 0; i < n; i++) {
    const T v = x[i];
    if (v < lo) lo = v;
    if (v > hi) hi = v;
  }
  printf("min_val: %.10f\n", (double)lo);
  printf("max_val: %.10f\n", (double)hi);
  printf("diff between max and min values: %.10f\n", (double)(hi - lo));
  return (hi - lo) * rel;
}

#if SLEEK_DECODE
// The point-wise bound check of the reference's MAXABS_f32 / MAXABS_d64
// (lossy_compressors_gpu/MAXABS_f32.h:40-95): non-finite values must come
// back unchanged (a NaN as any NaN), finite ones within eb; same report lines.
void maxabs(const T* orig, const T* recon, long long len, T eb)
{
  T max_diff = 0, lo = len ? orig[0] : 0, hi = lo;
  long long viol = 0, nonfinite = 0;
  for (long long i = 0; i < len; i++) {
    const T o = orig[i], r = recon[i];
    if (o < lo) lo = o;
    else if (o > hi) hi = o;
    if (!std::isfinite(o) || !std::isfinite(r)) {
      if (!(std::isnan(o) && std::isnan(r)) && !(o == r)) nonfinite++;
    } else {
      const T d = std::abs(o - r);
      if (d > max_diff) max_diff = d;
      if (r < o - eb || r > o + eb || d > eb) viol++;
    }
  }
  if (viol == 0 && nonfinite == 0) printf("MAXABS_%s verification passed\n", kTag);
  printf("min_val: %.10f\n", (double)lo);
  printf("max_val: %.10f\n", (double)hi);
  printf("diff between max and min values in original input: %.10f\n", (double)(hi - lo));
  printf("num_eb_violations: %lld (%.2f%%) \n", viol, len ? 100.0 * (double)viol / (double)len : 0.0);
  printf("max_diff between original and reconstructed values: %.10f\n", (double)max_diff);
  printf("inf_nan_violations: %lld (%.2f%%) \n", nonfinite, len ? 100.0 * (double)nonfinite / (double)len : 0.0);
  printf("Number of values in original file: %lld\n", len);
}
#endif
#endif

#if SLEEK_LOSSY
// The error-bound argument, parsed in the data's precision (std::stof for
// float, as the reference's drivers do).
T parse_eb(const char* s)
{
  if constexpr (SLEEK_F64) return std::stod(s);
  else return std::stof(s);
}
#endif

void ratio_line(long long num, long long den)
{
  const float cr = (float)((100.0 * (double)num) / (double)den);
  printf("ratio: %6.2f%% %7.3fx\n", cr, 100.0 / cr);
}

#if !SLEEK_DECODE
int compressor(int argc, char* argv[])
{
  const int want = SLEEK_LOSSY ? 4 : 3;
  if (argc != want) {
    printf("USAGE: %s input_file_name compressed_file_name%s\n\n", argv[0], SLEEK_LOSSY ? " error_bound" : "");
    return -1;
  }
  if (sleek_system_checks() != 0) return -1;
  bool ok = false;
  std::vector<unsigned char> in = read_file(argv[1], &ok);
  const long long insize = (long long)in.size();
  if (!ok || insize <= 0) {
    fprintf(stderr, "ERROR: input file too small\n\n");
    return -1;
  }
  printf("original size: %lld bytes\n", insize);
  if (insize % (long long)sizeof(T) != 0) {
    fprintf(stderr, "ERROR: size of input must be a multiple of %ld bytes\n", (long)sizeof(T));
    return -1;
  }
  const long long n = insize / (long long)sizeof(T);
  const long long maxsize = sleek_upper_bound_size(insize);
  T* d_in = nullptr;
  byte* d_enc = nullptr;
  long long* d_size = nullptr;
  hip_ok(hipMalloc((void**)&d_in, insize), "hipMalloc");
  hip_ok(hipMalloc((void**)&d_enc, maxsize), "hipMalloc");
  hip_ok(hipMalloc((void**)&d_size, sizeof(long long)), "hipMalloc");
  hip_ok(hipMemcpy(d_in, in.data(), insize, hipMemcpyHostToDevice), "hipMemcpy");
  const int blocks = sleek_thread_blocks();
#if SLEEK_LOSSY
#ifdef ARTIFACT
  const T eb = range_bound((const T*)in.data(), n, parse_eb(argv[3]));
#else
  const T eb = parse_eb(argv[3]);
#endif
  const auto info = sleek_error_bound(eb);
  auto call = [&] { sleek_compress_gpu(blocks, info, d_in, n, d_enc, d_size); };
#else
  auto call = [&] { sleek_compress_gpu(blocks, d_in, n, d_enc, d_size); };
#endif
#ifdef ARTIFACT
  call();   // warm-up
  hip_ok(hipDeviceSynchronize(), "warm-up");
  sleek_check(__LINE__);
  Timer timer;
  timer.start();
  call();
  const double runtime = timer.stop();
#else
  call();
#endif
  sleek_check(__LINE__);
  long long encsize = 0;
  hip_ok(hipMemcpy(&encsize, d_size, sizeof(long long), hipMemcpyDeviceToHost), "hipMemcpy");
  std::vector<unsigned char> enc((size_t)encsize);
  hip_ok(hipMemcpy(enc.data(), d_enc, encsize, hipMemcpyDeviceToHost), "hipMemcpy");
  printf("encoded size: %lld bytes\n", encsize);
  ratio_line(encsize, insize);
#ifdef ARTIFACT
  printf("encoding time: %.6f s\n", runtime);
  printf("encoding throughput: %8.3f Gbytes/s\n", (double)insize * 1e-9 / runtime);
#endif
  write_file(argv[2], enc.data(), encsize);
  (void)hipFree(d_in);
  (void)hipFree(d_enc);
  (void)hipFree(d_size);
  return 0;
}
#else
int decompressor(int argc, char* argv[])
{
#if SLEEK_LOSSY && defined(ARTIFACT)
  const int want = 5;
  const char* extra = " error_bound original_file_name";
#elif SLEEK_LOSSY
  const int want = 4;
  const char* extra = " error_bound";
#else
  const int want = 3;
  const char* extra = "";
#endif
  if (argc != want) {
    printf("USAGE: %s compressed_file_name decompressed_file_name%s\n\n", argv[0], extra);
    return -1;
  }
  if (sleek_system_checks() != 0) return -1;
  bool ok = false;
  std::vector<unsigned char> enc = read_file(argv[1], &ok);
  const long long insize = (long long)enc.size();
  if (!ok || insize < 8) {
    fprintf(stderr, "ERROR: compressed file too small\n\n");
    return -1;
  }
  long long pre_size = 0;   // the stream header: the original size in bytes
  memcpy(&pre_size, enc.data(), sizeof pre_size);
  if (pre_size < 0) {
    fprintf(stderr, "ERROR: corrupt stream header\n\n");
    return -1;
  }
  printf("encoded size: %lld bytes\n", insize);
  byte* d_enc = nullptr;
  T* d_dec = nullptr;
  long long* d_size = nullptr;
  hip_ok(hipMalloc((void**)&d_enc, insize), "hipMalloc");
  hip_ok(hipMalloc((void**)&d_dec, pre_size > 0 ? pre_size : 16), "hipMalloc");
  hip_ok(hipMalloc((void**)&d_size, sizeof(long long)), "hipMalloc");
  hip_ok(hipMemcpy(d_enc, enc.data(), insize, hipMemcpyHostToDevice), "hipMemcpy");
  const int blocks = sleek_thread_blocks();
#if SLEEK_LOSSY
#ifdef ARTIFACT
  std::vector<unsigned char> orig = read_file(argv[4], &ok);
  if (!ok || orig.empty()) {
    fprintf(stderr, "ERROR: original input file too small\n\n");
    return -1;
  }
  const long long norig = (long long)orig.size() / (long long)sizeof(T);
  const T eb = range_bound((const T*)orig.data(), norig, parse_eb(argv[3]));
#else
  const T eb = parse_eb(argv[3]);
#endif
  const auto info = sleek_error_bound(eb);
  auto call = [&](T* out) { sleek_decompress_gpu(blocks, info, d_enc, out, d_size); };
#else
  auto call = [&](T* out) { sleek_decompress_gpu(blocks, d_enc, out, d_size); };
#endif
#ifdef ARTIFACT
  {
    T* d_dummy = nullptr;   // warm-up into a scratch buffer
    hip_ok(hipMalloc((void**)&d_dummy, pre_size > 0 ? pre_size : 16), "hipMalloc");
    call(d_dummy);
    hip_ok(hipDeviceSynchronize(), "warm-up");
    sleek_check(__LINE__);
    (void)hipFree(d_dummy);
  }
  Timer timer;
  timer.start();
  call(d_dec);
  const double runtime = timer.stop();
#else
  call(d_dec);
#endif
  sleek_check(__LINE__);
  long long decsize = 0;
  hip_ok(hipMemcpy(&decsize, d_size, sizeof(long long), hipMemcpyDeviceToHost), "hipMemcpy");
  std::vector<unsigned char> dec((size_t)decsize);
  hip_ok(hipMemcpy(dec.data(), d_dec, decsize, hipMemcpyDeviceToHost), "hipMemcpy");
  printf("decoded size: %lld bytes\n", decsize);
#if SLEEK_LOSSY && defined(ARTIFACT)
  maxabs((const T*)orig.data(), (const T*)dec.data(), std::min(norig, decsize / (long long)sizeof(T)), eb);
#endif
  ratio_line(insize, decsize);
#ifdef ARTIFACT
  printf("decoding time: %.6f s\n", runtime);
  printf("decoding throughput: %8.3f Gbytes/s\n", (double)decsize * 1e-9 / runtime);
#endif
  write_file(argv[2], dec.data(), decsize);
  (void)hipFree(d_enc);
  (void)hipFree(d_dec);
  (void)hipFree(d_size);
  return 0;
}
#endif

}  // namespace

int main(int argc, char* argv[])
{
  printf("SLEEK for MI355X: %s-precision %s %s\n", kPrecision, SLEEK_LOSSY ? "lossy" : "lossless",
         SLEEK_DECODE ? "decompressor" : "compressor");
  printf("HIP codec on gfx950 behind the reference's command line\n\n");
#if SLEEK_DECODE
  return decompressor(argc, argv);
#else
  return compressor(argc, argv);
#endif
}
