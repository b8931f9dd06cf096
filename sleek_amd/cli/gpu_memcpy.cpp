// hipMemcpy host-to-device, device-to-device and device-to-host of a file's
// bytes, each timed with HIP events, with the reference's output lines
// (memcpy.cu:68-139): its harness reads line 8 ("DtD throughput: X Gbytes/s").
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

static void hip_ok(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    fprintf(stderr, "ERROR: %s: %s\n", what, hipGetErrorString(e));
    exit(-1);
  }
}

int main(int argc, char* argv[])
{
  printf("GPU_memcpy (HIP)\n");
  printf("hipMemcpy timed with HIP events\n\n");
  if (argc < 2) {
    printf("USAGE: %s file_name\n\n", argv[0]);
    return -1;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    fprintf(stderr, "ERROR: no HIP device detected\n\n");
    return -1;
  }
  hip_ok(hipSetDevice(0), "hipSetDevice");
  FILE* const fin = fopen(argv[1], "rb");
  if (fin == nullptr) {
    fprintf(stderr, "ERROR: input file too small\n\n");
    return -1;
  }
  fseek(fin, 0, SEEK_END);
  const long long insize = ftell(fin);
  if (insize <= 0) {
    fprintf(stderr, "ERROR: input file too small\n\n");
    fclose(fin);
    return -1;
  }
  std::vector<unsigned char> input((size_t)insize);
  fseek(fin, 0, SEEK_SET);
  const bool got = fread(input.data(), 1, (size_t)insize, fin) == (size_t)insize;
  fclose(fin);
  if (!got) {
    fprintf(stderr, "ERROR: cannot read %s\n\n", argv[1]);
    return -1;
  }
  printf("original size: %lld bytes\n", insize);
  unsigned char *d_a = nullptr, *d_b = nullptr;
  hip_ok(hipMalloc((void**)&d_a, insize), "hipMalloc");
  hip_ok(hipMalloc((void**)&d_b, insize), "hipMalloc");
  hipEvent_t beg, end;
  hip_ok(hipEventCreate(&beg), "hipEventCreate");
  hip_ok(hipEventCreate(&end), "hipEventCreate");
  struct Leg {
    const char* name;
    void* dst;
    const void* src;
    hipMemcpyKind kind;
  } legs[] = {{"HtD", d_a, input.data(), hipMemcpyHostToDevice},
              {"DtD", d_b, d_a, hipMemcpyDeviceToDevice},
              {"DtH", input.data(), d_a, hipMemcpyDeviceToHost}};
  for (const Leg& l : legs) {
    hip_ok(hipEventRecord(beg, 0), "hipEventRecord");
    hip_ok(hipMemcpy(l.dst, l.src, insize, l.kind), "hipMemcpy");
    hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hip_ok(hipEventRecord(end, 0), "hipEventRecord");
    hip_ok(hipEventSynchronize(end), "hipEventSynchronize");
    float ms = 0;
    hip_ok(hipEventElapsedTime(&ms, beg, end), "hipEventElapsedTime");
    const double runtime = 0.001 * ms;
    printf("%s time: %.6f s\n", l.name, runtime);
    printf("%s throughput: %8.3f Gbytes/s\n", l.name, (double)insize * 1e-9 / runtime);
  }
  (void)hipEventDestroy(beg);
  (void)hipEventDestroy(end);
  (void)hipFree(d_a);
  (void)hipFree(d_b);
  return 0;
}
