// Batched SLEEK kernels: many arrays (mixed float32/float64, lossless/lossy)
// in one launch per direction.
//
// A tile is one 16 KiB chunk of one array.  The host plan numbers every
// array's tiles consecutively (tile0, tiles) and gives it a workspace block
// with the single-array layout.  batch_prep_kernel clears the blocks and
// writes the tile map (tile -> array), so a launch needs no host work
// proportional to the tile count and records no memset.  The coders then
// walk the tiles in launch order (stride = grid), read the array's
// descriptor and run the same per-chunk code as the single-array kernels;
// every array's bytes are therefore those of a single-array call.  The
// encoder's look-back stays within an array (its own block), and it finishes
// even if a predecessor tile's workgroup is not resident, as in the
// single-array encoder.
#include "sleek_chunk.h"

namespace sleek {

// One workgroup per array: clear its workspace block (16-byte stores) and
// write its tiles into the map.
__global__ __launch_bounds__(256) void batch_prep_kernel(const BatchItem* items, uint32_t* tile_map, uint64_t* ws)
{
  const BatchItem* it = items + blockIdx.x;
  const long long w0 = it->ws_word, half = it->ws_words / 2, t0 = it->tile0, nt = it->tiles;
  uint4* const z = (uint4*)(ws + w0);
  for (long long k = threadIdx.x; k < half; k += 256) z[k] = make_uint4(0u, 0u, 0u, 0u);
  for (long long k = threadIdx.x; k < nt; k += 256) tile_map[t0 + k] = blockIdx.x;
}

__device__ __forceinline__ int item_kind(const BatchItem* it)
{
  return (it->elem_bytes == 8 ? 2 : 0) + (it->lossy ? 1 : 0);
}

// Encoder: one wave per tile (encode_chunk_wave).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3))) void batch_encode_kernel(
    const BatchItem* items, const uint32_t* tile_map, long long tiles, uint64_t* ws, int lb_shift, int lb_window)
{
  __shared__ uint32_t stage[kStageBytes / 4];
  for (long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    const BatchItem* it = items + uni(tile_map[t]);
    EncodeArgs a;
    a.in = it->src;
    a.insize = it->src_bytes;
    a.out = it->dst;
    a.outsize = it->d_size;
    a.ws = ws + it->ws_word;
    a.eb_e = it->eb_e;
    a.thr_e = it->thr_e;
    a.offs = it->offs;
    a.lb_shift = lb_shift;
    a.lb_window = lb_window;
    const long long nchunks = (a.insize + kChunk - 1) / kChunk;
    const long long c = t - it->tile0;
    const int kind = item_kind(it);
    if (nchunks == 0) {
      // an empty array: the 8-byte zero header
      if (threadIdx.x == 0) {
        *(long long*)a.out = 0;
        *a.outsize = 8;
      }
    } else {
      const ScanState st = scan_state(a.ws, nchunks, lb_shift);
      if (kind == 0) encode_chunk_wave<uint32_t, false>(a, st, c, nchunks, stage);
      else if (kind == 1) encode_chunk_wave<uint32_t, true>(a, st, c, nchunks, stage);
      else if (kind == 2) encode_chunk_wave<uint64_t, false>(a, st, c, nchunks, stage);
      else encode_chunk_wave<uint64_t, true>(a, st, c, nchunks, stage);
    }
  }
}

__device__ __forceinline__ DecodeArgs decode_args(const BatchItem* it, uint64_t* ws)
{
  DecodeArgs a;
  a.in = it->src;
  a.inlen = it->src_bytes;
  a.out = it->dst;
  a.cap = it->dst_bytes;
  a.decsize = it->d_size;
  a.ws = ws + it->ws_word;
  a.thr_e = it->thr_e;
  a.offs = it->offs;
  return a;
}

// Offsets: one workgroup per array.
__global__ __launch_bounds__(kScanThreads) void batch_offsets_kernel(const BatchItem* items, uint64_t* ws)
{
  __shared__ long long wsum[kScanThreads / 64];
  scan_offsets(decode_args(items + blockIdx.x, ws), wsum);
}

// Decoder: one workgroup per tile; tiles beyond the stream's chunk count
// (the plan sizes arrays by their capacity) are skipped.
__global__ __launch_bounds__(kThreads) void batch_decode_kernel(const BatchItem* items, const uint32_t* tile_map,
                                                                long long tiles, uint64_t* ws)
{
  __shared__ uint32_t stage[kStageBytes / 4];
  for (long long t = blockIdx.x; t < tiles; t += gridDim.x) {
    const BatchItem* it = items + uni(tile_map[t]);
    const DecodeArgs a = decode_args(it, ws);
    const long long hdr = uni(*(const long long*)a.in);
    const long long c = t - it->tile0;
    const int kind = item_kind(it);
    long long nchunks = 0;
    uint64_t why = 0;
    if (header_ok(a, hdr, &nchunks, &why) && c < nchunks) {   // a bad header is reported by the offsets kernel
      if (kind == 0) decode_chunk<uint32_t, false>(a, hdr, nchunks, c, stage);
      else if (kind == 1) decode_chunk<uint32_t, true>(a, hdr, nchunks, c, stage);
      else if (kind == 2) decode_chunk<uint64_t, false>(a, hdr, nchunks, c, stage);
      else decode_chunk<uint64_t, true>(a, hdr, nchunks, c, stage);
    }
  }
}

// Gathers each array's status word into the batch status words.
__global__ __launch_bounds__(256) void batch_status_kernel(const BatchItem* items, int n, uint64_t* ws)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) ws[kBatchHead + i] = ws[items[i].ws_word + kWsStatus];
}

__global__ __launch_bounds__(256) void zero_kernel(uint64_t* p, long long words)
{
  const long long stride = (long long)gridDim.x * 256;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < words; k += stride) p[k] = 0ull;
}

hipError_t zero_launch(uint64_t* p, long long words, hipStream_t stream)
{
  if (words <= 0) return hipSuccess;
  long long g = (words + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)g), dim3(256), 0, stream, p, words);
  return hipGetLastError();
}

int batch_resident_grid(bool decompress)
{
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const void* fn = decompress ? (const void*)batch_decode_kernel : (const void*)batch_encode_kernel;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, decompress ? kThreads : 64, 0) != hipSuccess) return 0;
  return cus * (per > 0 ? per : 1);
}

hipError_t batch_launch(bool decompress, const BatchItem* items, int n, long long tiles, uint64_t* ws, int grid,
                        int lb_shift, int lb_window, hipStream_t stream)
{
  if (n <= 0) return hipSuccess;
  uint32_t* const map = (uint32_t*)(ws + batch_map_word(n));
  const long long res = batch_resident_grid(decompress);
  long long g = grid > 0 ? grid : res;
  if (!decompress && res > 0 && g > res) g = res;   // the encoder's look-back: no more than the resident set
  if (g <= 0) g = 1024;
  if (g > tiles) g = tiles;
  hipLaunchKernelGGL(batch_prep_kernel, dim3(n), dim3(256), 0, stream, items, map, ws);
  if (decompress) {
    hipLaunchKernelGGL(batch_offsets_kernel, dim3(n), dim3(kScanThreads), 0, stream, items, ws);
    hipLaunchKernelGGL(batch_decode_kernel, dim3((unsigned)g), dim3(kThreads), 0, stream, items, map, tiles, ws);
  } else {
    hipLaunchKernelGGL(batch_encode_kernel, dim3((unsigned)g), dim3(64), 0, stream, items, map, tiles, ws, lb_shift,
                       lb_window);
  }
  return hipGetLastError();
}

hipError_t batch_status_launch(const BatchItem* items, int n, uint64_t* ws, hipStream_t stream)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(batch_status_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, items, n, ws);
  return hipGetLastError();
}

}  // namespace sleek
