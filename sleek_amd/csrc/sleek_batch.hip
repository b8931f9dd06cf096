// Batched SLEEK kernels: many arrays (mixed float32/float64, lossless/lossy)
// in one launch per direction.
//
// A tile is one 16 KiB chunk of one array (an empty array has none).  The
// host plan numbers every array's tiles consecutively (tile0, tiles) and
// gives it a workspace block with the single-array layout.  batch_prep_kernel clears the blocks and
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

// One workgroup per array: clear its workspace block (16-byte stores),
// write its tiles into the map, and give an empty array to be compressed
// its 8-byte zero header (it has no tile).  Workgroup 0 also clears the tile
// tickets when there are any.
__global__ __launch_bounds__(256) void batch_prep_kernel(const BatchItem* items, uint32_t* tile_map, uint64_t* ws,
                                                         int compress, uint64_t* tickets)
{
  if (tickets && blockIdx.x == 0)
    for (int k = threadIdx.x; k < kTicketWords / 2; k += 256) ((uint4*)tickets)[k] = make_uint4(0u, 0u, 0u, 0u);
  const BatchItem* it = items + blockIdx.x;
  const long long w0 = it->ws_word, half = it->ws_words / 2, t0 = it->tile0, nt = it->tiles;
  uint4* const z = (uint4*)(ws + w0);
  for (long long k = threadIdx.x; k < half; k += 256) z[k] = make_uint4(0u, 0u, 0u, 0u);
  for (long long k = threadIdx.x; k < nt; k += 256) tile_map[t0 + k] = blockIdx.x;
  if (compress && it->src_bytes == 0 && threadIdx.x == 0) {
    *(long long*)it->dst = 0;
    *it->d_size = 8;
  }
}

__device__ __forceinline__ int item_kind(const BatchItem* it)
{
  return (it->elem_bytes == 8 ? 2 : 0) + (it->lossy ? 1 : 0);
}

// A tile of the batched encoder: its array's arguments and scan state.
struct EncTile {
  EncodeArgs a;
  ScanState st;
  long long c, nchunks;
  int kind;
};

// Look-back settings shared by every tile of a batched launch.
struct LookCfg {
  int lb_shift, lb_window;
  unsigned long long* help;
  int ticket_shift;   // tile tickets: 2^ticket_shift sets (ticket_shift_for(grid))
};

__device__ __forceinline__ EncTile enc_tile(const BatchItem* __restrict__ items, const uint32_t* __restrict__ tile_map,
                                            long long t, uint64_t* ws, const LookCfg& lc)
{
  const BatchItem* it = items + uni(tile_map[t]);
  EncTile e;
  e.a.in = it->src;
  e.a.insize = it->src_bytes;
  e.a.out = it->dst;
  e.a.outsize = it->d_size;
  e.a.ws = ws + it->ws_word;
  e.a.eb_e = it->eb_e;
  e.a.thr_e = it->thr_e;
  e.a.offs = it->offs;
  e.a.lb_shift = lc.lb_shift;
  e.a.lb_window = lc.lb_window;
  e.a.tickets = nullptr;
  e.a.help = lc.help;
  e.nchunks = (e.a.insize + kChunk - 1) / kChunk;
  e.c = t - it->tile0;
  e.kind = item_kind(it);
  e.st = scan_state(e.a.ws, e.nchunks, lc.lb_shift);
  return e;
}

__device__ __forceinline__ ChunkShape enc_shape(const EncTile& e, uint2 slot[kSub])
{
  // subchunk by subchunk: the four element kinds' code paths leave no
  // registers for the butterfly's 32 inputs
  if (e.kind == 0) return shape_codes<uint32_t, false, false>(e.a, e.c, slot);
  if (e.kind == 1) return shape_codes<uint32_t, true, false>(e.a, e.c, slot);
  if (e.kind == 2) return shape_codes<uint64_t, false, false>(e.a, e.c, slot);
  return shape_codes<uint64_t, true, false>(e.a, e.c, slot);
}

__device__ __forceinline__ void enc_pack(const EncTile& e, const uint2 slot[kSub], const ChunkShape& sh, uint32_t* stage)
{
  if (e.kind == 0) pack_codes<uint32_t, false>(e.a, e.c, slot, sh, stage);
  else if (e.kind == 1) pack_codes<uint32_t, true>(e.a, e.c, slot, sh, stage);
  else if (e.kind == 2) pack_codes<uint64_t, false>(e.a, e.c, slot, sh, stage);
  else pack_codes<uint64_t, true>(e.a, e.c, slot, sh, stage);
}

__device__ __forceinline__ uint64_t enc_look_back(const EncTile& e)
{
  // the element type only matters to a look-back that sizes a missing chunk itself
  if (e.c == 0) return 0;
  if (e.kind == 0) return look_back<uint32_t, false>(e.a, e.st, e.c, e.nchunks);
  if (e.kind == 1) return look_back<uint32_t, true>(e.a, e.st, e.c, e.nchunks);
  if (e.kind == 2) return look_back<uint64_t, false>(e.a, e.st, e.c, e.nchunks);
  return look_back<uint64_t, true>(e.a, e.st, e.c, e.nchunks);
}

__device__ __forceinline__ void enc_finish(const EncTile& e, int csize, const uint32_t* stage)
{
  finish_chunk(e.a, e.st, e.c, e.nchunks, csize, stage, enc_look_back(e));
}

// ---- two waves per tile (the single-array two-wave encoder's layout) ----
__device__ __forceinline__ void enc_shape_half(const EncTile& e, int h, uint2 slot[kHalfSub], uint32_t lh[4])
{
  if (e.kind == 0) shape_half<uint32_t, false>(e.a, e.c, h, slot, lh);
  else if (e.kind == 1) shape_half<uint32_t, true>(e.a, e.c, h, slot, lh);
  else if (e.kind == 2) shape_half<uint64_t, false>(e.a, e.c, h, slot, lh);
  else shape_half<uint64_t, true>(e.a, e.c, h, slot, lh);
}

__device__ __forceinline__ ChunkShape enc_shape_of(const EncTile& e, const uint32_t* xl)
{
  return e.kind >= 2 ? shape_of<uint64_t>(e.a, e.c, xl) : shape_of<uint32_t>(e.a, e.c, xl);
}

__device__ __forceinline__ void enc_pack_half(const EncTile& e, int h, const uint2 code[kHalfSub], const ChunkShape& sh,
                                              uint32_t* stage)
{
  if (e.kind == 0) pack_half<uint32_t, false>(e.a, e.c, h, code, sh, stage);
  else if (e.kind == 1) pack_half<uint32_t, true>(e.a, e.c, h, code, sh, stage);
  else if (e.kind == 2) pack_half<uint64_t, false>(e.a, e.c, h, code, sh, stage);
  else pack_half<uint64_t, true>(e.a, e.c, h, code, sh, stage);
}

// Wave 0, once the tile's payload is staged: the look-back, the inclusive
// word, the stream header (first tile) and length (last tile); the tile's
// stream offset goes to *xexcl.
__device__ __forceinline__ void enc_settle(const EncTile& e, int csize, unsigned long long* xexcl)
{
  const uint64_t excl = enc_look_back(e);
  if ((threadIdx.x & 63) == 0) {
    if (e.c != 0) publish_inclusive(e.st, e.c, e.nchunks, excl + (uint64_t)csize);
    *xexcl = excl;
    if (e.c == 0) gstore((long long*)e.a.out, e.a.insize);
    if (e.c == e.nchunks - 1) gstore(e.a.outsize, 8 + 2 * e.nchunks + (long long)excl + csize);
  }
}

// Encoder: one wave per tile, pipelined over the wave's tiles like the
// single-array encoder: the next tile's slices load while the current one
// is packed, its codes and size are made before the current look-back.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void batch_encode_kernel(
    const BatchItem* __restrict__ items, const uint32_t* __restrict__ tile_map, long long tiles, uint64_t* ws,
    LookCfg lc)
{
  __shared__ __align__(16) uint32_t stage[kStageBytes / 4];
  long long t = blockIdx.x;
  if (t >= tiles) return;
  EncTile ec = enc_tile(items, tile_map, t, ws, lc);
  uint2 cur[kSub];
  load_raw(ec.a, ec.c, cur);
  ChunkShape sc = enc_shape(ec, cur);
  publish_chunk(ec.a, ec.st, ec.c, ec.nchunks, sc.csize);
  // The last tile is finished after the loop, so the loop body has no
  // branch on whether a next tile exists: with one, the compiler could not
  // tell that the next tile's loads were consumed, and waited for every
  // memory operation in flight (this tile's stores included) before the
  // next iteration.
  for (long long n = t + gridDim.x; n < tiles; n += gridDim.x) {
    const EncTile en = enc_tile(items, tile_map, n, ws, lc);
    uint2 nxt[kSub];
    load_raw(en.a, en.c, nxt);
    enc_pack(ec, cur, sc, stage);
    __builtin_amdgcn_s_setprio(2);   // others' look-backs wait on this size
    const ChunkShape sn = enc_shape(en, nxt);
    publish_chunk(en.a, en.st, en.c, en.nchunks, sn.csize);
    __builtin_amdgcn_s_setprio(0);
    enc_finish(ec, sc.csize, stage);
    ec = en;
    sc = sn;
#pragma unroll
    for (int s = 0; s < kSub; s++) cur[s] = nxt[s];
  }
  enc_pack(ec, cur, sc, stage);
  enc_finish(ec, sc.csize, stage);
}

// Encoder, two waves per tile: wave h holds subchunks 16h .. 16h+15 of the
// tile, as in the single-array two-wave encoder, and the next tile loads and
// is coded while the current one is packed, looked back and stored.  The
// last tile is finished after the loop, so the loop body has no branch on
// whether a next tile exists.  A workgroup takes every tile, its first
// included, from the tile tickets, as the single-array ticket encoder takes
// chunks, so an array's tiles start in arrival order and no tile waits on a
// workgroup that has not started.
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4, 4))) void batch_encode_pair_kernel(
    const BatchItem* __restrict__ items, const uint32_t* __restrict__ tile_map, long long tiles, uint64_t* ws,
    LookCfg lc, uint32_t* tickets)
{
  constexpr bool TICKET = true;
  __shared__ __align__(16) uint32_t stage[kStageBytes / 4];
  __shared__ uint32_t xl[2][8];
  __shared__ unsigned long long xexcl;
  __shared__ long long xnext, xfirst;
  const int h = uni((int)(threadIdx.x >> 6));   // wave-uniform: scalar branches and bit lengths
  const int lane = threadIdx.x & 63;
  long long t = blockIdx.x;
  if (TICKET) {
    if (h == 0) {
      const long long m = ticket_first(tickets, lc.ticket_shift, tiles);
      if (lane == 0) xfirst = m;
    }
    __syncthreads();
    t = uni(xfirst);
    if (t < 0) return;   // uniform: the home set is spent, the others take the tiles
  }
  if (t >= tiles) return;
  uint32_t tq = 0;
  if (TICKET && h == 0) tq = ticket_request(tickets, lc.ticket_shift, tq);
  EncTile ec = enc_tile(items, tile_map, t, ws, lc);
  uint2 cur[kHalfSub];
  uint32_t lh[4];
  load_half(ec.a, ec.c, h, cur);
  enc_shape_half(ec, h, cur, lh);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++) xl[0][4 * h + i] = lh[i];
  }
  if (TICKET && h == 0) {
    const long long m = ticket_chunk(tickets, lc.ticket_shift, tq, tiles);
    if (lane == 0) xnext = m;
  }
  __syncthreads();
  ChunkShape sc = enc_shape_of(ec, xl[0]);
  if (h == 0) publish_chunk(ec.a, ec.st, ec.c, ec.nchunks, sc.csize);
  int par = 1;
  long long n = TICKET ? uni(xnext) : t + gridDim.x;
  const uint32_t one = ticket_one();
  while (TICKET ? n >= 0 : n < tiles) {
    if (TICKET && h == 0) tq = ticket_request(tickets, lc.ticket_shift, tq, one);
    const EncTile en = enc_tile(items, tile_map, n, ws, lc);
    uint2 nxt[kHalfSub];
    load_half(en.a, en.c, h, nxt);
    enc_pack_half(ec, h, cur, sc, stage);
    __builtin_amdgcn_s_setprio(2);   // others' look-backs wait on this size
    enc_shape_half(en, h, nxt, lh);
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < 4; i++) xl[par][4 * opaque(h) + i] = lh[i];
    }
    if (h == 1) __builtin_amdgcn_s_setprio(0);
    __syncthreads();   // both halves packed; the next tile's bit lengths posted
    const ChunkShape sn = enc_shape_of(en, xl[par]);
    if (h == 0) {
      publish_chunk(en.a, en.st, en.c, en.nchunks, sn.csize);
      __builtin_amdgcn_s_setprio(0);
      enc_settle(ec, sc.csize, &xexcl);
      if (TICKET) {
        const long long m = ticket_chunk(tickets, lc.ticket_shift, tq, tiles);
        if (lane == 0) xnext = m;
      }
    }
    __syncthreads();   // the tile's stream offset known (and the tile after n)
    const long long nn = TICKET ? uni(xnext) : n + gridDim.x;
    pair_store(ec.a.out + 8 + 2 * ec.nchunks + (long long)uni((long long)xexcl), stage, sc.csize, h);
    __syncthreads();   // staging buffer read before it is packed again
    ec = en;
    sc = sn;
    par ^= 1;
#pragma unroll
    for (int s = 0; s < kHalfSub; s++) cur[s] = nxt[s];
    n = nn;
  }
  enc_pack_half(ec, h, cur, sc, stage);
  __syncthreads();
  if (h == 0) enc_settle(ec, sc.csize, &xexcl);
  __syncthreads();
  pair_store(ec.a.out + 8 + 2 * ec.nchunks + (long long)uni((long long)xexcl), stage, sc.csize, h);
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
  scan_offsets(decode_args(items + blockIdx.x, ws), wsum, 0, 1);
}

// Decoder: workgroups stride over the tiles, one tile at a time; tiles
// beyond the stream's chunk count (the plan sizes arrays by their capacity)
// are skipped.  At most 64 VGPRs (eight waves per SIMD, as the single-array
// decoder): the four element kinds' paths would otherwise take 81 and leave
// five.
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void batch_decode_kernel(
    const BatchItem* items, const uint32_t* tile_map, long long tiles, uint64_t* ws)
{
  __shared__ __align__(16) uint32_t stage[kStageBytes / 4];
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

int batch_resident_grid(bool decompress, bool pair)
{
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const void* fn = decompress ? (const void*)batch_decode_kernel
                              : (pair ? (const void*)batch_encode_pair_kernel : (const void*)batch_encode_kernel);
  const int threads = decompress ? kThreads : (pair ? 128 : 64);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, 0) != hipSuccess) return 0;
  return cus * (per > 0 ? per : 1);
}

hipError_t batch_launch(bool decompress, const BatchItem* items, int n, long long tiles, uint64_t* ws, int grid,
                        int lb_shift, int lb_window, bool pair, uint32_t* tickets, hipStream_t stream)
{
  LookCfg lc{lb_shift, lb_window, help_state(), 0};
  if (n <= 0) return hipSuccess;
  uint32_t* const map = (uint32_t*)(ws + batch_map_word(n));
  const long long res = batch_resident_grid(decompress, pair);
  // the decoder has no cross-tile waits: an oversubscribed grid (32 tiles'
  // worth of workgroups per resident slot, measured best of 1x..200x) lets
  // the hardware overlap the tiles' descriptor and header latencies
  long long g = grid > 0 ? grid : (decompress ? 32 * res : res);
  if (!decompress && res > 0 && g > res) g = res;   // the encoder's look-back: no more than the resident set
  if (g <= 0) g = 1024;
  if (g > tiles) g = tiles > 0 ? tiles : 1;
  lc.ticket_shift = ticket_shift_for(g);
  const bool tk = !decompress && pair && tiles > 0;
  if (tk && !tickets) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_prep_kernel, dim3(n), dim3(256), 0, stream, items, map, ws, decompress ? 0 : 1,
                     tk ? (uint64_t*)tickets : (uint64_t*)nullptr);
  if (decompress) {
    hipLaunchKernelGGL(batch_offsets_kernel, dim3(n), dim3(kScanThreads), 0, stream, items, ws);
    if (tiles > 0)
      hipLaunchKernelGGL(batch_decode_kernel, dim3((unsigned)g), dim3(kThreads), 0, stream, items, map, tiles, ws);
  } else if (tk) {
    hipLaunchKernelGGL(batch_encode_pair_kernel, dim3((unsigned)g), dim3(128), 0, stream, items, map, tiles, ws, lc,
                       tickets);
  } else if (tiles > 0) {
    hipLaunchKernelGGL(batch_encode_kernel, dim3((unsigned)g), dim3(64), 0, stream, items, map, tiles, ws, lc);
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
