// SLEEK encode / decode kernels for gfx950.
//
// encode: one workgroup (4 wave64s) codes one 16 KiB chunk at a time,
//   striding over the chunks in launch order.  Per chunk: 16-byte coalesced
//   loads (4 slices per thread) -> per-value code in registers -> per-
//   subchunk bit length by a half-wave max -> compressed/raw decision ->
//   publish the size -> OR-pack the codes into an LDS staging buffer -> wave
//   0 runs a two-level decoupled look-back for the chunk's stream offset ->
//   realigned 16-byte stores to the stream.
// decode: one workgroup of 1024 threads prefix-sums the uint16 size table
//   into per-chunk offsets; then workgroups stride over the chunks with no
//   cross-workgroup dependency: aligned 16-byte loads of the payload's cover
//   into LDS -> per-lane bit extraction -> inverse transform -> 16-byte
//   coalesced stores.
// Reference counterparts: d_encode / d_decode, sleek_common_gpu.h:474-768
// (lossless) and :989-1279 (lossy).
#include <mutex>
#include <vector>

#include "sleek_chunk.h"

namespace sleek {

// Look-back help state of this device: [0] help events since the library
// was loaded (each one a missing look-back word that a waiting wave sized
// itself), [1] the help threshold in naps (sleek_gpu_set_help_after).
// Kernels get its device address (help_state) as an argument.
__device__ unsigned long long help_words[2] = {0ull, kHelpNapsDefault};

// Scratch of sleek_gpu_compress_noalloc: the workspace of one compression
// of up to kNoAllocMaxBytes input bytes, in device memory that the code
// object brings (each device holds its own copy for as long as the library
// is loaded), so the call allocates nothing.  Host code never passes the
// array's name: it uses the device address noalloc_scratch() looked up for
// the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

void* noalloc_scratch()
{
  static std::mutex mu;
  static std::vector<void*> addr;   // per device
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
  if (!addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return nullptr;
    addr[dev] = p;
  }
  return addr[dev];
}

// The help state's per-device addresses and the threshold a device's state
// is given when its address is first looked up (kHelpNapsDefault until
// sleek_gpu_set_help_after).
static std::mutex help_mu;
static std::vector<unsigned long long*> help_addr;
static unsigned long long help_naps_host = kHelpNapsDefault;

static hipError_t write_help_naps(unsigned long long* state, unsigned long long naps)
{
  // a synchronous copy: kernels launched after it see the threshold
  return hipMemcpy(state + 1, &naps, sizeof naps, hipMemcpyHostToDevice);
}

unsigned long long* help_state()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(help_mu);
  if ((int)help_addr.size() <= dev) help_addr.resize(dev + 1, nullptr);
  if (!help_addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(help_words)) != hipSuccess || p == nullptr) return nullptr;
    if (help_naps_host != kHelpNapsDefault && write_help_naps((unsigned long long*)p, help_naps_host) != hipSuccess)
      return nullptr;
    help_addr[dev] = (unsigned long long*)p;
  }
  return help_addr[dev];
}

hipError_t set_help_naps(unsigned long long naps)
{
  std::lock_guard<std::mutex> lk(help_mu);
  help_naps_host = naps;
  int cur = 0;
  hipError_t e = hipSuccess;
  bool switched = false;
  for (int d = 0; d < (int)help_addr.size() && e == hipSuccess; d++) {
    if (!help_addr[d]) continue;   // set when the device's state is first looked up
    if (!switched) {
      e = hipGetDevice(&cur);
      if (e != hipSuccess) break;
      switched = true;
    }
    e = hipSetDevice(d);
    if (e == hipSuccess) e = write_help_naps(help_addr[d], naps);
  }
  if (switched) {
    const hipError_t r = hipSetDevice(cur);
    if (e == hipSuccess) e = r;
  }
  return e;
}

// Encoder: each workgroup walks its chunks (launch order, stride = grid) as
// a three-stage pipeline, one chunk per stage per iteration:
//   iteration k:  slices of chunk k+1 are loaded into registers (prefetch),
//                 chunk k is coded, its size published and its payload
//                 packed into LDS buffer k&1,
//                 chunk k-1's look-back runs (wave 0) and its payload, waiting
//                 in buffer (k-1)&1 since iteration k-1, is stored.
// Deferring the look-back by one iteration gives the chunk's predecessors a
// whole iteration to publish their sizes, so it rarely waits; the prefetch
// hides the load latency.  Two staging buffers put the occupancy at four
// workgroups per CU, which leaves room for the prefetch registers.
template <typename W, bool LOSSY, int WAVES>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES, 4))) void encode_kernel(EncodeArgs a)
{
  constexpr int PL = Word<W>::PerLane;
  __shared__ __align__(16) uint32_t stage[2][kStageBytes / 4];
  __shared__ uint32_t lgs[8];
  __shared__ unsigned long long s_excl;

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  const long long nchunks = (a.insize + kChunk - 1) / kChunk;
  const long long data_base = 8 + 2 * nchunks;
  uint16_t* const sizes = (uint16_t*)(a.out + 8);
  const ScanState st = scan_state(a.ws, nchunks, a.lb_shift);
  const W offs = (W)a.offs;
  const int woff = wv * (kWaveRounds * kSubBytes) + lane * 8;   // this lane's byte in round 0

  if (nchunks == 0) {
    if (blockIdx.x == 0 && t == 0) {
      *(long long*)a.out = 0;
      *a.outsize = 8;
    }
    return;
  }

  // prefetched slices of the chunk coded next
  uint2 nxt[kWaveRounds];
  long long c = blockIdx.x;
  if (c < nchunks) {
    const int os = (int)min((long long)kChunk, a.insize - c * kChunk);
#pragma unroll
    for (int r = 0; r < kWaveRounds; r++) nxt[r] = load_lane(a.in + c * kChunk, woff + r * kSubBytes, os);
  }
  long long prev = -1;     // chunk whose payload waits in stage[pb] (wave-uniform)
  int prev_csize = 0, pb = 0, cb = 0;

  for (;; c += gridDim.x) {
    const bool have = c < nchunks;
    int csize = 0;
    if (have) {
      uint32_t* const sb = stage[cb];
      for (int i = t; i < kChunk / 16 + 1; i += kThreads) ((uint4*)sb)[i] = make_uint4(0u, 0u, 0u, 0u);
      const long long base = c * kChunk;
      const int osize = (int)min((long long)kChunk, a.insize - base);

      W code[kWaveRounds][PL];
#pragma unroll
      for (int r = 0; r < kWaveRounds; r++) {
        W v[PL];
        lane_values(nxt[r], v);
#pragma unroll
        for (int k = 0; k < PL; k++) code[r][k] = v[k];
      }
      // prefetch the next chunk of this workgroup
      const long long cn = c + gridDim.x;
      if (cn < nchunks) {
        const int os = (int)min((long long)kChunk, a.insize - cn * kChunk);
#pragma unroll
        for (int r = 0; r < kWaveRounds; r++) nxt[r] = load_lane(a.in + cn * kChunk, woff + r * kSubBytes, os);
      }

      uint32_t lgw[kWaveRounds / 4];   // the wave's bit lengths, byte-packed (wave-uniform)
#pragma unroll
      for (int i = 0; i < kWaveRounds / 4; i++) lgw[i] = 0;
#pragma unroll
      for (int r = 0; r < kWaveRounds; r++) {
        W orv = 0;
#pragma unroll
        for (int k = 0; k < PL; k++) {
          code[r][k] = to_code<W, LOSSY>(code[r][k], a.eb_e, a.thr_e, offs);
          orv |= code[r][k];
        }
        lgw[r / 4] |= (uint32_t)wave_bit_length(orv) << (8 * (r % 4));
      }
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < kWaveRounds / 4; i++) lgs[wv * (kWaveRounds / 4) + i] = lgw[i];
      }
      __syncthreads();

      uint32_t L[8];
#pragma unroll
      for (int i = 0; i < 8; i++) L[i] = uni(lgs[i]);
      uint32_t nb = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) nb += byte_sum(L[i]);
      const int total_bits = (int)nb * Word<W>::PerSub;
      const int newsize = kSub + 2 + total_bits / 8;
      const bool packed = newsize < osize;   // osize <= kChunk, so also newsize < kChunk
      csize = packed ? newsize : osize;

      // publish the chunk's size at once; the group counter's reply is
      // looked at after the packing
      unsigned long long grp_old = 0;
      if (t == 0) {
        flag_max(st.chunk + c, (c == 0 ? kInc : kAgg) | (uint64_t)csize);
        grp_old = __hip_atomic_fetch_add(st.count + (c >> st.gshift), (1ull << 32) | (unsigned)csize, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        sizes[c] = (uint16_t)csize;
      }

      if (packed) {
        if (t < 8) sb[t] = L[t];                                // 32 bit-length bytes
        if (t == 64) sb[8 + total_bits / 32] = (uint32_t)osize;  // int16 trailer
        // wave-uniform subchunk index, bit length and bit offset per round
        int p0 = 8 * kSub + sub_bit_offset<W>(L, wv * kWaveRounds);
#pragma unroll
        for (int r = 0; r < kWaveRounds; r++) {
          const int ln = sub_length(L, wv * kWaveRounds + r);
          if (ln) or_field(sb, p0 + lane * PL * ln, lane_field(code[r], ln));
          p0 += ln * Word<W>::PerSub;
        }
      } else {
        // raw chunk: the original bytes (lossless codes invert exactly; lossy
        // ones do not, so the input is read again)
#pragma unroll
        for (int r = 0; r < kWaveRounds; r++) {
          uint2 q;
          if constexpr (LOSSY) {
            q = load_lane(a.in + base, woff + r * kSubBytes, osize);
          } else {
            W v[PL];
#pragma unroll
            for (int k = 0; k < PL; k++) v[k] = lossless_value<W>(code[r][k]);
            q = values_lane(v);
          }
          *(uint2*)((uint8_t*)sb + woff + r * kSubBytes) = q;
        }
      }

      if (t == 0) {
        // this chunk completes its group: publish the group's aggregate
        const long long g = c >> st.gshift;
        if ((long long)(grp_old >> 32) + 1 == group_members(st, g, nchunks))
          flag_max(st.group + g, (g == 0 ? kInc : kAgg) | ((grp_old & 0xffffffffull) + (unsigned)csize));
      }
    }

    if (prev >= 0) {
      if (t < 64) {
        // wave 0: the stream offset of the previous chunk
        const uint64_t excl = (prev == 0) ? 0ull : look_back<W, LOSSY>(a, st, prev, nchunks);
        if (t == 0) {
          if (prev != 0) publish_inclusive(st, prev, nchunks, excl + (uint64_t)prev_csize);
          s_excl = excl;
        }
      }
      __syncthreads();
      const long long excl = uni((long long)s_excl);
      store_unaligned(a.out + data_base + excl, stage[pb], prev_csize);
      if (t == 0) {
        if (prev == 0) *(long long*)a.out = a.insize;
        if (prev == nchunks - 1) *a.outsize = data_base + excl + prev_csize;
      }
    }
    __syncthreads();   // stage[pb] is reused only after every store read it
    if (!have) break;
    prev = c;
    prev_csize = csize;
    pb = cb;
    cb ^= 1;
  }
}

// Encoder, one wave per chunk, pipelined across a wave's chunks (launch
// order, stride = grid).  Iteration for chunk c, whose codes and size were
// made in the previous iteration:
//   issue the loads of the next chunk n -> pack c into LDS (covers the load
//   latency) -> code n, size it and publish its size -> look-back for c ->
//   store c.
// So a chunk's size is published a whole iteration before its look-back
// runs, and its predecessors' sizes, published by the other waves at about
// the same time, are there when it looks: the look-back rarely waits.  The
// coding and publishing of n run at a raised wave priority, since other
// waves' look-backs wait on it (2-5 % faster).  Two chunks' raw slices are
// live at once (128 VGPRs), two waves per SIMD.
//
// TICKET: the wave takes every chunk from the chunk tickets, as the two-wave
// ticket encoder does: the first before it starts, then the chunk after
// next, requested at the start of an iteration and read after the look-back.
template <typename W, bool LOSSY, bool TICKET>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void encode_wave2_kernel(EncodeArgs a)
{
  __shared__ __align__(16) uint32_t stage[kStageBytes / 4];
  const long long nchunks = (a.insize + kChunk - 1) / kChunk;
  if (nchunks == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *(long long*)a.out = 0;
      *a.outsize = 8;
    }
    return;
  }
  const ScanState st = scan_state(a.ws, nchunks, a.lb_shift);
  long long c = blockIdx.x;
  if (TICKET) c = ticket_first(a.tickets, a.ticket_shift, nchunks);   // uniform; -1: the home set is spent
  if (c < 0 || c >= nchunks) return;
  uint32_t tq = 0;
  if (TICKET) tq = ticket_request(a.tickets, a.ticket_shift, tq);
  uint2 cur[kSub];
  load_raw(a, c, cur);
  ChunkShape sc = shape_codes<W, LOSSY>(a, c, cur);
  publish_chunk(a, st, c, nchunks, sc.csize);
  long long n = TICKET ? ticket_chunk(a.tickets, a.ticket_shift, tq, nchunks) : c + gridDim.x;
  const uint32_t one = ticket_one();
  while (true) {
    const bool more = TICKET ? n >= 0 : n < nchunks;
    if (TICKET && more) tq = ticket_request(a.tickets, a.ticket_shift, tq, one);
    uint2 nxt[kSub];
    if (more) load_raw(a, n, nxt);
    pack_codes<W, LOSSY>(a, c, cur, sc, stage);
    ChunkShape sn;
    if (more) {
      // other waves' look-backs wait for n's size: code it ahead of the
      // SIMD's other wave
      __builtin_amdgcn_s_setprio(2);
      sn = shape_codes<W, LOSSY>(a, n, nxt);
      publish_chunk(a, st, n, nchunks, sn.csize);
      __builtin_amdgcn_s_setprio(0);
    }
    finish_chunk(a, st, c, nchunks, sc.csize, stage, (c == 0) ? 0ull : look_back<W, LOSSY>(a, st, c, nchunks));
    if (!more) break;
    const long long nn = TICKET ? ticket_chunk(a.tickets, a.ticket_shift, tq, nchunks) : n + gridDim.x;
    c = n;
    n = nn;
    sc = sn;
#pragma unroll
    for (int s = 0; s < kSub; s++) cur[s] = nxt[s];
  }
}

// Encoder, two waves per chunk, pipelined over the workgroup's chunks like
// encode_wave2_kernel: half the registers per wave (four waves per SIMD).
// Iteration for chunk c (codes and shape made in the previous iteration):
//   both waves: issue their half of chunk n's loads, pack their half of c;
//   code their half of n, post its bit lengths -> barrier -> n's shape;
//   wave 0: publish n's size, look back for c -> barrier -> both store
//   their lines of c -> barrier (staging buffer free).
//
// TICKET: a workgroup takes every chunk from the chunk tickets
// (EncodeArgs::tickets): the first one before it starts, then the chunk
// after next, requested at the start of an iteration and read after the
// look-back, so the atomic's round trip overlaps the iteration.  Chunks then
// start in arrival order, a chunk's predecessors have mostly published their
// sizes by its look-back however far individual workgroups drift apart, and
// no chunk waits on a workgroup that has not started.
template <typename W, bool LOSSY, bool TICKET>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4, 4))) void encode_pair_kernel(EncodeArgs a)
{
  __shared__ __align__(16) uint32_t stage[kStageBytes / 4];
  __shared__ uint32_t xl[2][8];
  __shared__ unsigned long long xexcl;
  __shared__ long long xnext, xfirst;
  const int h = uni((int)(threadIdx.x >> 6));   // wave-uniform: scalar branches and bit lengths
  const int lane = threadIdx.x & 63;
  // Where wave 0 publishes chunk c's inclusive word: before the store (the
  // compiler then drains it, with every other memory operation in flight,
  // before the store loop) or after it (later for the chunks that wait on
  // it).  After is faster for the float32 lossy encoders' short stores,
  // before for the others (measured).
  constexpr bool kLatePublish = LOSSY && sizeof(W) == 4;
  const long long nchunks = (a.insize + kChunk - 1) / kChunk;
  if (nchunks == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *(long long*)a.out = 0;
      *a.outsize = 8;
    }
    return;
  }
  const ScanState st = scan_state(a.ws, nchunks, a.lb_shift);
  const long long data_base = 8 + 2 * nchunks;
  long long c = blockIdx.x;
  if (TICKET) {
    if (h == 0) {
      const long long m = ticket_first(a.tickets, a.ticket_shift, nchunks);
      if (lane == 0) xfirst = m;
    }
    __syncthreads();
    c = uni(xfirst);
    if (c < 0) return;   // uniform: the home set is spent, the others take the chunks
  }
  if (c >= nchunks) return;
  uint2 cur[kHalfSub];
  uint32_t lh[4];
  uint32_t tq = 0;
  if (TICKET && h == 0) tq = ticket_request(a.tickets, a.ticket_shift, tq);
  load_half(a, c, h, cur);
  shape_half<W, LOSSY>(a, c, h, cur, lh);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++) xl[0][4 * h + i] = lh[i];
  }
  if (TICKET && h == 0) {
    const long long m = ticket_chunk(a.tickets, a.ticket_shift, tq, nchunks);
    if (lane == 0) xnext = m;
  }
  __syncthreads();
  ChunkShape sc = shape_of<W>(a, c, xl[0]);
  if (h == 0) publish_chunk(a, st, c, nchunks, sc.csize);
  long long n = TICKET ? uni(xnext) : c + gridDim.x;
  int par = 1;
  const uint32_t one = ticket_one();
  while (true) {
    const bool more = TICKET ? n >= 0 : n < nchunks;
    uint2 nxt[kHalfSub];
    if (TICKET && h == 0 && more) tq = ticket_request(a.tickets, a.ticket_shift, tq, one);
    if (more) load_half(a, n, h, nxt);
    pack_half<W, LOSSY>(a, c, h, cur, sc, stage);
    __builtin_amdgcn_s_setprio(2);   // n's size is waited on by others: as in encode_wave2_kernel
    if (more) {
      shape_half<W, LOSSY>(a, n, h, nxt, lh);
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 4; i++) xl[par][4 * opaque(h) + i] = lh[i];
      }
    }
    if (h == 1) __builtin_amdgcn_s_setprio(0);   // wave 0 keeps it until n's size is published
    __syncthreads();   // both halves packed; n's bit lengths posted
    ChunkShape sn;
    if (more) sn = shape_of<W>(a, n, xl[par]);
    if (h == 0) {
      if (more) publish_chunk(a, st, n, nchunks, sn.csize);
      __builtin_amdgcn_s_setprio(0);
      const uint64_t excl = (c == 0) ? 0ull : look_back<W, LOSSY>(a, st, c, nchunks);
      long long m = -1;
      if (TICKET && more) m = ticket_chunk(a.tickets, a.ticket_shift, tq, nchunks);
      if (lane == 0) {
        if (!kLatePublish && c != 0) publish_inclusive(st, c, nchunks, excl + (uint64_t)sc.csize);
        xexcl = excl;
        if (TICKET) xnext = m;
        if (c == 0) *(long long*)a.out = a.insize;
        if (c == nchunks - 1) *a.outsize = data_base + (long long)excl + sc.csize;
      }
    }
    __syncthreads();   // c's stream offset known (and the chunk after n)
    const long long excl = (long long)uni((long long)xexcl);
    const long long nn = TICKET ? uni(xnext) : n + gridDim.x;
    pair_store(a.out + data_base + excl, stage, sc.csize, h);
    if (kLatePublish && h == 0 && lane == 0 && c != 0)
      publish_inclusive(st, c, nchunks, (uint64_t)excl + (uint64_t)sc.csize);
    __syncthreads();   // staging buffer read before it is packed again
    if (!more) break;
    c = n;
    n = nn;
    sc = sn;
    par ^= 1;
#pragma unroll
    for (int s = 0; s < kHalfSub; s++) cur[s] = nxt[s];
  }
}

__global__ __launch_bounds__(kScanThreads) void offsets_kernel(DecodeArgs a)
{
  __shared__ long long wsum[kScanThreads / 64];
  scan_offsets(a, wsum, (int)blockIdx.x, (int)gridDim.x);
}

// Decoder without a separate offset scan: workgroup b decodes chunks b,
// b+G, b+2G, ... (G = grid <= kWinMax) and derives each chunk's stream
// offset from the previous one it decoded:
//   excl(c + G) = excl(c) + sizes[c] + ... + sizes[c + G - 1],
// a window of G table entries (L2-resident) whose loads are issued before
// chunk c is decoded and summed after.  The first chunk's offset is the sum
// of the entries before it.  No cross-workgroup dependency, no offsets
// table, one launch.
template <typename W, bool LOSSY>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void decode_fused_kernel(DecodeArgs a)
{
  __shared__ __align__(16) uint32_t stage[kStageBytes / 4];
  __shared__ long long red[2][kThreads / 64];
  const long long hdr = uni(*(const long long*)a.in);
  long long nchunks = 0;
  uint64_t why = 0;
  const bool ok = header_ok(a, hdr, &nchunks, &why);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (ok) *a.decsize = hdr;
    else raise_status(a.ws, why);
  }
  long long c = blockIdx.x;
  if (!ok || c >= nchunks) return;
  const uint16_t* const sizes = (const uint16_t*)(a.in + 8);
  uint32_t v[kWinLoads];
  table_load(sizes, 0, c, v);
  long long excl = table_sum(v, 0, c, red[0]);
  int par = 1;
  while (true) {
    const long long n = c + gridDim.x;
    const long long wend = min(n, nchunks);
    table_load(sizes, c, wend, v);
    decode_chunk_at<W, LOSSY>(a, hdr, nchunks, c, excl, stage);
    if (n >= nchunks) break;
    excl += table_sum(v, c, wend, red[par]);
    par ^= 1;
    c = n;
  }
}

// Decoder: workgroups stride over the chunks in any order.  A chunk's
// offset and size come from the offsets table (offs[c] and offs[c + 1], the
// offsets kernel having written entries 0 .. nchunks), loaded together with
// the stream header: one round trip before the payload's instead of two.
// The table is read before the header is checked, at indexes below the
// grid's chunk count plus one, which the workspace (sized for the capacity's
// chunks, as is the grid) holds; entries beyond the stream's chunks are not
// used.
template <typename W, bool LOSSY>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void decode_kernel(DecodeArgs a)
{
  __shared__ __align__(16) uint32_t stage[kStageBytes / 4];
  const long long* const offs_tab = (const long long*)(a.ws + kWsFlags);
  long long c = blockIdx.x;
  const long long hdr_raw = *(const long long*)a.in;
  long long o0 = offs_tab[c], o1 = offs_tab[c + 1];
  const long long hdr = uni(hdr_raw);
  long long nchunks = 0;
  uint64_t why = 0;
  if (!header_ok(a, hdr, &nchunks, &why)) return;   // reported by offsets_kernel
  for (; c < nchunks; c += gridDim.x) {
    if (c != (long long)blockIdx.x) {
      o0 = offs_tab[c];
      o1 = offs_tab[c + 1];
    }
    const long long excl = uni(o0);
    decode_chunk_at<W, LOSSY>(a, hdr, nchunks, c, excl, stage, (int)(uni(o1) - excl));
  }
}

// ---- launchers ----

template <typename W, bool LOSSY>
static hipError_t launch_encode(const EncodeArgs& a, int grid, int variant, hipStream_t stream)
{
  variant = resolve_variant(variant, (int)sizeof(W));
  if (variant == kEncodeWavePipe) hipLaunchKernelGGL((encode_wave2_kernel<W, LOSSY, false>), dim3(grid), dim3(64), 0, stream, a);
  else if (variant == kEncodeWaveTicket) hipLaunchKernelGGL((encode_wave2_kernel<W, LOSSY, true>), dim3(grid), dim3(64), 0, stream, a);
  else if (variant == kEncodePair) hipLaunchKernelGGL((encode_pair_kernel<W, LOSSY, false>), dim3(grid), dim3(128), 0, stream, a);
  else if (variant == kEncodePairTicket) hipLaunchKernelGGL((encode_pair_kernel<W, LOSSY, true>), dim3(grid), dim3(128), 0, stream, a);
  else if (variant == kEncodeFree) hipLaunchKernelGGL((encode_kernel<W, LOSSY, 1>), dim3(grid), dim3(kThreads), 0, stream, a);
  else hipLaunchKernelGGL((encode_kernel<W, LOSSY, 4>), dim3(grid), dim3(kThreads), 0, stream, a);
  return hipGetLastError();
}

template <typename W, bool LOSSY>
static hipError_t launch_decode(const DecodeArgs& a, int grid, int variant, hipStream_t stream, hipEvent_t ev_begin,
                                hipEvent_t ev_end)
{
  hipError_t e = hipSuccess;
  if (variant == kDecodeScan) {
    // parts of whole rounds, from the capacity's chunk count (the stream's own
    // count is read by the kernel; parts beyond it have nothing to do), at
    // most the decode grid
    const long long rounds = ((a.cap + kChunk - 1) / kChunk + kScanRound - 1) / kScanRound;
    const int parts = (int)std::max(1ll, std::min({(long long)kScanPartsMax, rounds, (long long)grid}));
    hipLaunchKernelGGL(offsets_kernel, dim3(parts), dim3(kScanThreads), 0, stream, a);
    if (ev_begin && (e = hipEventRecord(ev_begin, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL((decode_kernel<W, LOSSY>), dim3(grid), dim3(kThreads), 0, stream, a);
  } else {
    if (ev_begin && (e = hipEventRecord(ev_begin, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL((decode_fused_kernel<W, LOSSY>), dim3(grid < kWinMax ? grid : kWinMax), dim3(kThreads), 0,
                       stream, a);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (ev_end) e = hipEventRecord(ev_end, stream);
  return e;
}

hipError_t encode_launch(int elem_bytes, bool lossy, const EncodeArgs& a, int grid, int variant, hipStream_t stream)
{
  if (elem_bytes == 4) return lossy ? launch_encode<uint32_t, true>(a, grid, variant, stream) : launch_encode<uint32_t, false>(a, grid, variant, stream);
  return lossy ? launch_encode<uint64_t, true>(a, grid, variant, stream) : launch_encode<uint64_t, false>(a, grid, variant, stream);
}

hipError_t decode_launch(int elem_bytes, bool lossy, const DecodeArgs& a, int grid, int variant, hipStream_t stream,
                         hipEvent_t ev_begin, hipEvent_t ev_end)
{
  if (elem_bytes == 4)
    return lossy ? launch_decode<uint32_t, true>(a, grid, variant, stream, ev_begin, ev_end)
                 : launch_decode<uint32_t, false>(a, grid, variant, stream, ev_begin, ev_end);
  return lossy ? launch_decode<uint64_t, true>(a, grid, variant, stream, ev_begin, ev_end)
               : launch_decode<uint64_t, false>(a, grid, variant, stream, ev_begin, ev_end);
}

template <typename W, bool LOSSY>
static const void* encode_fn(int variant)
{
  variant = resolve_variant(variant, (int)sizeof(W));
  if (variant == kEncodeWavePipe) return (const void*)encode_wave2_kernel<W, LOSSY, false>;
  if (variant == kEncodeWaveTicket) return (const void*)encode_wave2_kernel<W, LOSSY, true>;
  if (variant == kEncodePair) return (const void*)encode_pair_kernel<W, LOSSY, false>;
  if (variant == kEncodePairTicket) return (const void*)encode_pair_kernel<W, LOSSY, true>;
  return variant == kEncodeFree ? (const void*)encode_kernel<W, LOSSY, 1> : (const void*)encode_kernel<W, LOSSY, 4>;
}

int resident_grid(int elem_bytes, bool lossy, bool decode, int variant)
{
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const void* fn;
  if (decode) {
    // the fused decoder's occupancy (the offset-scan decoder is launched one workgroup per chunk)
    fn = elem_bytes == 4 ? (lossy ? (const void*)decode_fused_kernel<uint32_t, true> : (const void*)decode_fused_kernel<uint32_t, false>)
                         : (lossy ? (const void*)decode_fused_kernel<uint64_t, true> : (const void*)decode_fused_kernel<uint64_t, false>);
  } else {
    fn = elem_bytes == 4 ? (lossy ? encode_fn<uint32_t, true>(variant) : encode_fn<uint32_t, false>(variant))
                         : (lossy ? encode_fn<uint64_t, true>(variant) : encode_fn<uint64_t, false>(variant));
  }
  const int rv = resolve_variant(variant, elem_bytes);
  const int threads = decode ? kThreads : ((rv == kEncodeWavePipe || rv == kEncodeWaveTicket) ? 64 : (rv == kEncodePair || rv == kEncodePairTicket ? 128 : kThreads));
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, 0) != hipSuccess) return 0;
  return cus * (per > 0 ? per : 1);
}

}  // namespace sleek
