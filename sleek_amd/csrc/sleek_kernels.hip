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
#include "sleek_device.h"
#include "sleek_internal.h"

namespace sleek {

template <typename W>
__device__ __forceinline__ void or_code(uint32_t* st, int b, W v);

template <>
__device__ __forceinline__ void or_code<uint32_t>(uint32_t* st, int b, uint32_t v)
{
  const int d = b >> 5;
  const uint32_t sh = (uint32_t)b & 31u;
  atomicOr(&st[d], v << sh);
  if (sh) {
    const uint32_t hi = v >> (32u - sh);
    if (hi) atomicOr(&st[d + 1], hi);
  }
}

template <>
__device__ __forceinline__ void or_code<uint64_t>(uint32_t* st, int b, uint64_t v)
{
  const int d = b >> 5;
  const uint32_t sh = (uint32_t)b & 31u;
  const uint64_t lo = v << sh;
  const uint32_t w0 = (uint32_t)lo, w1 = (uint32_t)(lo >> 32);
  const uint32_t w2 = sh ? (uint32_t)(v >> (64u - sh)) : 0u;
  if (w0) atomicOr(&st[d], w0);
  if (w1) atomicOr(&st[d + 1], w1);
  if (w2) atomicOr(&st[d + 2], w2);
}

__device__ __forceinline__ uint32_t extract_code(const uint32_t* st, int b, int ln, uint32_t*)
{
  const int d = b >> 5;
  const uint32_t v = __builtin_amdgcn_alignbit(st[d + 1], st[d], (uint32_t)b & 31u);
  return ln >= 32 ? v : (v & ((1u << ln) - 1u));
}

__device__ __forceinline__ uint64_t extract_code(const uint32_t* st, int b, int ln, uint64_t*)
{
  const int d = b >> 5;
  const uint32_t sh = (uint32_t)b & 31u;
  const uint32_t w0 = st[d], w1 = st[d + 1], w2 = st[d + 2];
  const uint64_t v = (uint64_t)__builtin_amdgcn_alignbit(w1, w0, sh) | ((uint64_t)__builtin_amdgcn_alignbit(w2, w1, sh) << 32);
  return ln >= 64 ? v : (v & ((1ull << ln) - 1ull));
}

// Load the 16-byte slice at byte off of a chunk of osize bytes, zero beyond
// osize.  The slice is 16-byte aligned, so when any of its bytes is valid the
// whole slice lies in a mapped page and can be loaded at once.
__device__ __forceinline__ uint4 load_slice(const uint8_t* src, int off, int osize)
{
  if (off >= osize) return make_uint4(0u, 0u, 0u, 0u);
  uint4 q = *(const uint4*)(src + off);
  const int valid = osize - off;   // bytes; a multiple of 4
  if (valid < 16) {
    q.w = 0u;
    if (valid < 12) q.z = 0u;
    if (valid < 8) q.y = 0u;
  }
  return q;
}

__device__ __forceinline__ void store_slice(uint8_t* dst, int off, int osize, uint4 q)
{
  if (off + 16 <= osize) {
    *(uint4*)(dst + off) = q;
    return;
  }
  const uint32_t d[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (off + 4 * i + 4 <= osize) ((uint32_t*)(dst + off))[i] = d[i];
}

// Compressed size of chunk j, computed by one whole wave from the input (the
// same bit lengths the chunk's own workgroup finds).  Used by a look-back
// that has waited too long on an unpublished predecessor.
template <typename W, bool LOSSY>
__device__ uint32_t wave_chunk_size(const EncodeArgs& a, long long j)
{
  const int lane = threadIdx.x & 63;
  const long long base = j * kChunk;
  const int osize = (int)min((long long)kChunk, a.insize - base);
  const uint8_t* const src = a.in + base;
  uint32_t nb = 0;
  for (int r = 0; r < kChunk / (64 * 16); r++) {   // 16 slices of 1 KiB: subchunks 2r and 2r+1
    W v[Word<W>::PerSlice];
    slice_values(load_slice(src, r * 1024 + lane * 16, osize), v);
    W orv = 0;
#pragma unroll
    for (int k = 0; k < Word<W>::PerSlice; k++) orv |= to_code<W, LOSSY>(v[k], a.eb_e, a.thr_e, (W)a.offs);
    const int bl = half_wave_max(bit_length(orv));
    nb += (uint32_t)(__shfl(bl, 0, 64) + __shfl(bl, 32, 64));
  }
  const int newsize = kSub + 2 + (int)(nb * Word<W>::PerSub / 8);
  return (uint32_t)(newsize < osize ? newsize : osize);
}

// Compressed size of group gj (sum over its chunks), by one whole wave:
// published aggregate words are used, every other chunk is sized from the
// input.
template <typename W, bool LOSSY>
__device__ uint64_t wave_group_size(const EncodeArgs& a, const ScanState& s, long long gj, long long nchunks)
{
  const int lane = threadIdx.x & 63;
  const long long first = gj * kGroup;
  const int members = (int)min((long long)kGroup, nchunks - first);
  uint64_t f = (lane < members) ? flag_load(s.chunk + first + lane) : kAgg;
  uint64_t mine = ((f >> 62) == 1) ? (f & kValMask) : 0;
  uint64_t todo = __ballot(lane < members && (f >> 62) != 1);
  while (todo) {
    const int l = __ffsll((unsigned long long)todo) - 1;
    const uint32_t sz = wave_chunk_size<W, LOSSY>(a, first + l);
    if (lane == l) mine = sz;
    todo &= todo - 1;
  }
  return wave_sum(mine);
}

// Help threshold of a look-back: after this long (100 MHz real-time ticks)
// without progress, the wave sizes the blocking chunk or group itself.
constexpr uint64_t kHelpTicks = 2000;   // 20 us

// Two-level decoupled look-back, run by one whole wave; returns the sum of
// the compressed sizes of chunks 0..c-1 in every lane.  Lanes watch the
// preceding chunks of c's own group and, at the same time, a 64-group window
// of earlier groups, so one round trip resolves up to 64 groups (4096
// chunks) and the inclusive frontier advances that far per round trip.
// Workgroups take chunks in launch order, so the words waited on are
// normally published within a round trip or two; a word still missing after
// kHelpTicks is computed from the input and published by the waiter, so
// the look-back finishes even when a predecessor's workgroup is not running.
template <typename W, bool LOSSY>
__device__ uint64_t look_back(const EncodeArgs& a, const ScanState& s, long long c, long long nchunks)
{
  const int lane = threadIdx.x & 63;
  const int q = (int)(c % kGroup);                 // chunks of the group before c
  long long gend = c / kGroup;                      // group window: groups gend-64 .. gend-1
  bool have_partial = (q == 0);
  uint64_t partial = 0, gacc = 0;
  uint64_t since = __builtin_amdgcn_s_memrealtime();
  while (true) {
    uint64_t f = kAgg;                              // lanes beyond the group: ready, value 0
    if (!have_partial && lane < q) f = flag_load(s.chunk + (c - 1 - lane));
    const long long gj = gend - 1 - lane;
    const uint64_t fg = (gj >= 0) ? flag_load(s.group + gj) : kInc;   // before group 0: inclusive 0
    if (!have_partial) {
      const uint64_t inc = __ballot((f >> 62) == 2);
      const uint64_t rdy = __ballot((f >> 62) != 0);
      if (inc) {
        const int p = __ffsll((unsigned long long)inc) - 1;
        const uint64_t need = (2ull << p) - 1;      // p < 63 here
        if ((rdy & need) == need) return wave_sum(lane <= p ? (f & kValMask) : 0);
      } else if (rdy == ~0ull) {
        partial = wave_sum(f & kValMask);
        have_partial = true;
        since = __builtin_amdgcn_s_memrealtime();
      }
    }
    if (have_partial) {
      const uint64_t inc = __ballot((fg >> 62) == 2);
      const uint64_t rdy = __ballot((fg >> 62) != 0);
      if (inc) {
        const int p = __ffsll((unsigned long long)inc) - 1;
        const uint64_t need = (p == 63) ? ~0ull : ((2ull << p) - 1);
        if ((rdy & need) == need) return partial + gacc + wave_sum(lane <= p ? (fg & kValMask) : 0);
      } else if (rdy == ~0ull) {
        gacc += wave_sum(fg & kValMask);
        gend -= 64;
        since = __builtin_amdgcn_s_memrealtime();
        continue;
      }
    }
    if (__builtin_amdgcn_s_memrealtime() - since > kHelpTicks) {
      // size the nearest unpublished word ourselves and publish it
      if (!have_partial) {
        const uint64_t miss = __ballot(lane < q && (f >> 62) == 0);
        if (miss) {
          const long long jc = c - __ffsll((unsigned long long)miss);
          const uint32_t sz = wave_chunk_size<W, LOSSY>(a, jc);
          if (lane == 0) flag_max(s.chunk + jc, kAgg | sz);
        }
      } else {
        const uint64_t miss = __ballot((fg >> 62) == 0);
        if (miss) {
          const long long gm = gend - __ffsll((unsigned long long)miss);
          const uint64_t sz = wave_group_size<W, LOSSY>(a, s, gm, nchunks);
          if (lane == 0) flag_max(s.group + gm, (gm == 0 ? kInc : kAgg) | sz);
        }
      }
      since = __builtin_amdgcn_s_memrealtime();
      continue;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

template <typename W, bool LOSSY, int WAVES>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void encode_kernel(EncodeArgs a)
{
  constexpr int PS = Word<W>::PerSlice;
  __shared__ uint32_t stage[kStageBytes / 4];
  __shared__ uint32_t lgs[8];
  __shared__ unsigned long long s_excl;

  const int t = threadIdx.x;
  const int lane32 = t & 31;
  const int half = t >> 5;
  const long long nchunks = (a.insize + kChunk - 1) / kChunk;
  const long long data_base = 8 + 2 * nchunks;
  uint16_t* const sizes = (uint16_t*)(a.out + 8);
  const ScanState st = scan_state(a.ws, nchunks);
  const W offs = (W)a.offs;

  if (nchunks == 0) {
    if (blockIdx.x == 0 && t == 0) {
      *(long long*)a.out = 0;
      *a.outsize = 8;
    }
    return;
  }

  // Workgroups stride over the chunks in launch order; no ticket counter
  // (a single same-address atomic per chunk serialises at ~12 ns each).
  for (long long c = blockIdx.x; c < nchunks; c += gridDim.x) {
    for (int i = t; i < kChunk / 16 + 1; i += kThreads) ((uint4*)stage)[i] = make_uint4(0u, 0u, 0u, 0u);
    const long long base = c * kChunk;
    const int osize = (int)min((long long)kChunk, a.insize - base);
    const uint8_t* const src = a.in + base;

    // Only the raw slices stay live across the barriers (16 VGPRs); the
    // codes are cheap and are computed again in the packing pass.
    uint4 raw[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r++) raw[r] = load_slice(src, r * kThreads * 16 + t * 16, osize);

#pragma unroll
    for (int r = 0; r < kRounds; r++) {
      W v[PS];
      slice_values(raw[r], v);
      W orv = 0;
#pragma unroll
      for (int k = 0; k < PS; k++) orv |= to_code<W, LOSSY>(v[k], a.eb_e, a.thr_e, offs);
      const int bl = half_wave_max(bit_length(orv));
      if (lane32 == 0) ((uint8_t*)lgs)[r * 8 + half] = (uint8_t)bl;
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kRounds; r++) reissue(raw[r]);
    uint32_t L[8];
#pragma unroll
    for (int i = 0; i < 8; i++) L[i] = uni(lgs[i]);
    uint32_t nb = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) nb += byte_sum(L[i]);
    const int total_bits = (int)nb * Word<W>::PerSub;
    const int newsize = kSub + 2 + total_bits / 8;
    const bool packed = newsize < osize;   // osize <= kChunk, so also newsize < kChunk
    const int csize = packed ? newsize : osize;

    if (t == 0) {
      // publish the chunk's size at once: successors can sum it while this
      // workgroup packs and looks back
      publish_size(st, c, nchunks, (uint32_t)csize);
      sizes[c] = (uint16_t)csize;
    }

    if (packed) {
      if (t < 8) stage[t] = L[t];                                // 32 bit-length bytes
      if (t == 64) stage[8 + total_bits / 32] = (uint32_t)osize;  // int16 trailer
#pragma unroll
      for (int r = 0; r < kRounds; r++) {
        const int s = r * 8 + half;
        const int ln = sub_length(L, s);
        if (ln == 0) continue;
        const int p0 = 8 * kSub + sub_bit_offset<W>(L, s) + lane32 * PS * ln;
        W v[PS];
        slice_values(raw[r], v);
#pragma unroll
        for (int k = 0; k < PS; k++) {
          const W cv = to_code<W, LOSSY>(v[k], a.eb_e, a.thr_e, offs);
          if (cv) or_code<W>(stage, p0 + k * ln, cv);
        }
      }
    } else {
      // raw chunk: the original bytes
#pragma unroll
      for (int r = 0; r < kRounds; r++) ((uint4*)stage)[r * kThreads + t] = raw[r];
    }

    if (t < 64) {
      // wave 0, after its share of the packing: the chunk's stream offset
      const uint64_t excl = (c == 0) ? 0ull : look_back<W, LOSSY>(a, st, c, nchunks);
      if (t == 0) {
        if (c != 0) publish_inclusive(st, c, nchunks, excl + (uint64_t)csize);
        s_excl = excl;
      }
    }
    __syncthreads();

    const long long excl = uni((long long)s_excl);
    store_unaligned(a.out + data_base + excl, stage, csize);
    if (t == 0) {
      if (c == 0) *(long long*)a.out = a.insize;
      if (c == nchunks - 1) *a.outsize = data_base + excl + csize;
    }
    __syncthreads();   // stage is cleared for the next chunk only after every store read it
  }
}

// Validity of a stream header against the caller's limits (uniform).
__device__ __forceinline__ bool header_ok(const DecodeArgs& a, long long hdr, long long* nchunks, uint64_t* why)
{
  if (hdr < 0) { *why = kStCorrupt; return false; }
  if (a.cap >= 0 && hdr > a.cap) { *why = kStCapacity; return false; }
  *nchunks = (hdr + kChunk - 1) / kChunk;
  if (a.inlen >= 0 && 8 + 2 * *nchunks > a.inlen) { *why = kStCorrupt; return false; }
  return true;
}

// Exclusive prefix sums of the uint16 size table into the workspace (one
// int64 per chunk plus the total), by one workgroup of 1024 threads: each
// thread takes 32 consecutive entries (8-byte loads), so one pass covers
// 32768 chunks (512 MiB of f32 input) and larger tables loop.  Also writes
// *decsize.
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 32;
__global__ __launch_bounds__(kScanThreads) void offsets_kernel(DecodeArgs a)
{
  __shared__ long long wsum[kScanThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const long long hdr = uni(*(const long long*)a.in);
  long long nchunks = 0;
  uint64_t why = 0;
  if (!header_ok(a, hdr, &nchunks, &why)) {
    if (t == 0) raise_status(a.ws, why);
    return;
  }
  if (t == 0) *a.decsize = hdr;
  const uint16_t* const tab = (const uint16_t*)(a.in + 8);
  long long* const offs = (long long*)(a.ws + kWsFlags);
  long long carry = 0;
  for (long long base = 0; base < nchunks; base += (long long)kScanThreads * kScanPer) {
    const long long i0 = base + (long long)t * kScanPer;
    uint32_t v[kScanPer];
    if (i0 + kScanPer <= nchunks) {
      const unsigned long long* p = (const unsigned long long*)(tab + i0);   // 8-byte aligned
#pragma unroll
      for (int k = 0; k < kScanPer / 4; k++) {
        const unsigned long long w = p[k];
        v[4 * k] = (uint32_t)(w & 0xffff);
        v[4 * k + 1] = (uint32_t)((w >> 16) & 0xffff);
        v[4 * k + 2] = (uint32_t)((w >> 32) & 0xffff);
        v[4 * k + 3] = (uint32_t)(w >> 48);
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanPer; k++) v[k] = (i0 + k < nchunks) ? (uint32_t)tab[i0 + k] : 0u;
    }
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) local += v[k];
    uint32_t inc = local;   // inclusive wave scan of the per-thread sums
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(inc, d, 64);
      if (lane >= d) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    long long before = 0, tile = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; w++) {
      const long long x = wsum[w];
      if (w < wave) before += x;
      tile += x;
    }
    long long run = carry + before + (long long)(inc - local);
    if (i0 + kScanPer <= nchunks) {
      longlong2* o = (longlong2*)(offs + i0);   // 16-byte aligned: i0 is a multiple of 32
#pragma unroll
      for (int k = 0; k < kScanPer / 2; k++) {
        const long long r0 = run;
        run += v[2 * k];
        o[k] = make_longlong2(r0, run);
        run += v[2 * k + 1];
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanPer; k++) {
        if (i0 + k < nchunks) offs[i0 + k] = run;
        run += v[k];
      }
    }
    carry += tile;
    __syncthreads();
  }
  if (t == 0) offs[nchunks] = carry;
}

// Chunk decode: no cross-workgroup dependency once the offsets are known,
// so workgroups stride over chunks in any order.
template <typename W, bool LOSSY>
__global__ __launch_bounds__(kThreads) void decode_kernel(DecodeArgs a)
{
  constexpr int PS = Word<W>::PerSlice;
  constexpr int TB = Word<W>::TB;
  __shared__ uint32_t stage[kStageBytes / 4];

  const int t = threadIdx.x;
  const int lane32 = t & 31;
  const int half = t >> 5;
  const long long hdr = uni(*(const long long*)a.in);
  long long nchunks = 0;
  uint64_t why = 0;
  if (!header_ok(a, hdr, &nchunks, &why)) return;   // reported by offsets_kernel
  const long long data_base = 8 + 2 * nchunks;
  const uint16_t* const sizes = (const uint16_t*)(a.in + 8);
  const long long* const offs_tab = (const long long*)(a.ws + kWsFlags);
  const W offs = (W)a.offs;
  if (a.inlen >= 0 && data_base + uni(offs_tab[nchunks]) > a.inlen) {
    if (blockIdx.x == 0 && t == 0) raise_status(a.ws, kStCorrupt);
    return;
  }

  for (long long c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const long long base = c * kChunk;
    const int osize = (int)min((long long)kChunk, hdr - base);
    const int csize = uni((int)sizes[c]);
    const long long excl = uni(offs_tab[c]);
    const bool bad = csize > osize || (csize < osize && csize < kSub + 2);

    if (!bad) {
      // aligned cover of the payload: payload byte i sits at stage byte mis + i
      const uint8_t* const src = a.in + data_base + excl;
      const int mis = (int)((uintptr_t)src & 15);
      const uint4* const cover = (const uint4*)(src - mis);
      const int nblk = (mis + csize + 15) >> 4;
      for (int j = t; j < nblk; j += kThreads) ((uint4*)stage)[j] = cover[j];
      __syncthreads();

      uint8_t* const dst = a.out + base;
      if (csize == osize) {
#pragma unroll
        for (int r = 0; r < kRounds; r++) {
          const int off = r * kThreads * 16 + t * 16;
          if (off < osize) store_slice(dst, off, osize, lds_read16(stage, mis + off));
        }
      } else {
        // 32 bit-length bytes, realigned from the cover
        uint32_t L[8];
        const int q = mis >> 2;
        const uint32_t sh = (uint32_t)(mis & 3) * 8u;
        bool over = false;
        uint32_t nb = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
          L[i] = uni(__builtin_amdgcn_alignbit(stage[q + i + 1], stage[q + i], sh));
          over |= any_byte_above(L[i], (uint32_t)TB);
          nb += byte_sum(L[i]);
        }
        const uint8_t* sb = (const uint8_t*)stage;
        const int tr = uni((int)(int16_t)((uint16_t)sb[mis + csize - 2] | ((uint16_t)sb[mis + csize - 1] << 8)));
        const bool corrupt = over || (long long)(kSub + 2) + (long long)nb * (Word<W>::PerSub / 8) != csize || tr != osize;
        if (corrupt) {
          if (t == 0) raise_status(a.ws, kStCorrupt);
        } else {
#pragma unroll
          for (int r = 0; r < kRounds; r++) {
            const int off = r * kThreads * 16 + t * 16;
            const int s = r * 8 + half;
            const int ln = sub_length(L, s);
            const int p0 = 8 * (mis + kSub) + sub_bit_offset<W>(L, s) + lane32 * PS * ln;
            W v[PS];
#pragma unroll
            for (int k = 0; k < PS; k++) v[k] = from_code<W, LOSSY>(extract_code(stage, p0 + k * ln, ln, (W*)nullptr), a.thr_e, offs);
            if (off < osize) store_slice(dst, off, osize, values_slice(v));
          }
        }
      }
    } else if (t == 0) {
      raise_status(a.ws, kStCorrupt);
    }
    __syncthreads();
  }
}

// ---- launchers ----

template <typename W, bool LOSSY>
static hipError_t launch_encode(const EncodeArgs& a, int grid, int waves, hipStream_t stream)
{
  if (waves >= 8) hipLaunchKernelGGL((encode_kernel<W, LOSSY, 8>), dim3(grid), dim3(kThreads), 0, stream, a);
  else hipLaunchKernelGGL((encode_kernel<W, LOSSY, 1>), dim3(grid), dim3(kThreads), 0, stream, a);
  return hipGetLastError();
}

template <typename W, bool LOSSY>
static hipError_t launch_decode(const DecodeArgs& a, int grid, hipStream_t stream)
{
  hipLaunchKernelGGL(offsets_kernel, dim3(1), dim3(kScanThreads), 0, stream, a);
  hipLaunchKernelGGL((decode_kernel<W, LOSSY>), dim3(grid), dim3(kThreads), 0, stream, a);
  return hipGetLastError();
}

hipError_t encode_launch(int elem_bytes, bool lossy, const EncodeArgs& a, int grid, int waves, hipStream_t stream)
{
  if (elem_bytes == 4) return lossy ? launch_encode<uint32_t, true>(a, grid, waves, stream) : launch_encode<uint32_t, false>(a, grid, waves, stream);
  return lossy ? launch_encode<uint64_t, true>(a, grid, waves, stream) : launch_encode<uint64_t, false>(a, grid, waves, stream);
}

hipError_t decode_launch(int elem_bytes, bool lossy, const DecodeArgs& a, int grid, hipStream_t stream)
{
  if (elem_bytes == 4) return lossy ? launch_decode<uint32_t, true>(a, grid, stream) : launch_decode<uint32_t, false>(a, grid, stream);
  return lossy ? launch_decode<uint64_t, true>(a, grid, stream) : launch_decode<uint64_t, false>(a, grid, stream);
}

template <typename W, bool LOSSY>
static const void* encode_fn(int waves)
{
  return waves >= 8 ? (const void*)encode_kernel<W, LOSSY, 8> : (const void*)encode_kernel<W, LOSSY, 1>;
}

int resident_grid(int elem_bytes, bool lossy, bool decode, int waves)
{
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const void* fn;
  if (decode) {
    fn = elem_bytes == 4 ? (lossy ? (const void*)decode_kernel<uint32_t, true> : (const void*)decode_kernel<uint32_t, false>)
                         : (lossy ? (const void*)decode_kernel<uint64_t, true> : (const void*)decode_kernel<uint64_t, false>);
  } else {
    fn = elem_bytes == 4 ? (lossy ? encode_fn<uint32_t, true>(waves) : encode_fn<uint32_t, false>(waves))
                         : (lossy ? encode_fn<uint64_t, true>(waves) : encode_fn<uint64_t, false>(waves));
  }
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, kThreads, 0) != hipSuccess) return 0;
  return cus * (per > 0 ? per : 1);
}

}  // namespace sleek
