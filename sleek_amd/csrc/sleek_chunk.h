// Per-chunk device code shared by the single-array and the batched kernels:
// byte movers, the encoder's look-back and one-wave chunk coder, the
// decoder's offset scan and chunk decoder.
#pragma once

#include "sleek_device.h"
#include "sleek_internal.h"

namespace sleek {

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

// NT: the non-temporal hint on whole slices (see ntload)
template <bool NT = false>
__device__ __forceinline__ void store_slice(uint8_t* dst, int off, int osize, uint4 q)
{
  if (off + 16 <= osize) {
    if (NT) ntstore((uint4*)(dst + off), q);
    else *(uint4*)(dst + off) = q;
    return;
  }
  const uint32_t d[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (off + 4 * i + 4 <= osize) ((uint32_t*)(dst + off))[i] = d[i];
}

// The 8-byte slice at byte off of a chunk of osize bytes (osize is a
// multiple of 4), zero beyond osize.  Branch-free: an invalid slice is
// loaded from the chunk's first slice (always inside the input) and
// zeroed, so the loads of a partial chunk are all in flight at once
// instead of one conditional load after another.
__device__ __forceinline__ uint2 load_lane(const uint8_t* src, int off, int osize)
{
  const bool ok = off < osize;
  uint2 q = *(const uint2*)(src + (ok ? off : 0));
  // masks the compiler cannot see through keep the load unconditional
  // (a known zero/all-ones mask would let it sink the load into a branch)
  uint32_t mx = 0u - (uint32_t)ok, my = 0u - (uint32_t)(osize - off >= 8);
  asm volatile("" : "+v"(mx), "+v"(my));
  q.x &= mx;
  q.y &= my;
  return q;
}

// Slice s (stride sub_stride 8-byte words) of a lane whose chunk has rem
// valid bytes from the lane's first slice on; zero where invalid.  Written
// with one lane pointer and constant offsets so that no per-slice offset is
// kept in registers across the chunk loop; an invalid slice is loaded from
// safe (the chunk's first slice) and zeroed, without a branch.
__device__ __forceinline__ uint2 load_lane_rem(const uint2* lp, const uint2* safe, int s, int sub_stride, int rem)
{
  const int b = s * sub_stride * 8;
  const bool ok = rem > b;
  uint2 q = *(ok ? lp + s * sub_stride : safe);
  uint32_t mx = 0u - (uint32_t)ok, my = 0u - (uint32_t)(rem - b >= 8);
  asm volatile("" : "+v"(mx), "+v"(my));   // opaque masks: see load_lane
  q.x &= mx;
  q.y &= my;
  return q;
}

template <bool NT = false>
__device__ __forceinline__ void store_lane(uint8_t* dst, int off, int osize, uint2 q)
{
  if (off + 8 <= osize) {
    if (NT) ntstore((uint2*)(dst + off), q);
    else *(uint2*)(dst + off) = q;
  } else if (off < osize) {
    *(uint32_t*)(dst + off) = q.x;
  }
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
  for (int s = 0; s < kSub; s++) {
    W v[Word<W>::PerLane];
    lane_values(load_lane(src, s * kSubBytes + lane * 8, osize), v);
    W orv = 0;
#pragma unroll
    for (int k = 0; k < Word<W>::PerLane; k++) orv |= to_code<W, LOSSY>(v[k], a.eb_e, a.thr_e, (W)a.offs);
    nb += (uint32_t)wave_bit_length(orv);
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
  const long long first = gj << s.gshift;
  const int members = (int)group_members(s, gj, nchunks);
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

// One help event (a look-back sized a missing word itself) on the device's
// counter (help[0]), if there is one.
__device__ __forceinline__ void count_help(unsigned long long* help)
{
  if (help) __hip_atomic_fetch_add(help, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The look-back's help threshold in naps (help[1]; the default without a
// help state), read only when a wait has gone on a while.
__device__ __forceinline__ uint32_t help_naps(const unsigned long long* help)
{
  return help ? (uint32_t)help[1] : kHelpNapsDefault;
}

// Two-level decoupled look-back, run by one whole wave; returns the sum of
// the compressed sizes of chunks 0..c-1 in every lane.  First the preceding
// chunks of c's own group are watched (one lane each); once they are all
// published, a window of earlier groups is watched, starting a.lb_window
// groups wide and doubling up to 64 as it slides back, so the inclusive
// frontier can advance up to 64 groups per round trip while a poll reads
// few words.  Groups hold 2^gshift chunks (8..64).
// Polls back off exponentially: the words live at device scope, and many
// waiting waves polling at full rate would compete with the data traffic.
// Workgroups take chunks in launch or ticket order, so the words waited on
// are normally published within a few round trips.  After help_naps() naps
// without progress (about 100 us of the wave's own sleep by default; time
// the wave spends suspended with its queue does not count) the waiter
// computes missing words from the input and publishes them, one after
// another without further waits, until others' words arrive again; so the
// look-back finishes even when a predecessor's workgroup is not running (an
// encoder with chunks in a static order whose grid exceeds what is
// resident).  Each word sized so is counted in a.help[0]: the ticket encoders
// hand chunks only to running workgroups and should never need it.
template <typename W, bool LOSSY>
__device__ uint64_t look_back(const EncodeArgs& a, const ScanState& s, long long c, long long nchunks)
{
  // the per-lane word addresses use opaque(lane): recomputed at each poll (a
  // few instructions) rather than kept live through the encoders' loops,
  // whose registers are all in use
  const int lane = threadIdx.x & 63;
  const int q = (int)(c & ((1ll << s.gshift) - 1));   // chunks of the group before c
  long long gend = c >> s.gshift;                       // group window: groups gend-gwin .. gend-1
  int gwin = a.lb_window;
  bool have_partial = (q == 0);
  uint64_t partial = 0, gacc = 0;
  uint32_t slept = 0;   // naps since the last progress
  int naps = 1;
  bool helping = false;   // the wait timed out: size missing words from the input, one after another
  while (true) {
    bool progressed = false, helped = false;
    if (!have_partial) {
      const uint64_t f = (lane < q) ? flag_load(s.chunk + (c - 1 - opaque(lane))) : kAgg;   // beyond the group: 0
      const uint64_t inc = __ballot((f >> 62) == 2);
      const uint64_t rdy = __ballot((f >> 62) != 0);
      if (inc) {
        const int p = __ffsll((unsigned long long)inc) - 1;
        const uint64_t need = (2ull << p) - 1;      // p < 63 here
        if ((rdy & need) == need) {
          const uint64_t r = wave_sum(lane <= p ? (f & kValMask) : 0);
          wait_vm_drained();   // free: the polls just used were issued last; nothing of the look-back stays pending
          return r;
        }
      } else if (rdy == ~0ull) {
        partial = wave_sum(f & kValMask);
        have_partial = true;
        progressed = true;
      }
      if (!progressed && (helping || slept >= help_naps(a.help))) {
        const uint64_t miss = __ballot(lane < q && (f >> 62) == 0);
        if (miss) {
          const long long jc = c - __ffsll((unsigned long long)miss);
          const uint32_t sz = wave_chunk_size<W, LOSSY>(a, jc);
          if (lane == 0) {
            flag_max(s.chunk + jc, kAgg | sz);
            count_help(a.help);
          }
        }
        helping = helped = true;
      }
    } else {
      // a group's word, or its completion count once every member has
      // reported (the count's bytes are then the group's aggregate)
      const long long gj = gend - 1 - opaque(lane);
      uint64_t fg = kAgg;                      // lanes beyond the window: 0
      if (lane < gwin) {
        if (gj < 0) {
          fg = kInc;                           // before group 0: inclusive 0
        } else {
          fg = flag_load(s.group + gj);
          const uint64_t n = flag_load((const uint64_t*)(s.count + gj));
          if ((fg >> 62) == 0 && (long long)(n >> 32) == group_members(s, gj, nchunks))
            fg = (gj == 0 ? kInc : kAgg) | (n & 0xffffffffull);
        }
      }
      const uint64_t inc = __ballot((fg >> 62) == 2);
      const uint64_t rdy = __ballot((fg >> 62) != 0);
      if (inc) {
        const int p = __ffsll((unsigned long long)inc) - 1;
        const uint64_t need = (2ull << p) - 1;      // p < gwin <= 64; lanes >= gwin carry 0
        if ((rdy & need) == need) {
          const uint64_t r = partial + gacc + wave_sum(lane <= p ? (fg & kValMask) : 0);
          wait_vm_drained();   // as above
          return r;
        }
      } else if (rdy == ~0ull) {
        gacc += wave_sum(fg & kValMask);
        gend -= gwin;
        gwin = min(2 * gwin, 64);
        progressed = true;
      }
      if (!progressed && (helping || slept >= help_naps(a.help))) {
        const uint64_t miss = __ballot((fg >> 62) == 0);
        if (miss) {
          const long long gm = gend - __ffsll((unsigned long long)miss);
          const uint64_t sz = wave_group_size<W, LOSSY>(a, s, gm, nchunks);
          if (lane == 0) {
            flag_max(s.group + gm, (gm == 0 ? kInc : kAgg) | sz);
            count_help(a.help);
          }
        }
        helping = helped = true;
      }
    }
    if (progressed) {   // others' words arrived: wait for them again rather than help
      slept = 0;
      naps = 1;
      helping = false;
      continue;
    }
    if (helped) continue;
    for (int i = 0; i < naps; i++) __builtin_amdgcn_s_sleep(2);   // 128 clocks each
    slept += naps;
    if (naps < 32) naps *= 2;
  }
}

// ---- one-wave chunk coder ----
// A single wave codes a chunk.  Lane l holds the 8-byte slice l of each of
// the chunk's 32 subchunks (64 VGPRs), so every per-subchunk quantity is
// wave-uniform.  The slices are turned into codes in place once; packing
// reads the codes.

// Issues the loads of chunk c's raw slices (64 VGPRs).  A slice beyond the
// input (last chunk only) is loaded from the chunk's first slice instead,
// so every load is unconditional and all 32 are in flight at once; such
// slices hold garbage until mask_tail zeroes them.
__device__ __forceinline__ void load_raw(const EncodeArgs& a, long long c, uint2 raw[kSub])
{
  const int lane = threadIdx.x & 63;
  const long long base = c * kChunk;
  const int osize = (int)min((long long)kChunk, a.insize - base);
  const uint2* const p = (const uint2*)(a.in + base) + lane;
  if (osize == kChunk) {
    // full chunk: one pointer, constant offsets, no masking
#pragma unroll
    for (int s = 0; s < kSub; s++) raw[s] = ntload(p + s * (kSubBytes / 8));   // streamed: see ntload
  } else {
    // an opaque copy of the lane's valid byte count chooses the addresses:
    // the compiler cannot tie it to the masking, so it cannot turn the
    // loads into conditional ones
    int rem = osize - lane * 8;
    asm volatile("" : "+v"(rem));
    const uint2* const safe = (const uint2*)(a.in + base);
#pragma unroll
    for (int s = 0; s < kSub; s++) raw[s] = ntload(rem > s * kSubBytes ? p + s * (kSubBytes / 8) : safe);
    wait_vm_drained();
  }
}

// Zeroes the bytes of chunk c's slices that lie beyond the input.
__device__ __forceinline__ void mask_tail(const EncodeArgs& a, long long c, uint2 raw[kSub])
{
  const int osize = (int)min((long long)kChunk, a.insize - c * kChunk);
  if (osize < kChunk) {
    // opaque, so the compiler derives the slices' limits from it here rather
    // than keeping 32 per-lane constants live across the whole kernel
    int rem = osize - (int)(threadIdx.x & 63) * 8;
    asm volatile("" : "+v"(rem));
#pragma unroll
    for (int s = 0; s < kSub; s++) {
      const int r = rem - s * kSubBytes;
      raw[s].x = r > 0 ? raw[s].x : 0u;
      raw[s].y = r >= 8 ? raw[s].y : 0u;
    }
  }
}

// What the stream holds for one chunk: bit lengths (byte-packed, uniform),
// payload bits, input bytes and stored bytes (csize < osize: packed).
struct ChunkShape {
  uint32_t L[8];
  int total_bits;
  int osize;
  int csize;
};

// Codes of the chunk's values, in place (slices as load_raw left them),
// and the chunk's shape.  BFLY: the subchunks' bit lengths by one
// transposing butterfly over all 32 (32 more VGPRs live at its start), else
// subchunk by subchunk (for kernels with no registers to spare).
template <typename W, bool LOSSY, bool BFLY = true>
__device__ __forceinline__ ChunkShape shape_codes(const EncodeArgs& a, long long c, uint2 slot[kSub])
{
  constexpr int PL = Word<W>::PerLane;
  const W offs = (W)a.offs;
  ChunkShape sh;
  sh.osize = (int)min((long long)kChunk, a.insize - c * kChunk);
  mask_tail(a, c, slot);
  uint32_t ov[kSub];
#pragma unroll
  for (int i = 0; i < 8; i++) sh.L[i] = 0;
#pragma unroll
  for (int s = 0; s < kSub; s++) {
    W v[PL];
    lane_values(slot[s], v);
    W orv = 0;
#pragma unroll
    for (int k = 0; k < PL; k++) {
      v[k] = to_code<W, LOSSY>(v[k], a.eb_e, a.thr_e, offs);
      orv |= v[k];
    }
    slot[s] = values_lane(v);
    if constexpr (BFLY) ov[s] = lane_length_word(orv);
    else sh.L[s / 4] |= (uint32_t)wave_bit_length(orv) << (8 * (s % 4));
  }
  if constexpr (BFLY) wave_lengths<kSub, sizeof(W) == 8>(ov, sh.L);
  uint32_t nb = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) nb += byte_sum(sh.L[i]);
  sh.total_bits = (int)nb * Word<W>::PerSub;
  const int newsize = kSub + 2 + sh.total_bits / 8;
  sh.csize = newsize < sh.osize ? newsize : sh.osize;
  return sh;
}

// Publishes chunk c's size (lane 0): its look-back word, its size-table
// entry and its group's completion count.  The count's reply is not waited
// for: look_back reads a complete count as the group's aggregate, so the
// aggregate is visible as soon as the last member's add lands.
__device__ __forceinline__ void publish_chunk(const EncodeArgs& a, const ScanState& st, long long c, long long nchunks,
                                              int csize)
{
  (void)nchunks;
  if ((threadIdx.x & 63) == 0) {
    flag_max(st.chunk + c, (c == 0 ? kInc : kAgg) | (uint64_t)csize);
    gstore((uint16_t*)(a.out + 8) + c, (uint16_t)csize);
    // no reply waited for: a complete count is read as the group's aggregate by look_back
    __hip_atomic_fetch_add(st.count + (c >> st.gshift), (1ull << 32) | (unsigned)csize, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The chunk's payload into the wave's LDS buffer: the packed form (32
// bit-length bytes, LSB-first fields, int16 trailer) from the codes, or
// the raw bytes: lossless codes invert exactly, lossy ones do not, so a
// raw lossy chunk is read again from the input.
template <typename W, bool LOSSY>
__device__ __forceinline__ void pack_codes(const EncodeArgs& a, long long c, const uint2 code[kSub],
                                           const ChunkShape& sh, uint32_t* stage)
{
  constexpr int PL = Word<W>::PerLane;
  const int lane = threadIdx.x & 63;
  if (sh.csize < sh.osize) {
    // clear the payload's dwords (rounded to 16 bytes), then OR the fields in
    for (int i = lane; i < (sh.csize + 15) / 16; i += 64) ((uint4*)stage)[i] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    if (lane == 0) {   // the 32 bit-length bytes (uniform values: no per-lane indexing)
      ((uint4*)stage)[0] = make_uint4(sh.L[0], sh.L[1], sh.L[2], sh.L[3]);
      ((uint4*)stage)[1] = make_uint4(sh.L[4], sh.L[5], sh.L[6], sh.L[7]);
    }
    int p0 = 8 * kSub;
#pragma unroll
    for (int s = 0; s < kSub; s++) {
      const int ln = sub_length(sh.L, s);
      if (ln) {
        W v[PL];
        lane_values(code[s], v);
        or_field(stage, p0 + lane * PL * ln, lane_field(v, ln));
      }
      p0 += ln * Word<W>::PerSub;
    }
    __syncthreads();
    if (lane == 0) stage[8 + sh.total_bits / 32] = (uint32_t)sh.osize;   // the trailer's dword follows the fields
  } else if constexpr (LOSSY) {
    // rare: streamed through a few registers at a time
    const uint2* const safe = (const uint2*)(a.in + c * kChunk);
    const uint2* const p = safe + lane;
    const int rem = sh.osize - lane * 8;
#pragma unroll 4
    for (int s = 0; s < kSub; s++)
      ((uint2*)stage)[lane + s * (kSubBytes / 8)] = load_lane_rem(p, safe, s, kSubBytes / 8, rem);
  } else {
#pragma unroll
    for (int s = 0; s < kSub; s++) {
      W v[PL];
      lane_values(code[s], v);
#pragma unroll
      for (int k = 0; k < PL; k++) v[k] = lossless_value<W>(v[k]);
      ((uint2*)stage)[lane + s * (kSubBytes / 8)] = values_lane(v);
    }
  }
  __syncthreads();
}

// Chunk c's inclusive word (excl: its exclusive offset) and the store of its
// payload (plus the header for chunk 0 and *outsize for the last chunk).
__device__ __forceinline__ void finish_chunk(const EncodeArgs& a, const ScanState& st, long long c, long long nchunks,
                                             int csize, const uint32_t* stage, uint64_t excl)
{
  const int lane = threadIdx.x & 63;
  const long long data_base = 8 + 2 * nchunks;
  if (lane == 0 && c != 0) publish_inclusive(st, c, nchunks, excl + (uint64_t)csize);
  wave_store_unaligned(a.out + data_base + (long long)excl, stage, csize);
  if (lane == 0) {
    if (c == 0) gstore((long long*)a.out, a.insize);
    if (c == nchunks - 1) gstore(a.outsize, data_base + (long long)excl + csize);
  }
  __syncthreads();   // stage read by the stores before it is reused
}

// ---- two-wave chunk coder ----
// Wave h of a 128-thread workgroup holds subchunks 16h .. 16h+15 of the
// chunk (slice l of each in lane l, 32 VGPRs per chunk), half the
// registers of the one-wave coder.  The halves' packed bits meet at a
// subchunk boundary, which is 8-byte aligned inside the payload, so each
// wave clears and packs its own qwords of the staging buffer.
constexpr int kHalfSub = kSub / 2;

// ---- chunk tickets (encode_pair_kernel<.., TICKET>) ----
// S = 2^shift counters, the largest power of two <= min(kTicketSets, grid)
// (ticket_shift_for, computed by the host), so that every set has a
// workgroup of its own (its home set, blockIdx.x mod S); ticket t of set x
// is chunk S * t + x.  A workgroup takes every chunk, its first included,
// from the counters, so a chunk is only ever held by a workgroup that runs.
// The home set is recomputed from blockIdx where it is used (a mask) rather
// than kept live across the encoder's loop, whose registers are all in use.
__device__ __forceinline__ int ticket_home(int shift) { return (int)(blockIdx.x & ((1u << shift) - 1u)); }

// Tickets of set x: chunks S * t + x below nchunks (nchunks < 2^32).
__device__ __forceinline__ uint32_t ticket_limit(uint32_t nchunks, int x, int shift)
{
  return nchunks > (uint32_t)x ? ((nchunks - (uint32_t)x - 1u) >> shift) + 1u : 0u;
}

// A ticket of the workgroup's home set, in lane 0; the other lanes keep
// `prev` (the previous ticket, in the register the atomic writes: a value of
// their own, not undefined, and no instruction that writes or reads the
// register before the atomic, which would wait for the previous atomic and
// so for every memory operation in flight).  The result is left in flight:
// it is read by ticket_chunk.  The address is made opaque so that the
// compiler does not turn the lane's atomic into a wave-wide one, whose
// result it would wait for at once.  Callers issue it before loads they wait
// for early (memory operations complete in issue order as far as those waits
// are concerned).
__device__ __forceinline__ uint32_t ticket_request(uint32_t* tk, int shift, uint32_t prev, uint32_t one = 1u)
{
  uint32_t t = prev;
  if ((threadIdx.x & 63) == 0) {
    uintptr_t p = (uintptr_t)(tk + ticket_home(shift) * kTicketStride);
    asm volatile("" : "+v"(p));
    t = __hip_atomic_fetch_add((__attribute__((address_space(1))) uint32_t*)p, one, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
  }
  return t;
}

// The increment of the encoders' in-loop ticket requests (1), in a register
// of its own defined before the loop: otherwise the compiler materialises it
// in the ticket's own register just before the atomic, and that write waits
// for every memory operation in flight (the previous chunk's stores).
__device__ __forceinline__ uint32_t ticket_one()
{
  uint32_t one = 1u;
  asm volatile("" : "+v"(one));
  return one;
}

// Inlined: as a call it cost far more than its registers, because after a
// call the compiler takes every register the callee may clobber as written
// by a load still in flight, and carried around the encoders' loops that
// put a wait for all memory in flight (the next chunk's prefetch included)
// before later writes of those registers.
// Once the workgroup's home set is spent: a ticket of the first other set
// whose counter still has tickets (all others read at once, one per lane), or
// -1 when every set is spent.  Counters only grow, so a set seen spent stays
// spent and each ticket below a limit goes to exactly one workgroup: every
// chunk is taken.  A set whose ticket came back spent is not tried again, so
// the loop runs at most S - 1 times whatever the reads return.
__device__ __forceinline__ long long ticket_fallback(uint32_t* tk, int shift, long long nchunks)
{
  typedef __attribute__((address_space(1))) uint32_t gu32;
  gu32* const g = (gu32*)tk;
  const int lane = threadIdx.x & 63;
  const int S = 1 << shift;
  const int home = ticket_home(shift);
  uint64_t spent = 1ull << home;
  while (true) {
    const int y = (home + 1 + lane) & (S - 1);
    bool open = false;
    if (lane < S - 1 && !((spent >> y) & 1ull))
      open = __hip_atomic_load(g + y * kTicketStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
             ticket_limit((uint32_t)nchunks, y, shift);
    const uint64_t m = __ballot(open);
    if (m == 0) return -1;
    const int z = (home + 1 + __builtin_ctzll(m)) & (S - 1);
    uint32_t u = 0;
    if (lane == 0) u = __hip_atomic_fetch_add(g + z * kTicketStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    u = uni(u);
    if (u < ticket_limit((uint32_t)nchunks, z, shift)) return ((long long)u << shift) + z;
    spent |= 1ull << z;
  }
}

// A workgroup's first chunk: a ticket of its home set (uniform), or -1 when
// that set is already spent, and the workgroup then has nothing to do (the
// chunks of every other set are taken by their own workgroups and their
// fallbacks).  No fallback here: a third inlined copy of it costs the
// encoders' loops registers.
__device__ __forceinline__ long long ticket_first(uint32_t* tk, int shift, long long nchunks)
{
  const int home = ticket_home(shift);
  const uint32_t t = uni(ticket_request(tk, shift, 0u));
  return t < ticket_limit((uint32_t)nchunks, home, shift) ? ((long long)t << shift) + home : -1;
}

// The chunk of a ticket from ticket_request (uniform), or -1 once every set
// is spent.
__device__ __forceinline__ long long ticket_chunk(uint32_t* tk, int shift, uint32_t t, long long nchunks)
{
  const int home = ticket_home(shift);
  t = uni(t);
  if (t < ticket_limit((uint32_t)nchunks, home, shift)) return ((long long)t << shift) + home;
  return ticket_fallback(tk, shift, nchunks);
}

// The wave's half of chunk c's slices, streamed (ntload), through
// global-address-space pointers also where the array pointer was read from
// memory (a batch item), whose accesses would otherwise be flat.
__device__ __forceinline__ void load_half(const EncodeArgs& a, long long c, int h, uint2 raw[kHalfSub])
{
  // opaque: the lane's offset is recomputed here rather than kept live (or
  // spilled: a reload here would wait for every memory operation in flight)
  const int lane = opaque(threadIdx.x & 63);
  const long long base = c * kChunk;
  const int osize = (int)min((long long)kChunk, a.insize - base);
  const uint2* const p = (const uint2*)(a.in + base + h * kHalfSub * kSubBytes) + lane;
  if (osize == kChunk) {
#pragma unroll
    for (int s = 0; s < kHalfSub; s++) raw[s] = ntload(p + s * (kSubBytes / 8));   // streamed: see ntload
  } else {
    int rem = osize - h * kHalfSub * kSubBytes - lane * 8;   // opaque: see load_raw
    asm volatile("" : "+v"(rem));
    const uint2* const safe = (const uint2*)(a.in + base);
#pragma unroll
    for (int s = 0; s < kHalfSub; s++) {
      const uint2* const q = rem > s * kSubBytes ? p + s * (kSubBytes / 8) : safe;
      raw[s] = ntload(q);
    }
    wait_vm_drained();
  }
}

// Codes of the wave's half in place (tail zeroed first); returns the half's
// 16 bit-length bytes.
template <typename W, bool LOSSY>
__device__ __forceinline__ void shape_half(const EncodeArgs& a, long long c, int h, uint2 slot[kHalfSub], uint32_t Lh[4])
{
  constexpr int PL = Word<W>::PerLane;
  const W offs = (W)a.offs;
  const int osize = (int)min((long long)kChunk, a.insize - c * kChunk);
  if (osize < kChunk) {
    // opaque, so the slices' limits are derived here rather than kept live
    // (or spilled) across the encoders' loops as 16 per-lane constants
    int rem = osize - h * kHalfSub * kSubBytes - (int)(threadIdx.x & 63) * 8;
    asm volatile("" : "+v"(rem));
#pragma unroll
    for (int s = 0; s < kHalfSub; s++) {
      const int r = rem - s * kSubBytes;
      slot[s].x = r > 0 ? slot[s].x : 0u;
      slot[s].y = r >= 8 ? slot[s].y : 0u;
    }
  }
  uint32_t ov[kHalfSub];
  if constexpr (LOSSY && sizeof(W) == 4) {
    const LossyF32 p = lossy_f32_params(a.thr_e, (uint32_t)a.offs);
#pragma unroll
    for (int s = 0; s < kHalfSub; s++) {
      const uint32_t c0 = lossy_code_f32(slot[s].x, p), c1 = lossy_code_f32(slot[s].y, p);
      slot[s] = make_uint2(c0, c1);
      ov[s] = c0 | c1;
    }
    wave_lengths<kHalfSub, false>(ov, Lh);
    return;
  }
#pragma unroll
  for (int s = 0; s < kHalfSub; s++) {
    W v[PL];
    lane_values(slot[s], v);
    W orv = 0;
#pragma unroll
    for (int k = 0; k < PL; k++) {
      v[k] = to_code<W, LOSSY>(v[k], a.eb_e, a.thr_e, offs);
      orv |= v[k];
    }
    slot[s] = values_lane(v);
    ov[s] = lane_length_word(orv);
  }
  wave_lengths<kHalfSub, sizeof(W) == 8>(ov, Lh);
}

// The chunk's shape from its 32 bit-length bytes (exchanged through LDS).
template <typename W>
__device__ __forceinline__ ChunkShape shape_of(const EncodeArgs& a, long long c, const uint32_t* xl)
{
  ChunkShape sh;
  sh.osize = (int)min((long long)kChunk, a.insize - c * kChunk);
  uint32_t nb = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    sh.L[i] = uni(xl[i]);
    nb += byte_sum(sh.L[i]);
  }
  sh.total_bits = (int)nb * Word<W>::PerSub;
  const int newsize = kSub + 2 + sh.total_bits / 8;
  sh.csize = newsize < sh.osize ? newsize : sh.osize;
  return sh;
}

// The wave's half of the payload into the shared staging buffer (each half
// starts and ends on a dword boundary): lossless fields, mostly 50-64 bits
// wide, written with put_codes, which needs no clearing pass; lossy ones,
// mostly a few bits wide, OR-ed into the cleared half with or_codes, whose
// one or two ORs cost less there than put_codes' extra instructions (each
// measured faster for its mode).  Wave 0 writes the bit-length bytes, wave 1
// the trailer.  Raw chunks: each wave its own half of the bytes.  No barrier
// inside: the waves' regions are disjoint.
template <typename W, bool LOSSY>
__device__ __forceinline__ void pack_half(const EncodeArgs& a, long long c, int h, const uint2 code[kHalfSub],
                                          const ChunkShape& sh, uint32_t* stage)
{
  constexpr int PL = Word<W>::PerLane;
  const int lane = threadIdx.x & 63;
  if (sh.csize < sh.osize) {
    const int mid = 8 * kSub + sub_bit_offset<W>(sh.L, kHalfSub);   // first bit of subchunk 16
    if constexpr (LOSSY) {   // narrow fields: OR-ed into a cleared buffer (see put_codes)
      const int q0 = h == 0 ? 4 : mid / 64;
      const int q1 = h == 0 ? mid / 64 : (8 * kSub + sh.total_bits) / 64 + 1;   // wave 1: through the trailer
      for (int q = q0 + lane; q < q1; q += 64) ((uint2*)stage)[q] = make_uint2(0u, 0u);
    }
    if (h == 0 && lane == 0) {
      ((uint4*)stage)[0] = make_uint4(sh.L[0], sh.L[1], sh.L[2], sh.L[3]);
      ((uint4*)stage)[1] = make_uint4(sh.L[4], sh.L[5], sh.L[6], sh.L[7]);
    }
    uint32_t lw[4];   // this half's bit lengths (h is uniform: no per-lane indexing)
#pragma unroll
    for (int i = 0; i < 4; i++) lw[i] = h == 0 ? sh.L[i] : sh.L[4 + i];
    int p0 = h == 0 ? 8 * kSub : mid;
#pragma unroll
    for (int s = 0; s < kHalfSub; s++) {
      const int ln = sub_length(lw, s);   // uniform (h is)
      if (ln) {
        W v[PL];
        lane_values(code[s], v);
        if constexpr (LOSSY) or_codes(stage, p0, lane, ln, v);
        else put_codes(stage, p0, lane, ln, v);
      }
      p0 += ln * Word<W>::PerSub;
    }
    if (h == 1 && lane == 0) stage[8 + sh.total_bits / 32] = (uint32_t)sh.osize;   // the dword after the payload
  } else if constexpr (LOSSY) {
    const uint2* const safe = (const uint2*)(a.in + c * kChunk);
    const uint2* const p = safe + h * kHalfSub * (kSubBytes / 8) + lane;
    const int rem = sh.osize - h * kHalfSub * kSubBytes - lane * 8;
#pragma unroll 4
    for (int s = 0; s < kHalfSub; s++)
      ((uint2*)stage)[lane + (h * kHalfSub + s) * (kSubBytes / 8)] = load_lane_rem(p, safe, s, kSubBytes / 8, rem);
  } else {
#pragma unroll
    for (int s = 0; s < kHalfSub; s++) {
      W v[PL];
      lane_values(code[s], v);
#pragma unroll
      for (int k = 0; k < PL; k++) v[k] = lossless_value<W>(v[k]);
      ((uint2*)stage)[lane + (h * kHalfSub + s) * (kSubBytes / 8)] = values_lane(v);
    }
  }
}

// The lines of pair_store for a head whose dword part is Q (uniform, so the
// dword selection is static): lines j = 64 h + lane + 128 i, two per trip
// with their four 128-bit reads waited for together.
template <int Q>
__device__ __forceinline__ void pair_store_lines(uint8_t* dst, uint32_t lds_base, int head, int nfull, int h,
                                                 uint32_t sh)
{
  const int lane = threadIdx.x & 63;
  int j = h * 64 + lane;
  for (; j + 128 < nfull; j += 256) {
    u32x4 a0, b0, a1, b1;
    lds_read_lines4(lds_base + (uint32_t)(j << 4), lds_base + (uint32_t)((j + 128) << 4), a0, b0, a1, b1);
    gstore((uint4*)(dst + head + (j << 4)), line_window<Q>(a0, b0, sh));
    gstore((uint4*)(dst + head + ((j + 128) << 4)), line_window<Q>(a1, b1, sh));
  }
  if (j < nfull) {
    u32x4 a0, b0;
    lds_read_lines2(lds_base + (uint32_t)(j << 4), a0, b0);
    gstore((uint4*)(dst + head + (j << 4)), line_window<Q>(a0, b0, sh));
  }
}

// len bytes of the staging buffer to dst (any byte address) by both waves:
// 16-byte lines alternate between the waves, wave 0 stores the head bytes
// before the first 16-byte boundary, wave 1 the tail bytes.  Line j of the
// output is staging bytes head + 16 j .. + 15, inside staging lines j and
// j + 1 (head < 16), which are read whole: reads of single dwords 16 bytes
// apart across the wave would conflict 8-way.
__device__ __forceinline__ void pair_store(uint8_t* dst, const uint32_t* lds, int len, int h)
{
  const int lane = threadIdx.x & 63;
  const int mis = (int)((uintptr_t)dst & 15);
  const int head = min(len, (16 - mis) & 15);
  const int nfull = (len - head) >> 4;
  const int tail = len - head - (nfull << 4);
  const uint8_t* lb = (const uint8_t*)lds;
  const uint32_t base = lds_addr(lds);
  const uint32_t sh = (uint32_t)(head & 3) * 8u;
  switch (uni(head >> 2)) {
    case 0: pair_store_lines<0>(dst, base, head, nfull, h, sh); break;
    case 1: pair_store_lines<1>(dst, base, head, nfull, h, sh); break;
    case 2: pair_store_lines<2>(dst, base, head, nfull, h, sh); break;
    default: pair_store_lines<3>(dst, base, head, nfull, h, sh); break;
  }
  // the edge bytes (disjoint from the lines: any order is correct)
  if (h == 0 && lane < head) gstore(dst + lane, lb[lane]);
  if (h == 1 && lane < tail) {
    const int i = head + (nfull << 4) + lane;
    gstore(dst + i, lb[i]);
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

// Exclusive prefix sums of a stream's uint16 size table into the workspace
// (one int64 per chunk plus the total), by `parts` workgroups of
// kScanThreads threads that need nothing from one another: the table's
// rounds of kScanRound entries are split into parts of whole rounds, and
// part p first sums every entry before its rounds itself (table reads that
// hit L2) and then scans its rounds from that carry.  Part 0 also writes
// *decsize, clears the status word and reports a bad header in it.
constexpr int kScanThreads = 1024;
constexpr int kScanRows = 2;                                           // rows of 256 entries per wave and round
constexpr long long kScanBlock = 256 * kScanRows;                      // entries per wave and round
constexpr long long kScanRound = kScanBlock * (kScanThreads / 64);     // entries per round
constexpr int kScanPartsMax = 64;                                      // workgroups per offsets launch

// Sum of the table's first n entries (n a multiple of 4) by the whole
// workgroup: 8-byte loads, a wave sum each, then the waves' sums.
__device__ __forceinline__ long long table_prefix(const uint16_t* tab, long long n, long long* wsum)
{
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  uint64_t acc = 0;
  for (long long i = 4ll * t; i < n; i += 4ll * kScanThreads) {
    const unsigned long long w = *(const unsigned long long*)(tab + i);   // 8-byte aligned: i is a multiple of 4
    acc += (w & 0xffff) + ((w >> 16) & 0xffff) + ((w >> 32) & 0xffff) + (w >> 48);
  }
  acc = wave_sum(acc);
  if (lane == 0) wsum[wave] = (long long)acc;
  __syncthreads();
  long long tot = 0;
#pragma unroll
  for (int v = 0; v < kScanThreads / 64; v++) tot += wsum[v];
  __syncthreads();   // wsum is reused by the rounds
  return tot;
}

__device__ __forceinline__ void scan_offsets(const DecodeArgs& a, long long* wsum, int part, int parts)
{
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // this launch opens the call: it clears the status word (the later
  // kernels of the call only OR bits into it), so the call needs no
  // separate clearing launch
  if (part == 0 && t == 0) flag_store(a.ws + kWsStatus, 0);
  const long long hdr = uni(*(const long long*)a.in);
  long long nchunks = 0;
  uint64_t why = 0;
  if (!header_ok(a, hdr, &nchunks, &why)) {
    if (part == 0 && t == 0) raise_status(a.ws, why);
    return;
  }
  if (part == 0 && t == 0) *a.decsize = hdr;
  const uint16_t* const tab = (const uint16_t*)(a.in + 8);
  long long* const offs = (long long*)(a.ws + kWsFlags);
  const long long rounds = (nchunks + kScanRound - 1) / kScanRound;
  const long long per = (rounds + parts - 1) / parts;   // rounds per part
  const long long start = (long long)part * per * kScanRound;
  const long long end = min(nchunks, start + per * kScanRound);
  if (start >= nchunks && !(part == 0 && nchunks == 0)) return;   // uniform: no rounds here
  long long carry = start > 0 ? table_prefix(tab, start, wsum) : 0;
  // Each wave takes kScanBlock consecutive entries per round, as kScanRows
  // rows of 256: lane l holds entries 4l .. 4l+3 of every row (one 8-byte
  // word), so the table loads and the offset stores are contiguous across
  // the wave.  A wave scan per row gives each lane its entries' offsets.
  for (long long base = start; base < end; base += kScanRound) {
    const long long wb = base + (long long)wave * kScanBlock;
    unsigned long long w[kScanRows];
    uint32_t rs[kScanRows], incl[kScanRows];
#pragma unroll
    for (int r = 0; r < kScanRows; r++) {
      const long long i0 = wb + r * 256 + lane * 4;
      if (i0 + 4 <= nchunks) {
        w[r] = *(const unsigned long long*)(tab + i0);   // 8-byte aligned: i0 is a multiple of 4
      } else {
        w[r] = 0;
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (i0 + k < nchunks) w[r] |= (unsigned long long)tab[i0 + k] << (16 * k);
      }
      rs[r] = (uint32_t)(w[r] & 0xffff) + (uint32_t)((w[r] >> 16) & 0xffff) + (uint32_t)((w[r] >> 32) & 0xffff) +
              (uint32_t)(w[r] >> 48);
    }
    uint32_t wtot = 0;
#pragma unroll
    for (int r = 0; r < kScanRows; r++) {
      incl[r] = wave_incl_scan(rs[r]);
      wtot += (uint32_t)__builtin_amdgcn_readlane((int)incl[r], 63);
    }
    if (lane == 0) wsum[wave] = wtot;
    __syncthreads();
    long long before = 0, tile = 0;
#pragma unroll
    for (int v = 0; v < kScanThreads / 64; v++) {
      const long long x = wsum[v];
      if (v < wave) before += x;
      tile += x;
    }
    long long row0 = carry + before;   // offset of the wave's first entry this round
#pragma unroll
    for (int r = 0; r < kScanRows; r++) {
      const long long i0 = wb + r * 256 + lane * 4;
      long long o[4];
      o[0] = row0 + (long long)(incl[r] - rs[r]);
#pragma unroll
      for (int k = 1; k < 4; k++) o[k] = o[k - 1] + (long long)((w[r] >> (16 * (k - 1))) & 0xffff);
      if (i0 + 4 <= nchunks) {
        longlong2* const q = (longlong2*)(offs + i0);   // 16-byte aligned: i0 is a multiple of 4
        q[0] = make_longlong2(o[0], o[1]);
        q[1] = make_longlong2(o[2], o[3]);
      } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (i0 + k < nchunks) offs[i0 + k] = o[k];
      }
      row0 += (long long)(uint32_t)__builtin_amdgcn_readlane((int)incl[r], 63);
    }
    carry += tile;
    __syncthreads();
  }
  if (t == 0 && end == nchunks) offs[nchunks] = carry;   // the part holding the last entry (part 0 when there are none)
}

// One chunk of the decoder, by one workgroup of kThreads threads, once
// the offsets are known (no cross-workgroup dependency): aligned 16-byte
// loads of the payload's cover into LDS -> per-lane bit extraction ->
// inverse transform -> coalesced stores.  hdr / nchunks come from a header
// that passed header_ok.  Damage found is reported in the status word.
// csize_known: the chunk's size-table entry when the caller has it (from
// the offsets table), else -1 and it is read here.
template <typename W, bool LOSSY>
__device__ __forceinline__ void decode_chunk_at(const DecodeArgs& a, long long hdr, long long nchunks, long long c,
                                                long long excl, uint32_t* stage, int csize_known = -1)
{
  constexpr int PL = Word<W>::PerLane;
  constexpr int TB = Word<W>::TB;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  const long long data_base = 8 + 2 * nchunks;
  const uint16_t* const sizes = (const uint16_t*)(a.in + 8);
  const W offs = (W)a.offs;
  const long long base = c * kChunk;
  const int osize = (int)min((long long)kChunk, hdr - base);
  const int csize = csize_known >= 0 ? csize_known : uni((int)sizes[c]);
  const bool bad = csize > osize || (csize < osize && csize < kSub + 2) ||
                   (a.inlen >= 0 && data_base + excl + csize > a.inlen);   // payload beyond the stream

  if (!bad) {
    // aligned cover of the payload: payload byte i sits at stage byte mis + i
    const uint8_t* const src = a.in + data_base + excl;
    const int mis = (int)((uintptr_t)src & 15);
    const uint4* const cover = (const uint4*)(src - mis);
    const int nblk = (mis + csize + 15) >> 4;
    // streamed payload reads for lossless streams, streamed output stores for
    // lossy ones (each measured best for its mode; see ntload).  A cover of
    // more than one block per thread (every lossless one, up to 16 KiB) goes
    // straight to LDS, all of a lane's loads in flight together: block
    // j = t + k * kThreads, which wave wv's k-th load writes at its base
    // stage block k * kThreads + 64 wv plus its lane.  A smaller lossy cover
    // takes one trip of the plain loop.
    if (!LOSSY || nblk > kThreads) {
      constexpr int kCoverLoads = (kStageBytes / 16 + kThreads - 1) / kThreads;
      static_assert((15 + kChunk + 15) / 16 <= kStageBytes / 16, "the cover fits the stage");
#pragma unroll
      for (int k = 0; k < kCoverLoads; k++) {
        const int j = t + k * kThreads;
        if (j < nblk) load_to_lds16<!LOSSY>(cover + j, (uint4*)stage + (k * kThreads + 64 * wv));
      }
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
      for (int j = t; j < nblk; j += kThreads) ((uint4*)stage)[j] = cover[j];
    }
    __syncthreads();

    uint8_t* const dst = a.out + base;
    if (csize == osize) {
#pragma unroll
      for (int r = 0; r < kRounds; r++) {
        const int off = r * kThreads * 16 + t * 16;
        if (off < osize) store_slice<LOSSY>(dst, off, osize, lds_read16(stage, mis + off));
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
        int p0 = 8 * (mis + kSub) + sub_bit_offset<W>(L, wv * kWaveRounds);
        const int woff = wv * (kWaveRounds * kSubBytes) + lane * 8;
        const LossyF32Inv pinv = lossy_f32_inv_params(a.thr_e, (uint32_t)a.offs);   // float32 lossy only
#pragma unroll
        for (int r = 0; r < kWaveRounds; r++) {
          const int ln = sub_length(L, wv * kWaveRounds + r);
          const uint64_t f = read_field(stage, p0 + lane * PL * ln);
          p0 += ln * Word<W>::PerSub;
          W v[PL];
          if constexpr (PL == 2 && LOSSY) {
            const uint32_t m = ln >= 32 ? 0xffffffffu : ((1u << ln) - 1u);
            v[0] = lossy_value_f32((uint32_t)f & m, pinv);
            v[1] = lossy_value_f32((uint32_t)(f >> ln) & m, pinv);
          } else if constexpr (PL == 2) {
            const uint32_t m = ln >= 32 ? 0xffffffffu : ((1u << ln) - 1u);
            v[0] = from_code<W, LOSSY>((W)f & m, a.thr_e, offs);
            v[1] = from_code<W, LOSSY>((W)(f >> ln) & m, a.thr_e, offs);
          } else {
            const uint64_t m = ln >= 64 ? ~0ull : ((1ull << ln) - 1ull);
            v[0] = from_code<W, LOSSY>((W)(f & m), a.thr_e, offs);
          }
          store_lane<LOSSY>(dst, woff + r * kSubBytes, osize, values_lane(v));
        }
      }
    }
  } else if (t == 0) {
    raise_status(a.ws, kStCorrupt);
  }
  __syncthreads();
}

// The same, with the chunk's offset from the offsets table (offset scan).
template <typename W, bool LOSSY>
__device__ __forceinline__ void decode_chunk(const DecodeArgs& a, long long hdr, long long nchunks, long long c,
                                             uint32_t* stage)
{
  const long long* const offs_tab = (const long long*)(a.ws + kWsFlags);
  decode_chunk_at<W, LOSSY>(a, hdr, nchunks, c, uni(offs_tab[c]), stage);
}

// Sum of the size-table entries [from, to), to - from <= kWinMax, by one
// workgroup: table_load issues the loads (entry from + t + kThreads*i for
// thread t, coalesced), table_sum reduces them once they are needed, so the
// loads' latency can hide behind other work.
constexpr int kWinLoads = 8;
constexpr int kWinMax = kThreads * kWinLoads;

// Entries beyond the window are loaded from entry 0 (always inside the
// table) and dropped in table_sum, so the loads are unconditional.
__device__ __forceinline__ void table_load(const uint16_t* sizes, long long from, long long to, uint32_t v[kWinLoads])
{
  const int t = threadIdx.x;
  long long lim = to - from - t;   // window entries at or after this thread's first
  asm volatile("" : "+v"(lim));    // opaque: keeps the loads unconditional
#pragma unroll
  for (int i = 0; i < kWinLoads; i++) v[i] = sizes[lim > (long long)kThreads * i ? from + t + (long long)kThreads * i : 0];
}

__device__ __forceinline__ long long table_sum(const uint32_t v[kWinLoads], long long from, long long to, long long* red)
{
  const long long lim = to - from - (long long)threadIdx.x;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kWinLoads; i++) s += lim > (long long)kThreads * i ? v[i] : 0u;
  const uint64_t w = wave_sum((uint64_t)s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = (long long)w;
  __syncthreads();
  long long tot = 0;
#pragma unroll
  for (int i = 0; i < kThreads / 64; i++) tot += red[i];
  return uni(tot);
}

}  // namespace sleek
