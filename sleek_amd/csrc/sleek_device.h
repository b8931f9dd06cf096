// Device-side building blocks of the SLEEK codec for gfx950 (CDNA4, wave64).
//
// Format constants, the per-value integer transforms, wave primitives and the
// LDS <-> global byte movers shared by the plain and the batched kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sleek {

// Loads and stores through global-address-space pointers.  Pointers that
// come from kernel arguments are proven global by the compiler; those read
// from memory (a batch item's arrays) are not, and their accesses would
// become flat instructions, which also count as LDS operations: each LDS
// wait would then wait for the array loads and stores in flight as well.
// The access is made with a plain scalar or vector type of the same size:
// HIP's vector types are classes, whose copy goes through a generic `this`
// and would make the access flat again.
typedef unsigned int graw2 __attribute__((ext_vector_type(2)));
typedef unsigned int graw4 __attribute__((ext_vector_type(4)));
template <int N> struct GRaw;
template <> struct GRaw<1> { typedef uint8_t T; };
template <> struct GRaw<2> { typedef uint16_t T; };
template <> struct GRaw<4> { typedef uint32_t T; };
template <> struct GRaw<8> { typedef graw2 T; };
template <> struct GRaw<16> { typedef graw4 T; };

template <typename T>
__device__ __forceinline__ T gload(const T* p)
{
  typedef typename GRaw<sizeof(T)>::T R;
  const R r = *(const __attribute__((address_space(1))) R*)p;
  T t;
  __builtin_memcpy(&t, &r, sizeof(T));
  return t;
}
// The same with the non-temporal hint (the `nt` bit: stream through the
// caches instead of keeping the line).  The encoders read their input this
// way: read once, 2x the caches' size, it would otherwise evict the lines
// the look-back polls and the counters live on (measured 7-14 % faster
// encoders on 512^3 float32).  Lossy decoders store their output this way
// (about 12 % faster); the lossless decoder reads its payload this way.
template <typename T>
__device__ __forceinline__ T ntload(const T* p)
{
  typedef typename GRaw<sizeof(T)>::T R;
  const R r = __builtin_nontemporal_load((const __attribute__((address_space(1))) R*)p);
  T t;
  __builtin_memcpy(&t, &r, sizeof(T));
  return t;
}
template <typename T>
__device__ __forceinline__ void ntstore(T* p, T v)
{
  typedef typename GRaw<sizeof(T)>::T R;
  R r;
  __builtin_memcpy(&r, &v, sizeof(T));
  __builtin_nontemporal_store(r, (__attribute__((address_space(1))) R*)p);
}

template <typename T>
__device__ __forceinline__ void gstore(T* p, T v)
{
  typedef typename GRaw<sizeof(T)>::T R;
  R r;
  __builtin_memcpy(&r, &v, sizeof(T));
  *(__attribute__((address_space(1))) R*)p = r;
}

// ---- format constants (sleek_common_gpu.h:54, :317-318 of the reference) ----
constexpr int kChunk = 16384;            // bytes of input per chunk
constexpr int kSub = 32;                 // subchunks per chunk
constexpr int kSubBytes = kChunk / kSub; // 512 bytes per subchunk

// ---- launch geometry ----
// A workgroup of 4 waves codes one chunk; wave w owns subchunks 8w..8w+7,
// one per round, and each lane holds one 8-byte slice of the subchunk, so
// every per-subchunk quantity (bit length, bit offset) is wave-uniform.
constexpr int kThreads = 256;                       // 4 waves of 64
constexpr int kWaveRounds = kSub / (kThreads / 64); // subchunks per wave
static_assert(kWaveRounds % 4 == 0, "a wave's bit lengths fill whole dwords");
constexpr int kRounds = kChunk / (kThreads * 16);   // 16-byte slices per thread and chunk
constexpr int kStageBytes = kChunk + 64;            // LDS staging buffer (payload + realignment slack)

// ---- scan state words ----
// Workspace layout (uint64 words): [1] status bits, then (from
// word 8) one look-back word per chunk, one per group of 2^gshift chunks
// (8..64, a launch parameter), and one completion counter per group.  A look-back word holds a byte count in
// bits 61..0 and its kind in bits 63..62: 0 = not published, 1 = aggregate
// (the chunk's or group's own compressed size), 2 = inclusive (stream offset
// just past it).  Kinds only ever increase, so inclusive words are written
// with an atomic max and are never downgraded.
constexpr int kWsStatus = 1;
constexpr int kWsFlags = 8;
constexpr uint64_t kAgg = 1ull << 62;
constexpr uint64_t kInc = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

// status bits reported by the kernels
constexpr uint64_t kStCorrupt = 1;   // malformed stream
constexpr uint64_t kStCapacity = 2;  // decoded size exceeds the caller's buffer

// ---- element traits ----
template <typename W> struct Word;
template <> struct Word<uint32_t> {
  static constexpr int TB = 32, EXPB = 8, MANB = 23;
  static constexpr int PerSlice = 4;    // values per 16-byte slice
  static constexpr int PerLane = 2;     // values per 8-byte slice (one lane's share of a subchunk)
  static constexpr int PerSub = 128;    // values per subchunk
};
template <> struct Word<uint64_t> {
  static constexpr int TB = 64, EXPB = 11, MANB = 52;
  static constexpr int PerSlice = 2;
  static constexpr int PerLane = 1;
  static constexpr int PerSub = 64;
};

__host__ __device__ __forceinline__ int clz_w(uint32_t x) { return x ? __builtin_clz(x) : 32; }
__host__ __device__ __forceinline__ int clz_w(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// ---- value transforms.  The steps are fixed by the stream format ----

// Lossless code (reference sleek_common_cpu.h:152-169): rotate the sign into
// bit 0, re-centre a non-zero exponent around 1.0, zigzag-fold.
//
// Closed form used here: with v = rotl1(x) and u = v - kOne (kOne = the
// lowest exponent bit after the rotation), the reference's case analysis
// on the exponent e (0 / 1..2^(E-1) / above) reduces to
//   code = (u << 1) + (t << (M + 2))  XOR  (t - 1),   t = top bit of u,
// since 1 <= e <= 2^(E-1) exactly when u's top bit is clear.  It equals
// the reference transform on all 2^32 float32 and all float64 exponent
// classes (checked against the oracle).
template <typename W>
__host__ __device__ __forceinline__ W lossless_code(W x)
{
  constexpr int TB = Word<W>::TB;
  constexpr W kOne = (W)1 << (Word<W>::MANB + 1);
  const W u = ((x << 1) | (x >> (TB - 1))) - kOne;
  // ts = -t (0 or all ones): t << (M + 2) = ts & 2^(M+2), t - 1 = ~ts
  typedef typename std::conditional<sizeof(W) == 4, int32_t, int64_t>::type S;
  const W ts = (W)((S)u >> (TB - 1));
  return ((u << 1) + (ts & ((W)1 << (Word<W>::MANB + 2)))) ^ ~ts;
}

// Inverse of lossless_code (reference sleek_common_cpu.h:317-335): bit 0
// of the code is the fold; an odd code undoes (u << 1) ^ ~0, an even one
// (u << 1) + 2 kOne, whose rotated value v gets its top bit back exactly
// when its exponent field is non-zero (c1 >= kOne).
template <typename W>
__host__ __device__ __forceinline__ W lossless_value(W c)
{
  constexpr int TB = Word<W>::TB;
  constexpr W kHalf = (W)1 << (TB - 1);
  constexpr W kOne = (W)1 << (Word<W>::MANB + 1);
  const W m = (W)0 - (c & 1);
  const W c1 = c >> 1;
  const W sel = (c1 + (kHalf - kOne)) & kHalf;
  const W v = (c1 ^ m) + ((m & (kHalf + kOne)) | (~m & sel));
  return (v >> 1) | (v << (TB - 1));
}

// Error-bounded code (reference sleek_common_cpu.h:393-419): magnitudes at or
// above 2^thr keep their bits (minus an offset), magnitudes in [2^eb, 2^thr)
// are rounded to bins of width 2^(eb+1), smaller ones become 0; the sign is
// folded into bit 0 with both zeros mapping to 0.  Unsigned arithmetic equals
// the reference's wrapping signed arithmetic modulo 2^TB.  Branch-free:
//  * below 2^eb the bin formula itself gives 0 (the shift, clamped to TB-1,
//    exceeds the mantissa's width), so no third case is selected;
//  * (q << 1 | !sign) - (q != 0) equals 2q - min(q, sign).
// Equal to the reference on all 2^32 float32 patterns for every class of
// bound (checked against the oracle).
template <typename W>
__host__ __device__ __forceinline__ W lossy_code(W val, int eb_e, int thr_e, W offs)
{
  (void)eb_e;   // implied by thr_e
  constexpr int TB = Word<W>::TB;
  constexpr int EM = Word<W>::EXPB + Word<W>::MANB;
  constexpr int M = Word<W>::MANB;
  const W mag = val & (((W)1 << EM) - 1);
  const int e = (int)(mag >> M);
  const int d = thr_e - 1 - e;                       // shift - 1
  const int sh1 = d < 0 ? 0 : (d > TB - 2 ? TB - 2 : d);
  const W mant = (val & (((W)1 << M) - 1)) | ((W)1 << M);
  const W binned = ((mant >> sh1) + 1) >> 1;         // (mant + 2^(sh-1)) >> sh
  const W q = (e >= thr_e) ? (W)(mag - offs) : binned;
  const W sign = val >> (TB - 1);
  return (q << 1) - (q < sign ? q : sign);
}

// Float32 error-bounded code: the same function as lossy_code<uint32_t> in
// 9 instructions instead of 16, for every bound (thr_e in [24, 254]).
//  * Bin: for e < thr_e, (mant >> sh1) = floor(y) with y = |x| * 2^(151 - thr_e)
//    for normal x, the product being exact (a power-of-two scaling of a value
//    that lands below 2^24; one that underflows is below 1 either way,
//    flushed or not).  The bin is ((mant >> sh1) + 1) >> 1 = floor((y + 1) / 2)
//    = floor(z + 1/2) with z = y / 2 = |x| * 2^(150 - thr_e), again exact:
//    that is one v_cvt_rpi_i32_f32 (floor(z + 0.5), taken exactly).  Zero and
//    denormal inputs (e = 0): the reference's bin is 1 when thr_e = 24 (a
//    denormal bound) and 0 above, while floor(z + 1/2) is 0 or 1; a floor
//    rmin (1 for thr_e = 24, else 0) makes them equal, and no normal input is
//    affected (its bin is >= 1 when thr_e = 24).
//  * Exact branch (|x| >= 2^(thr_e - 127), Inf and NaN: !(|x| < thr)):
//    q = mag - offs = val - offs modulo 2^31; the doubling below shifts the
//    sign bit (and the borrow into it) out.
//  * Sign fold: 2q - min(q, sign) = 2q -sat sign (2q is 0 exactly when q is).
struct LossyF32 {
  float hscale;       // 2^(150 - thr_e)
  float thr;          // 2^(thr_e - 127): the exact branch's threshold
  uint32_t negoffs;   // -offs mod 2^32
  int rmin;           // 1 for a denormal bound (thr_e = 24), else 0
};

__host__ __device__ __forceinline__ LossyF32 lossy_f32_params(int thr_e, uint32_t offs)
{
  LossyF32 p;
  const uint32_t sb = (uint32_t)(277 - thr_e) << 23, tb = (uint32_t)thr_e << 23;
  __builtin_memcpy(&p.hscale, &sb, 4);
  __builtin_memcpy(&p.thr, &tb, 4);
  p.negoffs = 0u - offs;
  p.rmin = thr_e == 24 ? 1 : 0;
  return p;
}

__host__ __device__ __forceinline__ uint32_t lossy_code_f32(uint32_t val, const LossyF32& p)
{
  float x;
  __builtin_memcpy(&x, &val, 4);
  const float ax = __builtin_fabsf(x);
  const float z = ax * p.hscale;
  int r;
#if defined(__HIP_DEVICE_COMPILE__)
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(z));   // floor(z + 0.5), saturating, NaN to 0
#else
  const double zr = __builtin_floor((double)z + 0.5);   // exact where it is used (z < 2^23)
  r = (z != z) ? 0 : (zr >= 2147483647.0 ? 2147483647 : (int)zr);
#endif
  r = r > p.rmin ? r : p.rmin;
  const uint32_t q = !(ax < p.thr) ? val + p.negoffs : (uint32_t)r;
  const uint32_t q2 = q << 1, s = val >> 31;
  return q2 > s ? q2 - s : 0u;   // saturating subtract
}

// Inverse of lossy_code (reference sleek_common_cpu.h:559-591), branch-free.
template <typename W>
__host__ __device__ __forceinline__ W lossy_value(W enc, int thr_e, W offs)
{
  constexpr int TB = Word<W>::TB;
  constexpr int M = Word<W>::MANB;
  constexpr int EM = Word<W>::EXPB + Word<W>::MANB;
  const W mag = (W)(enc + 1) >> 1;
  const int sh = min(max(clz_w(mag) - (TB - 1 - M), 0), TB - 1);   // in [1, M] where it is used
  const W binned = ((mag << sh) & (((W)1 << M) - 1)) | ((W)(thr_e - sh) << M);
  const W dec = (mag >= ((W)1 << M)) ? (W)(mag + offs) : binned;
  if constexpr (sizeof(W) == 4) {
    // mag = 0 for the codes 0 and 2^32 - 1 (enc + 1 wraps, as in the
    // reference's word type): no magnitude, only the sign bit
    return (mag == 0) ? (W)(enc << EM) : (W)(dec | (enc << EM));
  } else {
    // Deviation: the code 2^64 - 1, which no encoder writes (only a damaged
    // stream holds it), decodes to ((thr_e - 53) << 52) | 2^63 here and to
    // the sign bit alone in the reference.  The float64 decoders built from
    // the select above gave nondeterministic wrong values on the GPU for
    // valid streams (cause not found), so they keep this form.
    return (enc == 0) ? (W)0 : (W)(dec | (enc << EM));
  }
}

// Float32 inverse of lossy_code: the same function as lossy_value<uint32_t>
// on all 2^32 codes, in 8 instructions instead of 16.
//  * Binned values (mag in [1, 2^23)): the reference normalises mag into a
//    float with exponent field thr_e - sh, sh = 23 - floor(log2 mag).  That
//    float is mag * 2^(thr_e - 150), a normal number for every thr_e in
//    [24, 254] (its exponent lies in [thr_e - 150, thr_e - 128]): one
//    conversion and one multiply, both exact, give its bits.
//  * mag = 0 (the codes 0 and 2^32 - 1, enc + 1 wrapping as in the
//    reference): the product is +0, so only the sign bit remains.
struct LossyF32Inv {
  float scale;     // 2^(thr_e - 150)
  uint32_t offs;   // the exact branch's offset
};

__host__ __device__ __forceinline__ LossyF32Inv lossy_f32_inv_params(int thr_e, uint32_t offs)
{
  LossyF32Inv p;
  const uint32_t sb = (uint32_t)(thr_e - 23) << 23;
  __builtin_memcpy(&p.scale, &sb, 4);
  p.offs = offs;
  return p;
}

__host__ __device__ __forceinline__ uint32_t lossy_value_f32(uint32_t enc, const LossyF32Inv& p)
{
  const uint32_t mag = (enc + 1u) >> 1;
  const float fb = (float)mag * p.scale;   // exact where it is used (mag < 2^23)
  uint32_t b;
  __builtin_memcpy(&b, &fb, 4);
  const uint32_t dec = mag >= (1u << 23) ? mag + p.offs : b;
  return dec | (enc << 31);
}

template <typename W, bool LOSSY>
__device__ __forceinline__ W to_code(W v, int eb_e, int thr_e, W offs)
{
  if constexpr (LOSSY) return lossy_code<W>(v, eb_e, thr_e, offs);
  else return lossless_code<W>(v);
}

template <typename W, bool LOSSY>
__device__ __forceinline__ W from_code(W c, int thr_e, W offs)
{
  if constexpr (LOSSY) return lossy_value<W>(c, thr_e, offs);
  else return lossless_value<W>(c);
}

// ---- 16-byte slice <-> values ----
__device__ __forceinline__ void slice_values(const uint4& q, uint32_t v[4]) { v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w; }
__device__ __forceinline__ void slice_values(const uint4& q, uint64_t v[2])
{
  v[0] = (uint64_t)q.x | ((uint64_t)q.y << 32);
  v[1] = (uint64_t)q.z | ((uint64_t)q.w << 32);
}
__device__ __forceinline__ uint4 values_slice(const uint32_t v[4]) { return make_uint4(v[0], v[1], v[2], v[3]); }
__device__ __forceinline__ uint4 values_slice(const uint64_t v[2])
{
  return make_uint4((uint32_t)v[0], (uint32_t)(v[0] >> 32), (uint32_t)v[1], (uint32_t)(v[1] >> 32));
}

__device__ __forceinline__ void lane_values(const uint2& q, uint32_t v[2]) { v[0] = q.x; v[1] = q.y; }
__device__ __forceinline__ void lane_values(const uint2& q, uint64_t v[1]) { v[0] = (uint64_t)q.x | ((uint64_t)q.y << 32); }
__device__ __forceinline__ uint2 values_lane(const uint32_t v[2]) { return make_uint2(v[0], v[1]); }
__device__ __forceinline__ uint2 values_lane(const uint64_t v[1]) { return make_uint2((uint32_t)v[0], (uint32_t)(v[0] >> 32)); }

// A lane's codes as one bit field: code k at bit k*ln (ln <= word width,
// so two float32 codes fit in 64 bits).
__device__ __forceinline__ uint64_t lane_field(const uint32_t v[2], int ln) { return (uint64_t)v[0] | ((uint64_t)v[1] << ln); }
__device__ __forceinline__ uint64_t lane_field(const uint64_t v[1], int) { return v[0]; }

// OR over the 64 lanes of a wave, returned as a wave-uniform value: DPP
// swaps inside rows of 16 lanes, then the four row results are combined.
__device__ __forceinline__ uint32_t wave_or(uint32_t v)
{
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);  // row_mirror
  return (uint32_t)(__builtin_amdgcn_readlane((int)v, 0) | __builtin_amdgcn_readlane((int)v, 16) |
                    __builtin_amdgcn_readlane((int)v, 32) | __builtin_amdgcn_readlane((int)v, 48));
}

// Bit length of the largest code over a wave (the bit length of the OR of
// the codes), wave-uniform.
__device__ __forceinline__ int wave_bit_length(uint32_t orv) { const uint32_t o = wave_or(orv); return o ? 32 - __clz(o) : 0; }
// 64-bit codes: each lane's own bit length (no branch), then the wave's max.
__device__ __forceinline__ int wave_bit_length(uint64_t orv)
{
  const uint32_t hi = (uint32_t)(orv >> 32), lo = (uint32_t)orv;
  uint32_t b = hi ? 64u - (uint32_t)__clz(hi) : (lo ? 32u - (uint32_t)__clz(lo) : 0u);
  b = max(b, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)b, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  b = max(b, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)b, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  b = max(b, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)b, 0x141, 0xF, 0xF, false));  // row_half_mirror
  b = max(b, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)b, 0x140, 0xF, 0xF, false));  // row_mirror
  return (int)max(max((uint32_t)__builtin_amdgcn_readlane((int)b, 0), (uint32_t)__builtin_amdgcn_readlane((int)b, 16)),
                  max((uint32_t)__builtin_amdgcn_readlane((int)b, 32), (uint32_t)__builtin_amdgcn_readlane((int)b, 48)));
}

// ---- bit lengths of all subchunks at once ----
// The per-lane values of N subchunks (32 for a chunk, 16 for a half chunk):
// lane l's OR of its codes of subchunk s, or (MAX) their bit length.  They
// are combined over the wave by a transposing butterfly: in each exchange a
// lane keeps half of its values and sends the other half to a partner lane
// that keeps the opposite half, so a lane's values halve while the lanes
// they cover double (row_mirror, row_half_mirror and two quad_perms inside
// rows of 16, then an xor-16 swizzle, then an xor-32 permute).  N - 1
// exchanges cover all N subchunks; reducing each subchunk on its own costs
// 4 DPP steps and 4 lane reads per subchunk.  Afterwards subchunk s's value
// is in lane l for s = (l3 l2 l1 l0 l4) in binary when N = 32 (l_k: bit k
// of l) and s = (l3 l2 l1 l0) when N = 16.
template <bool MAX>
__device__ __forceinline__ uint32_t bfly_op(uint32_t a, uint32_t b)
{
  return MAX ? max(a, b) : (a | b);
}

template <int H, int BIT, int CTRL, bool MAX, int N>
__device__ __forceinline__ void bfly_step(uint32_t (&v)[N], int lane)
{
  const bool hi = (lane >> BIT) & 1;
#pragma unroll
  for (int i = 0; i < H; i++) {
    const uint32_t keep = hi ? v[i + H] : v[i];
    const uint32_t send = hi ? v[i] : v[i + H];
    v[i] = bfly_op<MAX>(keep, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, CTRL, 0xF, 0xF, false));
  }
}

__device__ __forceinline__ uint32_t swz_xor16(uint32_t v) { return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F); }

// The N subchunks' bit lengths, byte-packed little-endian into N/4 uniform
// dwords (L[i] byte j: subchunk 4i + j).  v is consumed.
template <int N, bool MAX>
__device__ __forceinline__ void wave_lengths(uint32_t (&v)[N], uint32_t* L)
{
  static_assert(N == 32 || N == 16, "a chunk or half a chunk");
  const int lane = threadIdx.x & 63;
  if constexpr (N == 32) {
    bfly_step<16, 3, 0x140, MAX>(v, lane);   // row_mirror
    bfly_step<8, 2, 0x141, MAX>(v, lane);    // row_half_mirror
    bfly_step<4, 1, 0x4E, MAX>(v, lane);     // quad_perm [2,3,0,1]
    bfly_step<2, 0, 0xB1, MAX>(v, lane);     // quad_perm [1,0,3,2]
    const bool hi = (lane >> 4) & 1;
    const uint32_t keep = hi ? v[1] : v[0];
    const uint32_t send = hi ? v[0] : v[1];
    v[0] = bfly_op<MAX>(keep, swz_xor16(send));
  } else {
    bfly_step<8, 3, 0x140, MAX>(v, lane);
    bfly_step<4, 2, 0x141, MAX>(v, lane);
    bfly_step<2, 1, 0x4E, MAX>(v, lane);
    bfly_step<1, 0, 0xB1, MAX>(v, lane);
    v[0] = bfly_op<MAX>(v[0], swz_xor16(v[0]));
  }
  const uint32_t r = bfly_op<MAX>(v[0], (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, (int)v[0]));
  const uint32_t bl = MAX ? r : 32u - (uint32_t)__clz(r);   // __clz(0) = 32
  if constexpr (N == 32) {
    // lanes l, l^1, l^16, l^17 hold bytes 0..3 of one dword: byte j = (l0 l4)
    uint32_t t = bl << (8 * (((lane & 1) << 1) | ((lane >> 4) & 1)));
    t |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0xB1, 0xF, 0xF, false);
    t |= swz_xor16(t);
#pragma unroll
    for (int i = 0; i < 8; i++)
      L[i] = (uint32_t)__builtin_amdgcn_readlane((int)t, (((i >> 2) & 1) << 3) | (((i >> 1) & 1) << 2) | ((i & 1) << 1));
  } else {
    // lanes l .. l^3 hold bytes 0..3 of one dword: byte j = (l1 l0)
    uint32_t t = bl << (8 * (lane & 3));
    t |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0xB1, 0xF, 0xF, false);
    t |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, 0x4E, 0xF, 0xF, false);
#pragma unroll
    for (int i = 0; i < 4; i++) L[i] = (uint32_t)__builtin_amdgcn_readlane((int)t, ((i >> 1) << 3) | ((i & 1) << 2));
  }
}

// A lane's contribution to its subchunk's bit length: the OR of its codes
// (32-bit words), or the bit length of that OR (64-bit words, combined by MAX).
__device__ __forceinline__ uint32_t lane_length_word(uint32_t orv) { return orv; }
__device__ __forceinline__ uint32_t lane_length_word(uint64_t orv)
{
  const uint32_t hi = (uint32_t)(orv >> 32), lo = (uint32_t)orv;
  return hi ? 64u - (uint32_t)__clz(hi) : 32u - (uint32_t)__clz(lo);
}

__device__ __forceinline__ int bit_length(uint32_t x) { return 32 - __clz(x); }
__device__ __forceinline__ int bit_length(uint64_t x) { return 64 - __clzll(x); }

// max over the 32 lanes of a half-wave (ds_swizzle bit mode, xor 1..16)
__device__ __forceinline__ int half_wave_max(int v)
{
  v = max(v, __builtin_amdgcn_ds_swizzle(v, 0x041F));
  v = max(v, __builtin_amdgcn_ds_swizzle(v, 0x081F));
  v = max(v, __builtin_amdgcn_ds_swizzle(v, 0x101F));
  v = max(v, __builtin_amdgcn_ds_swizzle(v, 0x201F));
  v = max(v, __builtin_amdgcn_ds_swizzle(v, 0x401F));
  return v;
}

// sum over the 64 lanes of a wave
__device__ __forceinline__ uint64_t wave_sum(uint64_t v)
{
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Broadcast lane 0's value: makes block-uniform values (read from LDS or
// global memory by every lane) scalar, so branches on them are scalar
// branches and loops containing barriers keep a single, uniform latch.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ long long uni(long long v)
{
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// Make the compiler treat q as redefined here (an empty asm with register
// operands), so values derived from q before this point are recomputed
// rather than kept live in registers.
__device__ __forceinline__ void reissue(uint4& q)
{
  asm volatile("" : "+v"(q.x), "+v"(q.y), "+v"(q.z), "+v"(q.w));
}
__device__ __forceinline__ void reissue(uint32_t& x) { asm volatile("" : "+v"(x)); }
// v, recomputed where it is used: values derived from it there are not
// hoisted out of the enclosing loop (and kept live, or spilled) by the compiler.
__device__ __forceinline__ int opaque(int v)
{
  asm volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ void reissue_pair(uint2& q) { asm volatile("" : "+v"(q.x), "+v"(q.y)); }
__device__ __forceinline__ void reissue(uint64_t& x)
{
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  asm volatile("" : "+v"(lo), "+v"(hi));
  x = (uint64_t)lo | ((uint64_t)hi << 32);
}

// Any byte of x above n (n <= 127), without carries between bytes.
__device__ __forceinline__ bool any_byte_above(uint32_t x, uint32_t n)
{
  return (((x + 0x01010101u * (127u - n)) | x) & 0x80808080u) != 0u;
}

// Inclusive sum over the wave's 64 lanes: Hillis-Steele within each row of
// 16 lanes (DPP row shifts by 1, 2, 4, 8; lanes without a source add 0),
// then row 0's total into row 1 and row 2's into row 3 (row_bcast:15), then
// the first two rows' total into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);   // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);   // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);   // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);   // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);   // row_bcast:15, rows 1 and 3
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);   // row_bcast:31, rows 2 and 3
  return x;
}

// Sum of the four bytes of x.
__device__ __forceinline__ uint32_t byte_sum(uint32_t x) { return __builtin_amdgcn_sad_u8(x, 0u, 0u); }

// Bit offset (relative to the first packed bit) of subchunk s, given the 32
// per-subchunk bit lengths as 8 little-endian byte-packed dwords.
template <typename W>
__device__ __forceinline__ int sub_bit_offset(const uint32_t lg[8], int s)
{
  uint32_t acc = 0;
  const int full = s >> 2;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    if (i < full) acc += byte_sum(lg[i]);
  }
  const int part = s & 3;
  if (part) {
    const uint32_t w = lg[full] & ((1u << (8 * part)) - 1u);
    acc += byte_sum(w);
  }
  return (int)acc * Word<W>::PerSub;
}

__device__ __forceinline__ int sub_length(const uint32_t lg[8], int s) { return (int)((lg[s >> 2] >> (8 * (s & 3))) & 0xffu); }

// OR a bit field of up to 64 bits into a zeroed LDS bit buffer at bit p
// (the field spans at most three dwords).
__device__ __forceinline__ void or_field(uint32_t* st, int p, uint64_t f)
{
  const int d = p >> 5;
  const uint32_t sh = (uint32_t)p & 31u;
  const uint64_t lo = f << sh;
  const uint32_t w2 = (uint32_t)(((f >> 32) << sh) >> 32);
  atomicOr(&st[d], (uint32_t)lo);
  atomicOr(&st[d + 1], (uint32_t)(lo >> 32));
  if (w2) atomicOr(&st[d + 2], w2);
}

// OR a lane's codes of one subchunk into a zeroed LDS bit buffer whose
// fields start at bit p0 (a multiple of 32).  ln (1..TB) is wave-uniform, so
// the case is a scalar branch: the lane's field (PL codes of ln bits, PL * ln
// bits) goes to bit p0 + PL * ln * lane.
//  * float32, ln <= 16: the 2 ln-bit field fits a dword; when 2 ln divides
//    32 no field straddles a dword boundary (one OR), else two ORs.
//  * otherwise: a field of up to 64 bits over up to three dwords (or_field).
__device__ __forceinline__ void or_codes(uint32_t* st, int p0, int lane, int ln, const uint32_t v[2])
{
  const uint32_t x = (uint32_t)lane * (uint32_t)(2 * ln);
  if (ln <= 16) {
    const uint32_t sh = x & 31u;
    uint32_t* const d = st + (p0 >> 5) + (x >> 5);
    const uint32_t f = v[0] | (v[1] << ln);
    if ((ln & (ln - 1)) == 0) {
      atomicOr(d, f << sh);
    } else {
      const uint64_t g = (uint64_t)f << sh;
      atomicOr(d, (uint32_t)g);
      atomicOr(d + 1, (uint32_t)(g >> 32));
    }
  } else {
    or_field(st, p0 + (int)x, (uint64_t)v[0] | ((uint64_t)v[1] << ln));
  }
}
__device__ __forceinline__ void or_codes(uint32_t* st, int p0, int lane, int ln, const uint64_t v[1])
{
  or_field(st, p0 + lane * ln, v[0]);
}

// A lane's codes of one subchunk into the LDS bit buffer, with no clearing
// beforehand.  The subchunk's fields start at bit p0 (a multiple of 32) and
// cover its dwords exactly: lane l's field of w = PL * ln bits sits at bit
// p0 + x, x = l * w.  Every one of those dwords has its bit 0 inside exactly
// one lane's field; that lane writes the dword whole (its field's bits from
// there up, zeros above them).  Then each lane whose field starts inside a
// dword ORs its first bits into it.  A wave's LDS operations execute in
// order, and the compiler barrier keeps the writes ahead of the ORs, so
// every OR lands on its dword's write: one conflict-free write per dword and
// at most one OR per lane, where OR-ing into a cleared buffer took a clearing
// pass and two or three ORs per lane.
__device__ __forceinline__ void put_field(uint32_t* st, int p0, uint32_t x, uint32_t w, uint64_t f)
{
  const uint32_t off = x & 31u, s1 = (0u - x) & 31u;   // s1: bits from the field to the next dword start
  uint32_t* const dw = st + (p0 >> 5) + ((x + 31u) >> 5);
  if (s1 < w) dw[0] = (uint32_t)(f >> s1);
  if (s1 + 32u < w) dw[1] = (uint32_t)(f >> (s1 + 32u));
  asm volatile("" ::: "memory");
  if (off) atomicOr(st + (p0 >> 5) + (x >> 5), (uint32_t)(f << off));
}
__device__ __forceinline__ void put_codes(uint32_t* st, int p0, int lane, int ln, const uint32_t v[2])
{
  const uint32_t w = (uint32_t)(2 * ln), x = (uint32_t)lane * w;
  if (ln <= 16) {   // uniform: a field of at most 32 bits
    const uint32_t off = x & 31u, s1 = (0u - x) & 31u;
    const uint32_t f = v[0] | (v[1] << ln);
    if (s1 < w) st[(p0 >> 5) + ((x + 31u) >> 5)] = f >> s1;
    asm volatile("" ::: "memory");
    if (off) atomicOr(st + (p0 >> 5) + (x >> 5), f << off);
  } else {
    put_field(st, p0, x, w, (uint64_t)v[0] | ((uint64_t)v[1] << ln));
  }
}
__device__ __forceinline__ void put_codes(uint32_t* st, int p0, int lane, int ln, const uint64_t v[1])
{
  put_field(st, p0, (uint32_t)(lane * ln), (uint32_t)ln, v[0]);
}

// The 64 bits of an LDS bit buffer starting at bit p.
__device__ __forceinline__ uint64_t read_field(const uint32_t* st, int p)
{
  const int d = p >> 5;
  const uint32_t sh = (uint32_t)p & 31u;
  const uint32_t w0 = st[d], w1 = st[d + 1], w2 = st[d + 2];
  return (uint64_t)__builtin_amdgcn_alignbit(w1, w0, sh) | ((uint64_t)__builtin_amdgcn_alignbit(w2, w1, sh) << 32);
}

// s_waitcnt vmcnt(0) (gfx9 encoding: expcnt and lgkmcnt left at their
// maxima).  Ends a rarely taken branch that loads into the same registers as
// the common path it rejoins (a partial last chunk): without it the compiler
// cannot tell at the join that those loads are done, and the common path's
// next loads into those registers would wait for every memory operation in
// flight, stores of the previous chunk included.
__device__ __forceinline__ void wait_vm_drained() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ---- scan state ----
__device__ __forceinline__ uint64_t flag_load(const uint64_t* p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flag_store(uint64_t* p, uint64_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void raise_status(uint64_t* ws, uint64_t bits)
{
  __hip_atomic_fetch_or((unsigned long long*)&ws[kWsStatus], (unsigned long long)bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void flag_max(uint64_t* p, uint64_t v)
{
  __hip_atomic_fetch_max((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Scan state of one encode call inside its workspace.
struct ScanState {
  uint64_t* chunk;               // one word per chunk
  uint64_t* group;               // one word per group of 2^gshift chunks
  unsigned long long* count;     // per group: (chunks reported << 32) | bytes
  int gshift;
};

__device__ __forceinline__ ScanState scan_state(uint64_t* ws, long long nchunks, int gshift)
{
  const long long ngroups = (nchunks + (1ll << gshift) - 1) >> gshift;
  ScanState s;
  s.chunk = ws + kWsFlags;
  s.group = s.chunk + nchunks;
  s.count = (unsigned long long*)(s.group + ngroups);
  s.gshift = gshift;
  return s;
}

// chunks of group g (the last group may be short)
__device__ __forceinline__ long long group_members(const ScanState& s, long long g, long long nchunks)
{
  return min(1ll << s.gshift, nchunks - (g << s.gshift));
}

// Publish chunk c's inclusive offset (and its group's, for the group's last chunk).
__device__ __forceinline__ void publish_inclusive(const ScanState& s, long long c, long long nchunks, uint64_t incl)
{
  flag_max(s.chunk + c, kInc | incl);
  const long long g = c >> s.gshift;
  if (((c + 1) & ((1ll << s.gshift) - 1)) == 0 || c == nchunks - 1) flag_max(s.group + g, kInc | incl);
}

// 16 bytes starting at an arbitrary byte offset of a 16-byte-aligned LDS
// buffer.  byte_off must be congruent modulo 16 across the wave's active
// lanes (callers step it by 16 per lane), so the two aligned 16-byte lines
// read per lane are consecutive across lanes (no bank conflicts; five
// dword reads per lane 16 bytes apart would conflict 8-way) and the word
// selection and bit shift are wave-uniform.  Reads up to 32 bytes from the
// line holding byte_off.
__device__ __forceinline__ uint4 lds_read16(const uint32_t* lds, int byte_off)
{
  const int sub = uni(byte_off & 15);
  const uint32_t sh = (uint32_t)(sub & 3) * 8u;
  const int line = byte_off >> 4;
  const uint4 a = ((const uint4*)lds)[line];
  const uint4 b = ((const uint4*)lds)[line + 1];
  uint32_t x0, x1, x2, x3, x4;
  if (sub < 4) {
    x0 = a.x; x1 = a.y; x2 = a.z; x3 = a.w; x4 = b.x;
  } else if (sub < 8) {
    x0 = a.y; x1 = a.z; x2 = a.w; x3 = b.x; x4 = b.y;
  } else if (sub < 12) {
    x0 = a.z; x1 = a.w; x2 = b.x; x3 = b.y; x4 = b.z;
  } else {
    x0 = a.w; x1 = b.x; x2 = b.y; x3 = b.z; x4 = b.w;
  }
  return make_uint4(__builtin_amdgcn_alignbit(x1, x0, sh), __builtin_amdgcn_alignbit(x2, x1, sh),
                    __builtin_amdgcn_alignbit(x3, x2, sh), __builtin_amdgcn_alignbit(x4, x3, sh));
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// The LDS byte address of a __shared__ buffer passed as a generic pointer.
__device__ __forceinline__ uint32_t lds_addr(const void* p)
{
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// Direct global -> LDS load of 16 bytes per lane (`global_load_lds_dwordx4`,
// gfx950): lane i's 16 bytes at src land at LDS byte wave_base + 16 i, with
// wave_base (the LDS pointer, taken through M0) the same for the whole wave;
// lanes switched off write nothing.  No registers hold the data, so a lane
// can have all its loads in flight at once; the loads count in vmcnt, and
// the LDS holds their bytes once a wave has waited for vmcnt(0).  NT sets
// the non-temporal hint (see ntload).
template <bool NT>
__device__ __forceinline__ void load_to_lds16(const void* src, void* wave_base)
{
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                   (__attribute__((address_space(3))) void*)wave_base, 16, 0, NT ? 2 : 0);
}

// Whole 16-byte LDS lines at byte addresses a, a + 16 (and c, c + 16):
// 128-bit reads, which the compiler would otherwise narrow to the dwords
// used (bank conflicts across the wave); the wait for them is in the same
// asm block, so the outputs are complete when it ends.
__device__ __forceinline__ void lds_read_lines2(uint32_t a, u32x4& x, u32x4& y)
{
  asm volatile("ds_read_b128 %0, %2\n\tds_read_b128 %1, %2 offset:16\n\ts_waitcnt lgkmcnt(0)"
               : "=&v"(x), "=&v"(y)
               : "v"(a));
}
__device__ __forceinline__ void lds_read_lines4(uint32_t a, uint32_t c, u32x4& x, u32x4& y, u32x4& z, u32x4& w)
{
  asm volatile("ds_read_b128 %0, %4\n\tds_read_b128 %1, %4 offset:16\n\t"
               "ds_read_b128 %2, %5\n\tds_read_b128 %3, %5 offset:16\n\ts_waitcnt lgkmcnt(0)"
               : "=&v"(x), "=&v"(y), "=&v"(z), "=&v"(w)
               : "v"(a), "v"(c));
}

// 16 bytes from a window of two consecutive lines (a, b): dwords Q .. Q + 4
// shifted right by sh bits (sh = 8 * (byte offset mod 4)).
template <int Q>
__device__ __forceinline__ uint4 line_window(const u32x4& a, const u32x4& b, uint32_t sh)
{
  const uint32_t d[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  return make_uint4(__builtin_amdgcn_alignbit(d[Q + 1], d[Q], sh), __builtin_amdgcn_alignbit(d[Q + 2], d[Q + 1], sh),
                    __builtin_amdgcn_alignbit(d[Q + 3], d[Q + 2], sh), __builtin_amdgcn_alignbit(d[Q + 4], d[Q + 3], sh));
}

// Copy len bytes from LDS to global memory at an arbitrary byte address, by
// one wave (the single-wave variant of store_unaligned below).
__device__ __forceinline__ void wave_store_unaligned(uint8_t* dst, const uint32_t* lds, int len)
{
  const int l = threadIdx.x & 63;
  const int mis = (int)((uintptr_t)dst & 15);
  const int head = min(len, (16 - mis) & 15);
  const int nfull = (len - head) >> 4;
  const int tail = len - head - (nfull << 4);
  const uint8_t* lb = (const uint8_t*)lds;
  if (l < head) gstore(dst + l, lb[l]);
  if (l >= 16 && l < 16 + tail) {
    const int i = head + (nfull << 4) + (l - 16);
    gstore(dst + i, lb[i]);
  }
  for (int j = l; j < nfull; j += 64) {
    const int p = head + (j << 4);
    gstore((uint4*)(dst + p), lds_read16(lds, p));
  }
}

// Copy len bytes from LDS (byte 0 of lds) to global memory at an arbitrary
// byte address: bytes up to the first 16-byte boundary and after the last
// one are stored singly (neighbouring chunks own the rest of those lines),
// the body with 16-byte stores.  Cooperative over the workgroup.
__device__ __forceinline__ void store_unaligned(uint8_t* dst, const uint32_t* lds, int len)
{
  const int t = threadIdx.x;
  const int mis = (int)((uintptr_t)dst & 15);
  const int head = min(len, (16 - mis) & 15);
  const int nfull = (len - head) >> 4;
  const int tail = len - head - (nfull << 4);
  const uint8_t* lb = (const uint8_t*)lds;
  if (t < head) gstore(dst + t, lb[t]);
  if (t >= 64 && t < 64 + tail) {
    const int i = head + (nfull << 4) + (t - 64);
    gstore(dst + i, lb[i]);
  }
  for (int j = t; j < nfull; j += kThreads) {
    const int p = head + (j << 4);
    gstore((uint4*)(dst + p), lds_read16(lds, p));
  }
}

}  // namespace sleek
