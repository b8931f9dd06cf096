/*
 * SLEEK CPU oracle, type-generic body -- TEST INFRASTRUCTURE ONLY.
 *
 * Included twice by sleek_oracle.c with
 *   W       unsigned word type (uint32_t for float32, uint64_t for float64)
 *   WB      bits per word (32 / 64)
 *   EXPB    exponent bits (8 / 11)
 *   MANB    mantissa bits (23 / 52)
 *   SFX(x)  name suffixing macro (x##_u32 / x##_u64)
 *
 * Every function cites the reference lines it restates.  The arithmetic is
 * done in unsigned W, which is what the reference's mix of signed and
 * unsigned expressions computes modulo 2^WB (its signed shifts and
 * decrements wrap on the compilers it targets).
 */

#define VALS_PER_SUB (SLEEK_ORACLE_CS / (int)sizeof(W) / SLEEK_ORACLE_SC) /* 128 or 64 */
#define VALS_PER_CHUNK (SLEEK_ORACLE_CS / (int)sizeof(W))
#define TOPMASK ((W)((((W)1 << EXPB) - 1)) << (MANB + 1))   /* exponent field after rotate */
#define HALF ((W)1 << (WB - 1))
#define ONE_EXP ((W)1 << (MANB + 1))                         /* one unit of exponent after rotate */

/* lossless forward transform: sleek_common_cpu.h:152-160 (f32), :162-169 (f64).
 * rotate the sign to bit 0, re-centre a non-zero exponent, then zigzag. */
W SFX(sleek_oracle_fwd)(W v)
{
  v = (v << 1) | (v >> (WB - 1));
  if ((v & TOPMASK) != 0) {
    v -= HALF;
    if (((v & TOPMASK) == 0) || (v >= HALF)) v -= ONE_EXP;
  }
  /* zigzag: shift left, xor with the replicated top bit */
  W s = (v >> (WB - 1)) ? ~(W)0 : (W)0;
  return (v << 1) ^ s;
}

/* lossless inverse transform: sleek_common_cpu.h:317-325 (f32), :327-334 (f64) */
W SFX(sleek_oracle_inv)(W v)
{
  W s = (v & 1) ? ~(W)0 : (W)0;
  v = (v >> 1) ^ s;
  if ((v & TOPMASK) != 0) {
    if (v >= HALF) v += ONE_EXP;
    v += HALF;
  }
  return (v << (WB - 1)) | (v >> 1);
}

/* quantize: sleek_common_cpu.h:393-419 */
W SFX(sleek_oracle_quantize)(W val, int eb_e, int thr_e, W offs)
{
  const W absv = val & (((W)1 << (EXPB + MANB)) - 1);
  const int val_e = (int)(absv >> MANB);
  W enc = 0;
  if (val_e >= thr_e) {
    enc = absv - offs;
  } else if (val_e >= eb_e) {
    W mant = (val & (((W)1 << MANB) - 1)) | ((W)1 << MANB);
    const int shift = thr_e - val_e;
    mant += (W)1 << (shift - 1);
    enc = mant >> shift;
  }
  enc = (enc << 1) | ((~val) >> (EXPB + MANB));
  if (enc != 0) enc--;
  return enc;
}

static inline int SFX(clz_w)(W x)
{
#if WB == 32
  return __builtin_clz(x);
#else
  return __builtin_clzll(x);
#endif
}

/* deQuantize: sleek_common_cpu.h:559-591 */
W SFX(sleek_oracle_dequantize)(W enc, int thr_e, W offs)
{
  W dec = 0;
  if (enc != 0) {
    const W absv = (W)(enc + 1) >> 1;
    if (absv >= ((W)1 << MANB)) {
      dec = absv + offs;
    } else if (absv > 0) {
      const int shift = SFX(clz_w)(absv) - (WB - 1 - MANB);
      dec = absv << shift;
      dec &= ((W)1 << MANB) - 1;
      dec |= (W)(thr_e - shift) << MANB;
    }
    dec |= enc << (EXPB + MANB);
  }
  return dec;
}

/* Code one (zero-padded) chunk: per-value codes and per-subchunk bit
 * lengths.  sleek_common_cpu.h:134-184 (lossless), :455-485 (lossy).
 * Returns the total number of packed bits. */
static int SFX(code_chunk)(const uint8_t* src, int osize, W* codes, uint8_t logn[SLEEK_ORACLE_SC],
                           int lossy, int eb_e, int thr_e, W offs)
{
  W vals[VALS_PER_CHUNK];
  memset(vals, 0, sizeof(vals));
  memcpy(vals, src, (size_t)osize);          /* zero padding of a partial chunk */
  int bits = 0;
  for (int s = 0; s < SLEEK_ORACLE_SC; s++) {
    W mx = 0;
    for (int j = s * VALS_PER_SUB; j < (s + 1) * VALS_PER_SUB; j++) {
      const W c = lossy ? SFX(sleek_oracle_quantize)(vals[j], eb_e, thr_e, offs) : SFX(sleek_oracle_fwd)(vals[j]);
      codes[j] = c;
      if (c > mx) mx = c;
    }
    const int ln = (mx == 0) ? 0 : WB - SFX(clz_w)(mx);
    logn[s] = (uint8_t)ln;
    bits += ln * VALS_PER_SUB;
  }
  return bits;
}

/* Pack one compressed chunk into dst (newsize bytes).
 * sleek_common_cpu.h:204-238: logn bytes, LSB-first codes in W words, int16 osize trailer. */
static void SFX(pack_chunk)(const W* codes, const uint8_t logn[SLEEK_ORACLE_SC], int bits, int osize, uint8_t* dst)
{
  W words[VALS_PER_CHUNK + 1];
  const int nwords = bits / WB;
  memset(words, 0, sizeof(W) * (size_t)(nwords + 1));
  int pos = 0;                                  /* bit position in the packed area */
  for (int s = 0; s < SLEEK_ORACLE_SC; s++) {
    const int ln = logn[s];
    for (int j = s * VALS_PER_SUB; j < (s + 1) * VALS_PER_SUB && ln > 0; j++) {
      const W c = codes[j];
      const int wi = pos / WB, sh = pos % WB;
      words[wi] |= c << sh;
      if (sh != 0 && WB - sh < ln) words[wi + 1] |= c >> (WB - sh);
      pos += ln;
    }
  }
  memcpy(dst, logn, SLEEK_ORACLE_SC);
  memcpy(dst + SLEEK_ORACLE_SC, words, (size_t)bits / 8);
  const int16_t tr = (int16_t)osize;
  memcpy(dst + SLEEK_ORACLE_SC + bits / 8, &tr, 2);
}

/* Whole-array encoder: sleek_common_cpu.h:101-256 (lossless), :422-556 (lossy).
 * The reference overlaps a spin-waited carry chain with packing; here the
 * chunk sizes are found first, prefix-summed, and the chunks are then
 * written, which yields the same bytes. */
long long SFX(sleek_oracle_encode)(const uint8_t* in, long long insize, uint8_t* out,
                                   int lossy, int eb_e, int thr_e, W offs)
{
  if (insize < 0 || out == NULL || (insize > 0 && in == NULL)) return SLEEK_ORACLE_EARG;
  const long long chunks = (insize + SLEEK_ORACLE_CS - 1) / SLEEK_ORACLE_CS;
  uint8_t* const data = out + 8 + 2 * chunks;
  long long* const start = (long long*)malloc(sizeof(long long) * (size_t)(chunks + 1));
  uint16_t* const csz = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(chunks + 1));
  if (!start || !csz) { free(start); free(csz); return SLEEK_ORACLE_EARG; }

  /* pass 1: chunk sizes (the fit test of sleek_common_cpu.h:198 / :498) */
#pragma omp parallel for schedule(dynamic, 4)
  for (long long c = 0; c < chunks; c++) {
    W codes[VALS_PER_CHUNK];
    uint8_t logn[SLEEK_ORACLE_SC];
    const long long base = c * SLEEK_ORACLE_CS;
    const int osize = (int)((insize - base) < SLEEK_ORACLE_CS ? (insize - base) : SLEEK_ORACLE_CS);
    const int bits = SFX(code_chunk)(in + base, osize, codes, logn, lossy, eb_e, thr_e, offs);
    const int newsize = (SLEEK_ORACLE_SC * 8 + bits + 16) / 8;
    csz[c] = (uint16_t)((newsize < SLEEK_ORACLE_CS && newsize < osize) ? newsize : osize);
  }
  long long acc = 0;
  for (long long c = 0; c < chunks; c++) { start[c] = acc; acc += csz[c]; }

  /* pass 2: payloads */
#pragma omp parallel for schedule(dynamic, 4)
  for (long long c = 0; c < chunks; c++) {
    const long long base = c * SLEEK_ORACLE_CS;
    const int osize = (int)((insize - base) < SLEEK_ORACLE_CS ? (insize - base) : SLEEK_ORACLE_CS);
    if (csz[c] == osize) {
      memcpy(data + start[c], in + base, (size_t)osize);   /* raw chunk keeps the original bytes */
    } else {
      W codes[VALS_PER_CHUNK];
      uint8_t logn[SLEEK_ORACLE_SC];
      const int bits = SFX(code_chunk)(in + base, osize, codes, logn, lossy, eb_e, thr_e, offs);
      SFX(pack_chunk)(codes, logn, bits, osize, data + start[c]);
    }
  }
  /* header (sleek_common_cpu.h:251) and size table (:202, :245), little-endian */
  const int64_t hdr = insize;
  memcpy(out, &hdr, 8);
  memcpy(out + 8, csz, 2 * (size_t)chunks);
  free(start);
  free(csz);
  return 8 + 2 * chunks + acc;
}

/* Whole-array decoder: sleek_common_cpu.h:259-385 (lossless), :594-685 (lossy).
 * Structural checks beyond the reference's trailer test (:376-378) make a
 * malformed stream an error instead of undefined behaviour. */
int SFX(sleek_oracle_decode)(const uint8_t* in, long long inlen, uint8_t* out, long long outcap,
                             long long* outsize, int lossy, int thr_e, W offs)
{
  if (in == NULL || inlen < 8) return SLEEK_ORACLE_EARG;
  int64_t hdr;
  memcpy(&hdr, in, 8);
  if (hdr < 0) return SLEEK_ORACLE_ECORRUPT;
  if (outsize) *outsize = hdr;
  if (hdr > outcap) return SLEEK_ORACLE_ECAP;
  if (hdr > 0 && out == NULL) return SLEEK_ORACLE_EARG;
  const long long chunks = (hdr + SLEEK_ORACLE_CS - 1) / SLEEK_ORACLE_CS;
  if (8 + 2 * chunks > inlen) return SLEEK_ORACLE_ECORRUPT;
  const uint8_t* const data = in + 8 + 2 * chunks;
  long long* const start = (long long*)malloc(sizeof(long long) * (size_t)(chunks + 1));
  if (!start) return SLEEK_ORACLE_EARG;
  long long acc = 0;
  int bad = 0;
  for (long long c = 0; c < chunks; c++) {
    uint16_t cs;
    memcpy(&cs, in + 8 + 2 * c, 2);
    const long long base = c * SLEEK_ORACLE_CS;
    const int osize = (int)((hdr - base) < SLEEK_ORACLE_CS ? (hdr - base) : SLEEK_ORACLE_CS);
    if (cs > osize || (cs < osize && cs < SLEEK_ORACLE_SC + 2)) bad = 1;
    start[c] = acc;
    acc += cs;
  }
  if (bad || 8 + 2 * chunks + acc > inlen) { free(start); return SLEEK_ORACLE_ECORRUPT; }

#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
  for (long long c = 0; c < chunks; c++) {
    uint16_t cs;
    memcpy(&cs, in + 8 + 2 * c, 2);
    const long long base = c * SLEEK_ORACLE_CS;
    const int osize = (int)((hdr - base) < SLEEK_ORACLE_CS ? (hdr - base) : SLEEK_ORACLE_CS);
    const uint8_t* const p = data + start[c];
    if (cs == osize) { memcpy(out + base, p, (size_t)osize); continue; }
    int bits = 0;
    for (int s = 0; s < SLEEK_ORACLE_SC; s++) {
      if (p[s] > WB) bad = 1;
      bits += p[s] * VALS_PER_SUB;
    }
    if (bad || SLEEK_ORACLE_SC + 2 + bits / 8 != cs) { bad = 1; continue; }
    int16_t tr;
    memcpy(&tr, p + cs - 2, 2);
    if (tr != osize) { bad = 1; continue; }
    W words[VALS_PER_CHUNK + 1];
    memcpy(words, p + SLEEK_ORACLE_SC, (size_t)bits / 8);
    words[bits / WB] = 0;
    W vals[VALS_PER_CHUNK];
    int pos = 0;
    for (int s = 0; s < SLEEK_ORACLE_SC; s++) {
      const int ln = p[s];
      const W mask = (ln == WB) ? ~(W)0 : (((W)1 << ln) - 1);
      for (int j = s * VALS_PER_SUB; j < (s + 1) * VALS_PER_SUB; j++) {
        W r = 0;
        if (ln > 0) {
          const int wi = pos / WB, sh = pos % WB;
          r = words[wi] >> sh;
          if (sh != 0 && WB - sh < ln) r |= words[wi + 1] << (WB - sh);
          r &= mask;
        }
        pos += ln;
        vals[j] = lossy ? SFX(sleek_oracle_dequantize)(r, thr_e, offs) : SFX(sleek_oracle_inv)(r);
      }
    }
    memcpy(out + base, vals, (size_t)osize);
  }
  free(start);
  return bad ? SLEEK_ORACLE_ECORRUPT : SLEEK_ORACLE_OK;
}

#undef VALS_PER_SUB
#undef VALS_PER_CHUNK
#undef TOPMASK
#undef HALF
#undef ONE_EXP
