// Host-visible declarations shared by the kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sleek {

// look-back groups of 2^shift chunks: 8 .. 64 (one lane per chunk)
constexpr int kGroupShiftMin = 3;
constexpr int kGroupShiftMax = 6;

struct EncodeArgs {
  const uint8_t* in;      // input array (16-byte aligned)
  long long insize;       // bytes
  uint8_t* out;           // stream (8-byte aligned, >= upper-bound size)
  long long* outsize;     // device: stream length in bytes
  uint64_t* ws;           // zeroed workspace: status and look-back words
  int eb_e, thr_e;        // lossy parameters (ignored for lossless)
  unsigned long long offs;
  int lb_shift;           // look-back groups of 2^lb_shift chunks (3..6)
  int lb_window;          // initial look-back window in groups (1..64)
  unsigned int* tickets;  // zeroed chunk counters of the ticket encoders (kTicketSets, kTicketStride apart)
  int ticket_shift;       // 2^ticket_shift ticket sets are in use: ticket_shift_for(grid)
  unsigned long long* help;     // device help state (help_state()): [0] help events, [1] help threshold in naps
};

// Chunk tickets (kEncodePairTicket): every chunk, a workgroup's first
// included, comes from one of S counters, S the largest power of two up to
// min(kTicketSets, grid) (home set x = blockIdx.x mod S): ticket t of set x
// is chunk S * t + x.  A chunk is therefore only ever held by a running
// workgroup, and every set has one, so the look-back never waits on a
// workgroup that has not started.  Many sets keep the counters' queues
// short: each takes about grid / S requests per chunk time, and all of them
// at once when the launch starts.  The counters sit kTicketStride dwords (a
// 128-byte line) apart, after the look-back words of the workspace.
constexpr int kTicketSets = 64;
constexpr int kTicketStride = 32;
constexpr long long kTicketWords = kTicketSets * kTicketStride / 2;   // 8-byte workspace words
static_assert(kTicketSets <= 64 && (kTicketSets & (kTicketSets - 1)) == 0,
              "ticket sets: a power of two, at most one per lane of the fallback's wave");
// log2 of the ticket sets a launch of `grid` workgroups uses: the largest
// power of two up to min(kTicketSets, grid), so that every set has a
// workgroup of its own
inline int ticket_shift_for(long long grid)
{
  int s = 0;
  while ((2ll << s) <= grid && (2 << s) <= kTicketSets) s++;
  return s;
}

struct DecodeArgs {
  const uint8_t* in;      // stream (8-byte aligned)
  long long inlen;        // stream bytes available, or -1 if unknown
  uint8_t* out;           // decoded array (16-byte aligned)
  long long cap;          // capacity of out in bytes, or -1 if unknown
  long long* decsize;     // device: decoded length in bytes
  uint64_t* ws;           // zeroed workspace
  int thr_e;
  unsigned long long offs;
};

// Encoder variants.  kEncodePipelined: 256-thread workgroups, three-stage
// pipeline, register budget of 4 waves per SIMD (the most its two LDS
// staging buffers allow); kEncodeFree: the same kernel at the compiler's
// register choice; variant 3 (one unpipelined wave per chunk) is retired and
// rejected; kEncodeWavePipe: one wave per chunk, pipelined over the wave's
// chunks; kEncodePair: two waves per chunk (128 threads), pipelined the same
// way; kEncodePairTicket: kEncodePair with chunks handed out in arrival
// order (chunk tickets); kEncodeWaveTicket: kEncodeWavePipe with chunk
// tickets.  kEncodeDefault resolves to kEncodePairTicket for both element
// types: with its default 64-group first look-back window it is about 5 %
// faster than kEncodePair with that one's default 8; both two-wave forms run
// 4 waves per SIMD and beat kEncodeWavePipe by 4-8 %.  The ticket encoders
// hand chunks to running workgroups only; the others take chunks in launch
// order and rely on the look-back's help when a workgroup holding a chunk
// has not started.  Output bytes never depend on the variant or on the grid
// size.
constexpr int kEncodeDefault = 0;
constexpr int kEncodePipelined = 1;
constexpr int kEncodeFree = 2;
constexpr int kEncodeRetired = 3;
constexpr int kEncodeWavePipe = 4;
constexpr int kEncodePair = 5;
constexpr int kEncodePairTicket = 6;
constexpr int kEncodeWaveTicket = 7;
constexpr int kEncodeVariants = 8;
inline bool ticket_variant(int resolved) { return resolved == kEncodePairTicket || resolved == kEncodeWaveTicket; }
inline int resolve_variant(int v, int elem_bytes)
{
  (void)elem_bytes;   // the same choice for both element types
  return v != kEncodeDefault ? v : kEncodePairTicket;
}

// The first look-back window (in groups) used when none is set, for the
// single-array and the batched encoders alike: 64 with chunk (tile) tickets,
// where the nearest inclusive group word is about one iteration (64 groups)
// back, so that one poll usually settles it; 8 for the others (each measured
// best for its encoder).
inline int default_window(int resolved_variant) { return ticket_variant(resolved_variant) ? 64 : 8; }

// A look-back's naps (s_sleep 2: 128 clocks, about 53 ns at 2.4 GHz) per
// microsecond of its own waiting, and the default wait before it sizes a
// missing predecessor itself: about 100 us.
constexpr unsigned int kHelpNapsPerUs = 19;
constexpr unsigned int kHelpNapsDefault = 100 * kHelpNapsPerUs;
// The device address of the calling device's help state (null on a HIP
// error): [0] the count of help events (kernels add one per missing word
// they size themselves), [1] the help threshold in naps (a look-back that has
// slept this many naps of 128 clocks without progress sizes the missing words
// itself).
unsigned long long* help_state();
// Sets the help threshold (naps) of every device: those whose help state has
// been looked up now (a synchronous copy each), the others when it is.
hipError_t set_help_naps(unsigned long long naps);

// The no-allocation compression (sleek_gpu_compress_noalloc): inputs up to
// kNoAllocMaxBytes (4 GiB, 262,144 chunks), whose workspace fits the
// library's static device scratch of kNoAllocScratchBytes.
constexpr long long kNoAllocMaxBytes = 4ll << 30;
constexpr long long kNoAllocChunks = kNoAllocMaxBytes / 16384;
constexpr long long kNoAllocScratchBytes =
    (16 + kNoAllocChunks + 2 * (kNoAllocChunks >> kGroupShiftMin) + kTicketWords) * 8;
// The current device's address of that scratch (null on a HIP error).
void* noalloc_scratch();

hipError_t encode_launch(int elem_bytes, bool lossy, const EncodeArgs& a, int grid, int variant, hipStream_t stream);
// Clears `words` 8-byte words with vector stores (no memset node when the
// caller captures a graph).  p is 8-byte aligned.
hipError_t zero_launch(uint64_t* p, long long words, hipStream_t stream);
// Decoder variants: kDecodeScan (default): an offset-scan launch, then one
// workgroup per chunk; kDecodeFused: one persistent launch that derives
// chunk offsets from windows of the size table (faster for lossless
// streams, slower for small lossy payloads).  Same bytes.
constexpr int kDecodeScan = 0;
constexpr int kDecodeFused = 1;
constexpr int kDecodeVariants = 2;
// ev_begin / ev_end (may be null): recorded on the stream right before and
// right after the decoder kernel (after the offset scan).
hipError_t decode_launch(int elem_bytes, bool lossy, const DecodeArgs& a, int grid, int variant, hipStream_t stream,
                         hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
int resident_grid(int elem_bytes, bool lossy, bool decode, int variant);

// One array of a batched call; the layout of sleek_gpu_batch_item.
struct BatchItem {
  const uint8_t* src;
  uint8_t* dst;
  long long src_bytes;
  long long dst_bytes;
  long long* d_size;
  unsigned long long offs;
  int eb_e, thr_e;
  int elem_bytes;
  int lossy;
  long long tile0;
  long long tiles;
  long long ws_word;
  long long ws_words;
};
static_assert(sizeof(BatchItem) == 96, "BatchItem layout");

// Batch workspace: kBatchHead header words, one status word per array,
// the tile map (uint32 per tile), then each array's block (same layout as a
// single-array workspace).  All regions start at even words (16 bytes).
constexpr long long kBatchHead = 16;
inline long long even_words(long long w) { return (w + 1) & ~1ll; }
inline long long batch_map_word(int n) { return kBatchHead + even_words(n); }

// pair: the two-wave batched encoder with tile tickets (counters: tickets,
// kTicketWords words); else the one-wave batched encoder, tiles in launch
// order
hipError_t batch_launch(bool decompress, const BatchItem* items, int n, long long tiles, uint64_t* ws, int grid,
                        int lb_shift, int lb_window, bool pair, uint32_t* tickets, hipStream_t stream);
hipError_t batch_status_launch(const BatchItem* items, int n, uint64_t* ws, hipStream_t stream);
int batch_resident_grid(bool decompress, bool pair);

}  // namespace sleek
