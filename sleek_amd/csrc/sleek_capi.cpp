// C ABI of the SLEEK GPU codec (declared in include/sleek_gpu.h).
//
// Argument checks, workspace management and kernel launches.  The
// synchronous entry points mirror the reference host API
// (sleek_lossless_gpu.h:47-100, sleek_lossy_gpu.h:56-127) with status codes
// in place of exit() / device traps; the *_async entry points are
// stream-ordered and take a caller-owned workspace.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/sleek_gpu.h"
#include "sleek_internal.h"

namespace {

thread_local char g_last_error[256] = "";

int hip_fail(hipError_t e, const char* what)
{
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
  return SLEEK_EHIP;
}

long long chunks_of(long long bytes) { return (bytes + SLEEK_CHUNK_BYTES - 1) / SLEEK_CHUNK_BYTES; }

// workspace: 8 header words (word 1: status), then per chunk one word, and
// per look-back group (8 chunks at the smallest setting) two words (look-back
// state when encoding; chunk offsets and their total when decoding); 8 spare
// words, an even total.  A batch item's workspace is this much.
long long ws_words_for(long long insize)
{
  const long long ch = chunks_of(insize);
  const long long gmin = 1ll << sleek::kGroupShiftMin;
  return sleek::even_words(16 + ch + 2 * ((ch + gmin - 1) / gmin));
}

// a single-array call's workspace: the above, then the chunk-ticket counters
long long ws_bytes_for(long long insize) { return (ws_words_for(insize) + sleek::kTicketWords) * 8; }

// tuning knobs (sleek_gpu_set_tuning); none changes the bytes produced.
// Atomic: calls on other host threads read them while a setter runs.
std::atomic<int> g_encode_variant{sleek::kEncodeDefault};
std::atomic<int> g_encode_grid{0};
std::atomic<int> g_decode_grid{0};
std::atomic<int> g_decode_variant{sleek::kDecodeScan};
std::atomic<int> g_lb_shift{5};    // look-back groups of 32 chunks
std::atomic<int> g_lb_window{0};   // initial look-back window, in groups (0: the encoder's default)

// Per-device scratch of the synchronous entry points (grow-only).
struct DeviceScratch {
  void* ptr = nullptr;
  long long bytes = 0;
};
std::mutex g_mu;
std::vector<DeviceScratch> g_scratch;

int scratch_for(long long need, void** out)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
  DeviceScratch& s = g_scratch[dev];
  if (s.bytes < need) {
    if (s.ptr) {
      e = hipFree(s.ptr);
      if (e != hipSuccess) return hip_fail(e, "hipFree");
      s.ptr = nullptr;
      s.bytes = 0;
    }
    const long long sz = need < (1 << 20) ? (1 << 20) : need;
    e = hipMalloc(&s.ptr, (size_t)sz);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    s.bytes = sz;
  }
  *out = s.ptr;
  return SLEEK_OK;
}

// sleek_gpu_set_kernel_events: events the calling thread's next codec call
// records around its coding kernel (used once, then cleared).
struct KernelEvents {
  hipEvent_t begin = nullptr, end = nullptr;
};
thread_local KernelEvents t_kernel_events;

KernelEvents take_kernel_events()
{
  const KernelEvents k = t_kernel_events;
  t_kernel_events = KernelEvents();
  return k;
}

int grid_for(int blocks, int elem_bytes, bool lossy, bool decode, long long nchunks)
{
  long long g = blocks;
  if (g <= 0) g = decode ? g_decode_grid : g_encode_grid;
  if (decode) {
    if (g <= 0) g = g_decode_variant == sleek::kDecodeScan ? nchunks : sleek::resident_grid(elem_bytes, lossy, true, 0);
  } else {
    // the encoder's look-back waits on chunks in launch order: a grid beyond
    // the resident set only adds waits, so it is capped there
    const int res = sleek::resident_grid(elem_bytes, lossy, false, g_encode_variant);
    if (g <= 0 || (res > 0 && g > res)) g = res;
  }
  if (g <= 0) g = 1024;
  if (nchunks >= 0 && g > nchunks) g = nchunks > 0 ? nchunks : 1;
  if (g > (1 << 20)) g = 1 << 20;
  return (int)g;
}

bool aligned(const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; }

int compress_impl(int elem_bytes, bool lossy, int eb_e, int thr_e, unsigned long long offs, const void* d_input,
                  long long insize, unsigned char* d_encoded, long long* d_encsize, void* ws, long long ws_bytes,
                  int blocks, hipStream_t stream)
{
  if ((elem_bytes != 4 && elem_bytes != 8) || insize < 0 || insize % elem_bytes != 0) return SLEEK_EINVAL;
  if (chunks_of(insize) > 0xffffffffll) return SLEEK_EINVAL;   // chunk indices are 32-bit in the kernels' tickets
  if (!d_encoded || !d_encsize || !ws || (insize > 0 && !d_input)) return SLEEK_EINVAL;
  if (!aligned(d_input, 16) || !aligned(d_encoded, 8) || !aligned(d_encsize, 8) || !aligned(ws, 8)) return SLEEK_EALIGN;
  const long long need = ws_bytes_for(insize);
  if (ws_bytes < need) return SLEEK_EWORKSPACE;
  // cleared with stores, not a memset, so that callers may capture the call
  hipError_t e = sleek::zero_launch((uint64_t*)ws, need / 8, stream);
  if (e != hipSuccess) return hip_fail(e, "workspace clear");
  sleek::EncodeArgs a;
  a.in = (const uint8_t*)d_input;
  a.insize = insize;
  a.out = d_encoded;
  a.outsize = d_encsize;
  a.ws = (uint64_t*)ws;
  a.eb_e = eb_e;
  a.thr_e = thr_e;
  a.offs = offs;
  a.lb_shift = g_lb_shift;
  a.lb_window = g_lb_window > 0 ? (int)g_lb_window : sleek::default_window(sleek::resolve_variant(g_encode_variant, elem_bytes));
  a.tickets = (unsigned int*)((uint64_t*)ws + ws_words_for(insize));
  a.help = sleek::help_state();
  const int grid = grid_for(blocks, elem_bytes, lossy, false, chunks_of(insize));
  a.ticket_shift = sleek::ticket_shift_for(grid);
  const KernelEvents ev = take_kernel_events();
  if (ev.begin && (e = hipEventRecord(ev.begin, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  e = sleek::encode_launch(elem_bytes, lossy, a, grid, g_encode_variant, stream);
  if (e != hipSuccess) return hip_fail(e, "encode launch");
  if (ev.end && (e = hipEventRecord(ev.end, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  return SLEEK_OK;
}

int decompress_impl(int elem_bytes, bool lossy, int thr_e, unsigned long long offs, const unsigned char* d_encoded,
                    long long enclen, void* d_decoded, long long capacity, long long* d_decsize, void* ws,
                    long long ws_bytes, int blocks, hipStream_t stream)
{
  if ((elem_bytes != 4 && elem_bytes != 8) || !d_encoded || !d_decsize || !ws) return SLEEK_EINVAL;
  if (enclen >= 0 && enclen < 8) return SLEEK_ECORRUPT;
  if (!aligned(d_decoded, 16) || !aligned(d_encoded, 8) || !aligned(d_decsize, 8) || !aligned(ws, 8)) return SLEEK_EALIGN;
  if (capacity >= 0) {
    if (ws_bytes < ws_bytes_for(capacity)) return SLEEK_EWORKSPACE;
  }
  // without a capacity the workspace bounds the chunk count
  const long long ws_chunks = ws_bytes / 8 - 8;
  const long long cap = capacity >= 0 ? capacity : (ws_chunks - 8) * SLEEK_CHUNK_BYTES;
  if (cap < 0) return SLEEK_EWORKSPACE;
  // only the status word needs clearing (the offsets are written); the
  // scan decoder's offset kernel clears it itself
  hipError_t e = hipSuccess;
  const int variant = g_decode_variant;   // one value for the clear and the launch
  if (variant != sleek::kDecodeScan) {
    e = sleek::zero_launch((uint64_t*)ws, 8, stream);
    if (e != hipSuccess) return hip_fail(e, "workspace clear");
  }
  sleek::DecodeArgs a;
  a.in = d_encoded;
  a.inlen = enclen;
  a.out = (uint8_t*)d_decoded;
  a.cap = cap;
  a.decsize = d_decsize;
  a.ws = (uint64_t*)ws;
  a.thr_e = thr_e;
  a.offs = offs;
  const int grid = grid_for(blocks, elem_bytes, lossy, true, chunks_of(cap));
  const KernelEvents ev = take_kernel_events();
  e = sleek::decode_launch(elem_bytes, lossy, a, grid, variant, stream, ev.begin, ev.end);
  if (e != hipSuccess) return hip_fail(e, "decode launch");
  return SLEEK_OK;
}

int read_status(const void* ws, hipStream_t stream)
{
  hipError_t e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  unsigned long long st = 0;
  e = hipMemcpy(&st, (const uint64_t*)ws + 1, 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy status");
  if (st & 2) return SLEEK_ECAPACITY;
  if (st & 1) return SLEEK_ECORRUPT;
  return SLEEK_OK;
}

int compress_sync(int elem_bytes, bool lossy, int eb_e, int thr_e, unsigned long long offs, const void* d_input,
                  long long num_elements, unsigned char* d_encoded, long long* d_encsize, int blocks)
{
  if (num_elements < 0) return SLEEK_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  const long long insize = num_elements * elem_bytes;
  void* ws = nullptr;
  int rc = scratch_for(ws_bytes_for(insize), &ws);
  if (rc) return rc;
  rc = compress_impl(elem_bytes, lossy, eb_e, thr_e, offs, d_input, insize, d_encoded, d_encsize, ws,
                     ws_bytes_for(insize), blocks, 0);
  if (rc) return rc;
  return read_status(ws, 0);
}

int decompress_sync(int elem_bytes, bool lossy, int thr_e, unsigned long long offs, const unsigned char* d_encoded,
                    void* d_decoded, long long* d_decsize, int blocks)
{
  if (!d_encoded) return SLEEK_EINVAL;
  if (!aligned(d_encoded, 8)) return SLEEK_EALIGN;
  std::lock_guard<std::mutex> lk(g_mu);
  // the decoded length sizes the scratch (the reference's callers read it the same way)
  long long hdr = 0;
  hipError_t e = hipMemcpy(&hdr, d_encoded, 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy header");
  if (hdr < 0) return SLEEK_ECORRUPT;
  void* ws = nullptr;
  int rc = scratch_for(ws_bytes_for(hdr), &ws);
  if (rc) return rc;
  rc = decompress_impl(elem_bytes, lossy, thr_e, offs, d_encoded, -1, d_decoded, -1, d_decsize, ws,
                       ws_bytes_for(hdr), blocks, 0);
  if (rc) return rc;
  return read_status(ws, 0);
}

}  // namespace

extern "C" {

long long sleek_gpu_upper_bound_size(long long insize)
{
  const long long chunks = chunks_of(insize);
  return 3 * 4 + chunks * 2 + chunks * SLEEK_CHUNK_BYTES;
}

int sleek_gpu_system_checks(void)
{
  const int one = 1;
  if (*(const char*)&one != 1) return SLEEK_EINVAL;
  if (sizeof(long long) != 8 || sizeof(int) != 4 || sizeof(short) != 2) return SLEEK_EINVAL;
  return SLEEK_OK;
}

int sleek_gpu_thread_blocks(void)
{
  const int g = sleek::resident_grid(4, false, false, g_encode_variant);
  return g > 0 ? g : SLEEK_EHIP;
}

int sleek_gpu_error_bound_f32(float eb, int* eb_e, int* thr_e, unsigned int* offs_e)
{
  if (!eb_e || !thr_e || !offs_e) return SLEEK_EINVAL;
  uint32_t bits;
  memcpy(&bits, &eb, 4);
  const int e = (int)((bits >> 23) & 0xffu);
  const int t = e + 24;
  if (t >= 255) return SLEEK_EEB;
  *eb_e = e;
  *thr_e = t;
  *offs_e = ((uint32_t)t << 23) - (1u << 23);
  return SLEEK_OK;
}

int sleek_gpu_error_bound_f64(double eb, int* eb_e, int* thr_e, unsigned long long* offs_e)
{
  if (!eb_e || !thr_e || !offs_e) return SLEEK_EINVAL;
  uint64_t bits;
  memcpy(&bits, &eb, 8);
  const int e = (int)((bits >> 52) & 0x7ffu);
  const int t = e + 53;
  if (t >= 2047) return SLEEK_EEB;
  *eb_e = e;
  *thr_e = t;
  *offs_e = ((unsigned long long)t << 52) - (1ull << 52);
  return SLEEK_OK;
}

const char* sleek_gpu_strerror(int status)
{
  switch (status) {
    case SLEEK_OK: return "ok";
    case SLEEK_EINVAL: return "invalid argument";
    case SLEEK_EHIP: return "HIP runtime error";
    case SLEEK_ECORRUPT: return "malformed compressed stream";
    case SLEEK_EEB: return "error bound is too large";
    case SLEEK_ECAPACITY: return "decoded size exceeds the output capacity";
    case SLEEK_EWORKSPACE: return "workspace too small";
    case SLEEK_EALIGN: return "pointer not aligned (16 B for arrays, 8 B for streams)";
    case SLEEK_ETOOLARGE: return "input larger than the no-allocation entry point's scratch allows (4 GiB)";
    default: return "unknown status";
  }
}

const char* sleek_gpu_last_error(void) { return g_last_error; }

int sleek_gpu_compress_f32(int blocks, const float* d_input, long long n, unsigned char* d_encoded, long long* d_encsize)
{
  return compress_sync(4, false, 0, 0, 0, d_input, n, d_encoded, d_encsize, blocks);
}

int sleek_gpu_decompress_f32(int blocks, const unsigned char* d_encoded, float* d_decoded, long long* d_decsize)
{
  return decompress_sync(4, false, 0, 0, d_encoded, d_decoded, d_decsize, blocks);
}

int sleek_gpu_compress_f64(int blocks, const double* d_input, long long n, unsigned char* d_encoded, long long* d_encsize)
{
  return compress_sync(8, false, 0, 0, 0, d_input, n, d_encoded, d_encsize, blocks);
}

int sleek_gpu_decompress_f64(int blocks, const unsigned char* d_encoded, double* d_decoded, long long* d_decsize)
{
  return decompress_sync(8, false, 0, 0, d_encoded, d_decoded, d_decsize, blocks);
}

int sleek_gpu_compress_lossy_f32(int blocks, int eb_e, int thr_e, unsigned int offs_e, const float* d_input,
                                 long long n, unsigned char* d_encoded, long long* d_encsize)
{
  return compress_sync(4, true, eb_e, thr_e, offs_e, d_input, n, d_encoded, d_encsize, blocks);
}

int sleek_gpu_decompress_lossy_f32(int blocks, int eb_e, int thr_e, unsigned int offs_e,
                                   const unsigned char* d_encoded, float* d_decoded, long long* d_decsize)
{
  (void)eb_e;
  return decompress_sync(4, true, thr_e, offs_e, d_encoded, d_decoded, d_decsize, blocks);
}

int sleek_gpu_compress_lossy_f64(int blocks, int eb_e, int thr_e, unsigned long long offs_e, const double* d_input,
                                 long long n, unsigned char* d_encoded, long long* d_encsize)
{
  return compress_sync(8, true, eb_e, thr_e, offs_e, d_input, n, d_encoded, d_encsize, blocks);
}

int sleek_gpu_decompress_lossy_f64(int blocks, int eb_e, int thr_e, unsigned long long offs_e,
                                   const unsigned char* d_encoded, double* d_decoded, long long* d_decsize)
{
  (void)eb_e;
  return decompress_sync(8, true, thr_e, offs_e, d_encoded, d_decoded, d_decsize, blocks);
}

int sleek_gpu_set_tuning(int encode_variant, int encode_grid, int decode_grid)
{
  if (encode_variant < 0 || encode_variant >= sleek::kEncodeVariants || encode_variant == sleek::kEncodeRetired ||
      encode_grid < 0 || decode_grid < 0)
    return SLEEK_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  g_encode_variant = encode_variant;
  g_encode_grid = encode_grid;
  g_decode_grid = decode_grid;
  return SLEEK_OK;
}

int sleek_gpu_set_decoder(int variant)
{
  if (variant < 0 || variant >= sleek::kDecodeVariants) return SLEEK_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  g_decode_variant = variant;
  return SLEEK_OK;
}

int sleek_gpu_set_lookback(int group_chunks, int window_groups)
{
  if (group_chunks < 0 || window_groups < 0 || window_groups > 64) return SLEEK_EINVAL;
  int shift = 5;   // the default: 32 chunks
  if (group_chunks > 0) {
    for (shift = sleek::kGroupShiftMin; shift <= sleek::kGroupShiftMax && (1 << shift) != group_chunks; shift++) {
    }
    if (shift > sleek::kGroupShiftMax) return SLEEK_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_lb_shift = shift;
  g_lb_window = window_groups;   // 0: the encoder's default
  return SLEEK_OK;
}

int sleek_gpu_set_help_after(int microseconds)
{
  if (microseconds < 0 || microseconds > 10000000) return SLEEK_EINVAL;
  const hipError_t e = sleek::set_help_naps((unsigned long long)microseconds * sleek::kHelpNapsPerUs);
  return e == hipSuccess ? SLEEK_OK : hip_fail(e, "help threshold");
}

int sleek_gpu_set_kernel_events(void* begin, void* end)
{
  t_kernel_events.begin = (hipEvent_t)begin;
  t_kernel_events.end = (hipEvent_t)end;
  return SLEEK_OK;
}

long long sleek_gpu_help_events(void)
{
  unsigned long long* p = sleek::help_state();
  if (!p) return hip_fail(hipErrorInvalidSymbol, "help counter");
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  unsigned long long v = 0;
  e = hipMemcpy(&v, p, 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy help counter");
  return (long long)v;
}

long long sleek_gpu_workspace_size(long long max_insize)
{
  if (max_insize < 0) return SLEEK_EINVAL;
  return ws_bytes_for(max_insize);
}

int sleek_gpu_compress_async(int elem_bytes, int lossy, int eb_e, int thr_e, unsigned long long offs_e,
                             const void* d_input, long long insize, unsigned char* d_encoded, long long* d_encsize,
                             void* d_workspace, long long workspace_bytes, int blocks, void* stream)
{
  return compress_impl(elem_bytes, lossy != 0, eb_e, thr_e, offs_e, d_input, insize, d_encoded, d_encsize,
                       d_workspace, workspace_bytes, blocks, (hipStream_t)stream);
}

int sleek_gpu_decompress_async(int elem_bytes, int lossy, int thr_e, unsigned long long offs_e,
                               const unsigned char* d_encoded, long long enclen, void* d_decoded, long long capacity,
                               long long* d_decsize, void* d_workspace, long long workspace_bytes, int blocks,
                               void* stream)
{
  return decompress_impl(elem_bytes, lossy != 0, thr_e, offs_e, d_encoded, enclen, d_decoded, capacity, d_decsize,
                         d_workspace, workspace_bytes, blocks, (hipStream_t)stream);
}

int sleek_gpu_compress_noalloc(int elem_bytes, int lossy, int eb_e, int thr_e, unsigned long long offs_e,
                               const void* d_input, long long insize, unsigned char* d_encoded, long long* d_encsize,
                               int blocks, void* stream)
{
  if (insize < 0) return SLEEK_EINVAL;
  if (insize > sleek::kNoAllocMaxBytes) return SLEEK_ETOOLARGE;
  static_assert(sleek::kNoAllocScratchBytes % 16 == 0, "scratch size");
  void* const ws = sleek::noalloc_scratch();
  if (!ws) return hip_fail(hipErrorInvalidSymbol, "no-allocation scratch");
  const long long need = ws_bytes_for(insize);
  if (need > sleek::kNoAllocScratchBytes) return SLEEK_ETOOLARGE;   // cannot happen below the maximum
  return compress_impl(elem_bytes, lossy != 0, eb_e, thr_e, offs_e, d_input, insize, d_encoded, d_encsize, ws, need,
                       blocks, (hipStream_t)stream);
}

int sleek_gpu_workspace_status(const void* d_workspace, void* stream)
{
  if (!d_workspace) return SLEEK_EINVAL;
  return read_status(d_workspace, (hipStream_t)stream);
}

// ---- batched calls ----

static_assert(sizeof(sleek_gpu_batch_item) == sizeof(sleek::BatchItem), "batch item layout");
static_assert(offsetof(sleek_gpu_batch_item, offs) == offsetof(sleek::BatchItem, offs), "batch item layout");
static_assert(offsetof(sleek_gpu_batch_item, lossy) == offsetof(sleek::BatchItem, lossy), "batch item layout");
static_assert(offsetof(sleek_gpu_batch_item, ws_words) == offsetof(sleek::BatchItem, ws_words), "batch item layout");

int sleek_gpu_batch_plan(sleek_gpu_batch_item* items, int n, int decompress, long long* tiles, long long* workspace_bytes)
{
  if (n < 0 || (n > 0 && !items) || !tiles || !workspace_bytes) return SLEEK_EINVAL;
  long long tile = 0;
  for (int i = 0; i < n; i++) {
    sleek_gpu_batch_item& it = items[i];
    if ((it.elem_bytes != 4 && it.elem_bytes != 8) || (it.lossy != 0 && it.lossy != 1) || !it.d_size)
      return SLEEK_EINVAL;
    // an empty array may come with a null data pointer
    if (decompress ? (!it.src || (!it.dst && it.dst_bytes > 0)) : (!it.dst || (!it.src && it.src_bytes > 0)))
      return SLEEK_EINVAL;
    if (!aligned(it.d_size, 8)) return SLEEK_EALIGN;
    long long bytes;
    if (decompress) {
      if (it.dst_bytes < 0 || (it.src_bytes >= 0 && it.src_bytes < 8)) return SLEEK_EINVAL;
      if (!aligned(it.src, 8) || !aligned(it.dst, 16)) return SLEEK_EALIGN;
      bytes = it.dst_bytes;
    } else {
      if (it.src_bytes < 0 || it.src_bytes % it.elem_bytes != 0) return SLEEK_EINVAL;
      if (!aligned(it.src, 16) || !aligned(it.dst, 8)) return SLEEK_EALIGN;
      bytes = it.src_bytes;
    }
    const long long ch = chunks_of(bytes);
    it.tile0 = tile;
    it.tiles = ch;
    it.ws_words = ws_words_for(bytes);
    tile += it.tiles;
    if (tile > 0xffffffffll) return SLEEK_EINVAL;
  }
  // blocks follow the header, the status words and the tile map
  long long word = sleek::batch_map_word(n) + sleek::even_words((tile + 1) / 2);
  for (int i = 0; i < n; i++) {
    items[i].ws_word = word;
    word += items[i].ws_words;
  }
  *tiles = tile;
  *workspace_bytes = (word + sleek::kTicketWords) * 8;   // then the tile-ticket counters
  return SLEEK_OK;
}

static int batch_async(bool decompress, const sleek_gpu_batch_item* d_items, int n, long long tiles, void* ws,
                       long long ws_bytes, int grid, void* stream)
{
  if (n < 0 || tiles < 0 || (n > 0 && (!d_items || !ws))) return SLEEK_EINVAL;
  if (n == 0) return SLEEK_OK;
  if (!aligned(ws, 16) || !aligned(d_items, 8)) return SLEEK_EALIGN;
  if (ws_bytes < (sleek::batch_map_word(n) + (tiles + 1) / 2) * 8) return SLEEK_EWORKSPACE;
  // the batched encoder: two waves per tile with tile tickets (counters in
  // the workspace's last words), whatever the single-array variant, except
  // that encode variant 4 selects the one-wave encoder with tiles in launch
  // order (a test form that relies on the look-back's help under co-running
  // kernels)
  const int rv = sleek::resolve_variant(g_encode_variant, 4);
  const bool pair = rv != sleek::kEncodeWavePipe;
  const long long even_end = ws_bytes / 16 * 2;   // the counters end at a 16-byte boundary
  if (!decompress && pair && even_end < sleek::batch_map_word(n) + (tiles + 1) / 2 + sleek::kTicketWords)
    return SLEEK_EWORKSPACE;
  uint32_t* const tickets = (uint32_t*)((uint64_t*)ws + even_end - sleek::kTicketWords);
  const int window = g_lb_window > 0 ? (int)g_lb_window : sleek::default_window(pair ? sleek::kEncodePairTicket : rv);
  const hipError_t e = sleek::batch_launch(decompress, (const sleek::BatchItem*)d_items, n, tiles, (uint64_t*)ws, grid,
                                           g_lb_shift, window, pair, tickets, (hipStream_t)stream);
  return e == hipSuccess ? SLEEK_OK : hip_fail(e, "batch launch");
}

int sleek_gpu_batch_compress_async(const sleek_gpu_batch_item* d_items, int n, long long tiles, void* d_workspace,
                                   long long workspace_bytes, int grid, void* stream)
{
  return batch_async(false, d_items, n, tiles, d_workspace, workspace_bytes, grid, stream);
}

int sleek_gpu_batch_decompress_async(const sleek_gpu_batch_item* d_items, int n, long long tiles, void* d_workspace,
                                     long long workspace_bytes, int grid, void* stream)
{
  return batch_async(true, d_items, n, tiles, d_workspace, workspace_bytes, grid, stream);
}

int sleek_gpu_batch_status(const sleek_gpu_batch_item* d_items, int n, void* d_workspace, long long workspace_bytes,
                           int* status, void* stream)
{
  if (n < 0 || (n > 0 && (!d_items || !d_workspace || !status))) return SLEEK_EINVAL;
  if (n == 0) return SLEEK_OK;
  if (workspace_bytes < (sleek::kBatchHead + n) * 8) return SLEEK_EWORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = sleek::batch_status_launch((const sleek::BatchItem*)d_items, n, (uint64_t*)d_workspace, s);
  if (e != hipSuccess) return hip_fail(e, "batch status");
  std::vector<unsigned long long> words((size_t)n);
  e = hipMemcpyAsync(words.data(), (const uint64_t*)d_workspace + sleek::kBatchHead, (size_t)n * 8,
                     hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync status");
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  for (int i = 0; i < n; i++)
    status[i] = (words[i] & 2) ? SLEEK_ECAPACITY : ((words[i] & 1) ? SLEEK_ECORRUPT : SLEEK_OK);
  return SLEEK_OK;
}

int sleek_gpu_batch_graph_create(int decompress, const sleek_gpu_batch_item* d_items, int n, long long tiles,
                                 void* d_workspace, long long workspace_bytes, int grid, void** graph)
{
  if (!graph) return SLEEK_EINVAL;
  *graph = nullptr;
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  int rc = SLEEK_OK;
  e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    rc = hip_fail(e, "hipStreamBeginCapture");
  } else {
    rc = batch_async(decompress != 0, d_items, n, tiles, d_workspace, workspace_bytes, grid, s);
    e = hipStreamEndCapture(s, &g);
    if (rc == SLEEK_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
    if (rc == SLEEK_OK) {
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
    }
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(s);
  if (rc == SLEEK_OK) *graph = (void*)x;
  return rc;
}

int sleek_gpu_graph_launch(void* graph, void* stream)
{
  if (!graph) return SLEEK_EINVAL;
  const hipError_t e = hipGraphLaunch((hipGraphExec_t)graph, (hipStream_t)stream);
  return e == hipSuccess ? SLEEK_OK : hip_fail(e, "hipGraphLaunch");
}

int sleek_gpu_graph_destroy(void* graph)
{
  if (!graph) return SLEEK_OK;
  const hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph);
  return e == hipSuccess ? SLEEK_OK : hip_fail(e, "hipGraphExecDestroy");
}

}  // extern "C"
