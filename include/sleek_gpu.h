/*
 * SLEEK for MI355X (gfx950): C ABI of the GPU codec.
 *
 * Plain C, no HIP or torch types: device pointers are void* / byte pointers,
 * streams are passed as void* (a hipStream_t, or NULL for the null stream).
 * Every function returns a status (SLEEK_OK or a negative SLEEK_E* code)
 * unless noted.  The C++ drop-in headers sleek_lossless_gpu.h and
 * sleek_lossy_gpu.h wrap these with the reference's overloads.
 *
 * Stream format: identical, byte for byte, to the reference codec's
 * (int64 length header, uint16 per-chunk size table, byte-packed chunk
 * payloads; 16 KiB chunks).
 */
#ifndef SLEEK_GPU_H
#define SLEEK_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define SLEEK_OK 0
#define SLEEK_EINVAL (-1)      /* bad argument */
#define SLEEK_EHIP (-2)        /* HIP runtime error (see sleek_gpu_last_error) */
#define SLEEK_ECORRUPT (-3)    /* malformed stream detected while decoding */
#define SLEEK_EEB (-4)         /* error bound too large */
#define SLEEK_ECAPACITY (-5)   /* decoded size exceeds the output capacity */
#define SLEEK_EWORKSPACE (-6)  /* workspace too small */
#define SLEEK_EALIGN (-7)      /* pointer not aligned as required */
#define SLEEK_ETOOLARGE (-8)   /* input above sleek_gpu_compress_noalloc's maximum */

#define SLEEK_CHUNK_BYTES 16384

/* ---- helpers ---- */

/* Bytes a stream of insize input bytes may need; same value as the
 * reference's sleek_upper_bound_size (sleek_common_gpu.h:79-83). */
long long sleek_gpu_upper_bound_size(long long insize);

/* Platform checks of sleek_system_checks (sleek_common_gpu.h:86-105): 0 if ok. */
int sleek_gpu_system_checks(void);

/* Suggested persistent grid size for the current device (the role of
 * sleek_thread_blocks, sleek_common_gpu.h:108-119).  Output bytes never
 * depend on the grid size.  Returns a positive count, or a negative status. */
int sleek_gpu_thread_blocks(void);

/* Error-bound parameters of an absolute bound (sleek_lossy_gpu.h:56-71).
 * Returns SLEEK_EEB where the reference exits the process. */
int sleek_gpu_error_bound_f32(float eb, int* eb_e, int* thr_e, unsigned int* offs_e);
int sleek_gpu_error_bound_f64(double eb, int* eb_e, int* thr_e, unsigned long long* offs_e);

/* Text of a status code, and the last HIP error message of this thread. */
const char* sleek_gpu_strerror(int status);
const char* sleek_gpu_last_error(void);

/* ---- synchronous codec calls (library-owned scratch, null stream) ----
 * Counterparts of sleek_compress_gpu / sleek_decompress_gpu:
 * lossless sleek_lossless_gpu.h:47-100, lossy sleek_lossy_gpu.h:74-127.
 * d_input / d_decoded must be 16-byte aligned, d_encoded 8-byte aligned;
 * d_encoded holds at least sleek_gpu_upper_bound_size(insize) bytes.
 * *d_encsize / *d_decsize are device int64 results.  blocks <= 0 picks the
 * grid; any positive value is a persistent grid size (bytes do not change).
 * On return the results are complete; a malformed stream gives
 * SLEEK_ECORRUPT (the reference traps the device there). */
int sleek_gpu_compress_f32(int blocks, const float* d_input, long long num_elements,
                           unsigned char* d_encoded, long long* d_encsize);
int sleek_gpu_decompress_f32(int blocks, const unsigned char* d_encoded, float* d_decoded,
                             long long* d_decsize);
int sleek_gpu_compress_f64(int blocks, const double* d_input, long long num_elements,
                           unsigned char* d_encoded, long long* d_encsize);
int sleek_gpu_decompress_f64(int blocks, const unsigned char* d_encoded, double* d_decoded,
                             long long* d_decsize);
int sleek_gpu_compress_lossy_f32(int blocks, int eb_e, int thr_e, unsigned int offs_e,
                                 const float* d_input, long long num_elements,
                                 unsigned char* d_encoded, long long* d_encsize);
int sleek_gpu_decompress_lossy_f32(int blocks, int eb_e, int thr_e, unsigned int offs_e,
                                   const unsigned char* d_encoded, float* d_decoded,
                                   long long* d_decsize);
int sleek_gpu_compress_lossy_f64(int blocks, int eb_e, int thr_e, unsigned long long offs_e,
                                 const double* d_input, long long num_elements,
                                 unsigned char* d_encoded, long long* d_encsize);
int sleek_gpu_decompress_lossy_f64(int blocks, int eb_e, int thr_e, unsigned long long offs_e,
                                   const unsigned char* d_encoded, double* d_decoded,
                                   long long* d_decsize);

/* ---- stream-ordered calls with a caller-owned workspace ----
 * No reference counterpart: the reference allocates and frees its scratch
 * inside every call (sleek_lossless_gpu.h:52-59), which serialises the
 * device.  These enqueue on `stream` and return immediately; the workspace
 * must not be shared by calls that may overlap.
 * elem_bytes: 4 (float32) or 8 (float64).  lossy: 0 or 1.
 * enclen / capacity: stream bytes available and decoded-buffer bytes, or -1
 * if unknown (capacity -1 needs a workspace sized for the stream header's
 * length).  Device-detected stream errors are read back, after the stream
 * has drained, by sleek_gpu_workspace_status. */
long long sleek_gpu_workspace_size(long long max_insize);
int sleek_gpu_compress_async(int elem_bytes, int lossy, int eb_e, int thr_e, unsigned long long offs_e,
                             const void* d_input, long long insize, unsigned char* d_encoded,
                             long long* d_encsize, void* d_workspace, long long workspace_bytes,
                             int blocks, void* stream);
int sleek_gpu_decompress_async(int elem_bytes, int lossy, int thr_e, unsigned long long offs_e,
                               const unsigned char* d_encoded, long long enclen, void* d_decoded,
                               long long capacity, long long* d_decsize, void* d_workspace,
                               long long workspace_bytes, int blocks, void* stream);
/* Launch-shape knobs; no reference counterpart.  encode_variant: 0 =
 * default (currently 6), 1 = pipelined 256-thread encoder at a
 * 4-waves-per-SIMD register budget, 2 = the same at the compiler's register
 * choice, 4 = one wave per chunk pipelined over the wave's chunks, 5 = two
 * waves per chunk pipelined the same way, 6 = 5 with every chunk handed out
 * in arrival order by ticket counters, 7 = 4 with chunk tickets (3, an
 * unpipelined one-wave encoder, is retired).  The batched calls use two
 * waves per tile with tile tickets, except under variant 4 (one wave per
 * tile, tiles in launch order).  encode_grid / decode_grid: grid sizes
 * (0 = automatic).  They
 * change speed only, never the bytes.  Applies to calls made after it
 * returns; SLEEK_EINVAL for an unknown variant or a negative size. */
int sleek_gpu_set_tuning(int encode_variant, int encode_grid, int decode_grid);
/* Look-back help (no reference counterpart; the reference's look-back only
 * waits, sleek_common_gpu.h:220-253): a look-back that has waited about this
 * many microseconds for a predecessor's size (its own sleep time: time its
 * queue spends suspended does not count) computes it from the input itself
 * (0 = at once, for tests; default 100; at most 10^7).  The ticket encoders
 * (variants 6 and 7, and the batched encoder except under variant 4) hand
 * chunks to running workgroups only and should never need it.  Speed only,
 * never the bytes; applies to later calls. */
int sleek_gpu_set_help_after(int microseconds);
/* Help events (missing look-back words a waiting wave sized itself) on the
 * current device since the library was loaded; synchronises the device.
 * Returns the count, or a negative status. */
long long sleek_gpu_help_events(void);
/* Kernel timing (no reference counterpart; for benchmarks): the next
 * compress or decompress call (single-array, sync or async) that this host
 * thread makes records the caller's HIP events (hipEvent_t, either may be
 * null) on the call's stream: `begin` right before its coding kernel (the
 * encoder, or the decoder after its offset scan) and `end` right after it,
 * so hipEventElapsedTime(begin, end) is that kernel's own duration, without
 * the workspace clear or the offset scan.  The setting is used once and
 * then cleared.  Returns SLEEK_OK. */
int sleek_gpu_set_kernel_events(void* begin, void* end);
/* Look-back geometry of the encoders (no reference counterpart; the
 * reference watches a fixed window of 32 chunks, sleek_common_gpu.h:220-253):
 * group_chunks = chunks per look-back group, 8, 16, 32 or 64 (0 = default,
 * 32); window_groups = groups watched by the first group-level poll, 1..64
 * (0 = default: 64 for encode_variant 6, 8 otherwise).  Speed only, never
 * the bytes; applies to later calls. */
int sleek_gpu_set_lookback(int group_chunks, int window_groups);
/* Decoder shape (no reference counterpart): 0 = an offset-scan launch, then
 * one workgroup per chunk (the default); 1 = one persistent launch that
 * derives each chunk's stream offset from the size table itself.  Same
 * bytes; applies to later calls. */
int sleek_gpu_set_decoder(int variant);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  sleek_gpu_compress_async with the library's static device
 * scratch as the workspace: it calls no hipMalloc, hipFree or allocating
 * routine.  Inputs up to 4 GiB (4,294,967,296 bytes); larger ones give
 * SLEEK_ETOOLARGE.  The scratch (2,625,664 bytes) is held by each device that
 * the library runs on until the process ends, and every call on a device
 * shares it: calls that may overlap on one device (other streams, other
 * host threads) race and must not be issued.  The call clears what must
 * start at zero with a kernel, so it can be captured.  Same bytes as
 * sleek_gpu_compress_async; the encoder reports no stream errors. */
int sleek_gpu_compress_noalloc(int elem_bytes, int lossy, int eb_e, int thr_e, unsigned long long offs_e,
                               const void* d_input, long long insize, unsigned char* d_encoded,
                               long long* d_encsize, int blocks, void* stream);
/* Synchronises `stream`, then returns SLEEK_OK, SLEEK_ECORRUPT or
 * SLEEK_ECAPACITY from the workspace's status word. */
int sleek_gpu_workspace_status(const void* d_workspace, void* stream);

/* ---- batched calls: many arrays in one launch per direction ----
 * No reference counterpart (the reference codes one array per call).  Every
 * array's stream / decoded bytes are exactly those of the single-array
 * calls.  Arrays may mix float32 and float64, lossless and lossy.
 *
 * The caller fills the first block of each item; sleek_gpu_batch_plan
 * fills the second and returns the tile count and workspace size.  The
 * caller then copies the planned items to device memory (d_items) and
 * launches with a workspace of at least that size, which must not be shared
 * by launches that may overlap.  A tile is one 16 KiB chunk of one array;
 * workgroups take tiles in launch order through a device tile map. */
typedef struct sleek_gpu_batch_item {
  const void* src;          /* compress: input array (16-byte aligned); decompress: stream (8-byte aligned) */
  void* dst;                /* compress: stream buffer of sleek_gpu_upper_bound_size(src_bytes) bytes
                               (8-byte aligned); decompress: output array (16-byte aligned) */
  long long src_bytes;      /* compress: input bytes (a multiple of elem_bytes);
                               decompress: stream bytes available, or -1 if unknown */
  long long dst_bytes;      /* decompress: capacity in bytes (required); compress: ignored */
  long long* d_size;        /* device int64 result: stream length / decoded length */
  unsigned long long offs;  /* lossy parameters, from sleek_gpu_error_bound_f32/_f64 */
  int eb_e, thr_e;
  int elem_bytes;           /* 4 (float32) or 8 (float64) */
  int lossy;                /* 0 or 1 */
  /* filled by sleek_gpu_batch_plan */
  long long tile0;          /* first tile of this array */
  long long tiles;          /* its tiles (16 KiB chunks; 0 for an empty array) */
  long long ws_word;        /* first 8-byte word of its workspace block */
  long long ws_words;       /* words in that block */
} sleek_gpu_batch_item;     /* 96 bytes */

/* Plans n items for compression (decompress = 0) or decompression (1).
 * SLEEK_EINVAL names nothing: a bad item (size, type, missing capacity),
 * SLEEK_EALIGN a misaligned pointer. */
int sleek_gpu_batch_plan(sleek_gpu_batch_item* items, int n, int decompress, long long* tiles,
                         long long* workspace_bytes);
/* One launch sequence per direction on `stream`: a preparation kernel
 * (clears each array's workspace block with stores, writes the tile map)
 * and the coder; the decoder adds a per-array offset scan.  grid <= 0
 * picks a persistent grid.  No allocation, no memset: capturable. */
int sleek_gpu_batch_compress_async(const sleek_gpu_batch_item* d_items, int n, long long tiles, void* d_workspace,
                                   long long workspace_bytes, int grid, void* stream);
int sleek_gpu_batch_decompress_async(const sleek_gpu_batch_item* d_items, int n, long long tiles,
                                     void* d_workspace, long long workspace_bytes, int grid, void* stream);
/* Per-array status after a batched launch: synchronises `stream`, then
 * writes SLEEK_OK / SLEEK_ECORRUPT / SLEEK_ECAPACITY into status[0..n). */
int sleek_gpu_batch_status(const sleek_gpu_batch_item* d_items, int n, void* d_workspace, long long workspace_bytes,
                           int* status, void* stream);
/* The batched launch sequence captured once into an executable HIP graph
 * (*graph), replayed by sleek_gpu_graph_launch on any stream; the items,
 * buffers and workspace it names must stay allocated until
 * sleek_gpu_graph_destroy. */
int sleek_gpu_batch_graph_create(int decompress, const sleek_gpu_batch_item* d_items, int n, long long tiles,
                                 void* d_workspace, long long workspace_bytes, int grid, void** graph);
int sleek_gpu_graph_launch(void* graph, void* stream);
int sleek_gpu_graph_destroy(void* graph);

#ifdef __cplusplus
}
#endif

#endif
