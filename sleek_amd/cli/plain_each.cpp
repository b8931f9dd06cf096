// Per-array plain launches of a batch, from a C++ loop over the stream-ordered
// C ABI: the comparison the batched launch is measured against.  One
// sleek_gpu_compress_async / sleek_gpu_decompress_async call per array on one
// stream, each with its own workspace (so every call's status can be read
// back afterwards), no host synchronisation inside the loop.  Built as a small
// shared library for the benchmark (ctypes), so the loop's per-call cost is
// the C ABI's, not an interpreter's.
#include "../../include/sleek_gpu.h"

extern "C" {

// items: host copies of the batch items (src, dst, sizes, d_size, type and
// error-bound fields as for sleek_gpu_batch_plan; the planned fields are
// ignored).  ws[i] / ws_bytes[i]: array i's workspace.  Returns SLEEK_OK, or
// the first failing call's status (the calls before it are enqueued).
int sleek_plain_each(int decompress, const sleek_gpu_batch_item* items, int n, void* const* ws,
                     const long long* ws_bytes, void* stream)
{
  for (int i = 0; i < n; i++) {
    const sleek_gpu_batch_item& it = items[i];
    const int rc = decompress
        ? sleek_gpu_decompress_async(it.elem_bytes, it.lossy, it.thr_e, it.offs, (const unsigned char*)it.src,
                                     it.src_bytes, it.dst, it.dst_bytes, it.d_size, ws[i], ws_bytes[i], 0, stream)
        : sleek_gpu_compress_async(it.elem_bytes, it.lossy, it.eb_e, it.thr_e, it.offs, it.src, it.src_bytes,
                                   (unsigned char*)it.dst, it.d_size, ws[i], ws_bytes[i], 0, stream);
    if (rc != SLEEK_OK) return rc;
  }
  return SLEEK_OK;
}

}  // extern "C"
