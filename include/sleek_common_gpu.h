/*
 * SLEEK for MI355X: common declarations of the drop-in GPU headers.
 *
 * Keeps the names a caller of the reference's include/sleek_common_gpu.h
 * uses (SLEEK_CS, byte, sleek_upper_bound_size, sleek_system_checks,
 * sleek_thread_blocks -- reference lines :53-119) on top of the C ABI in
 * sleek_gpu.h.  Plain C++: compiles with g++ and links against
 * libsleek_amd.so; no HIP headers are needed by the caller.
 */
#pragma once

#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "sleek_gpu.h"

#ifndef SLEEK_NO_BYTE_ALIAS
using byte = unsigned char;   // the reference's alias (sleek_common_gpu.h:53)
#endif

static const int SLEEK_CS = SLEEK_CHUNK_BYTES;   // chunk size in bytes (format constant)
static const int SLEEK_TPB = 256;                // threads per workgroup of the gfx950 kernels

// Stream bytes an input of insize bytes may need (sleek_common_gpu.h:79-83).
static inline long long sleek_upper_bound_size(const long long insize)
{
  return sleek_gpu_upper_bound_size(insize);
}

// 0 if the platform matches the format's assumptions (sleek_common_gpu.h:86-105).
static inline int sleek_system_checks()
{
  if (sleek_gpu_system_checks() != SLEEK_OK) {
    fprintf(stderr, "ERROR: SLEEK needs a little-endian platform with 8-byte long long, 4-byte int, 2-byte short\n\n");
    return -1;
  }
  return 0;
}

// Persistent grid size for the current device (sleek_common_gpu.h:108-119).
static inline int sleek_thread_blocks()
{
  const int b = sleek_gpu_thread_blocks();
  if (b <= 0) {
    fprintf(stderr, "ERROR: no usable GPU detected (%s)\n\n", sleek_gpu_last_error());
    exit(-1);
  }
  return b;
}

// Status of the most recent drop-in call on this thread (SLEEK_OK, or the
// SLEEK_E* code the C ABI returned).  The reference functions return void.
static inline int& sleek_last_status()
{
  static thread_local int st = SLEEK_OK;
  return st;
}

static inline void sleek_report(const int status, const char* what)
{
  sleek_last_status() = status;
  if (status != SLEEK_OK) {
    fprintf(stderr, "ERROR: %s: %s", what, sleek_gpu_strerror(status));
    if (status == SLEEK_EHIP) fprintf(stderr, " (%s)", sleek_gpu_last_error());
    fprintf(stderr, "\n\n");
  }
}

// Replaces the reference's CUDA-named CheckCuda(line) (sleek_common_gpu.h:68-76):
// exits if the most recent drop-in call on this thread failed.
static inline void sleek_check(const int line)
{
  if (sleek_last_status() != SLEEK_OK) {
    fprintf(stderr, "SLEEK error %d on line %d: %s\n\n", sleek_last_status(), line,
            sleek_gpu_strerror(sleek_last_status()));
    exit(-1);
  }
}
