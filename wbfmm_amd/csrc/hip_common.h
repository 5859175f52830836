// hip_common.h -- shared helpers of the gfx950 kernels (device side) and
// of the launch layer (stream, error capture, per-kernel event timing).

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wbfmm_hip.h"

namespace wbfmm {

hipStream_t stream();
// stream-ordered device allocation on stream(): the memory comes from the
// device's default pool and returns there when the stream reaches the free,
// so neither call waits for the device
int dmalloc(void **p, size_t bytes);
void dfree(void *p);
// record an error message; returns WBFMM_HIP_ERR_RUNTIME
int fail(const char *what, hipError_t e);
// record an error message of the library's own; returns code
int fail_msg(const char *msg, int code);
// begin/end a timed launch of kernel `name` (no-ops unless timing is on)
void timing_begin(const char *name);
void timing_end(const char *name);

#define WBFMM_CHECK(call)                                  \
  do {                                                     \
    hipError_t _e = (call);                                \
    if (_e != hipSuccess) return ::wbfmm::fail(#call, _e); \
  } while (0)

#define WBFMM_LAUNCH(name, kernel, grid, block, shmem, ...)                  \
  do {                                                                        \
    ::wbfmm::timing_begin(name);                                              \
    hipLaunchKernelGGL(kernel, grid, block, shmem, ::wbfmm::stream(),        \
                       __VA_ARGS__);                                          \
    hipError_t _e = hipGetLastError();                                        \
    ::wbfmm::timing_end(name);                                                \
    if (_e != hipSuccess) return ::wbfmm::fail(name, _e);                    \
  } while (0)

// ---- device helpers ------------------------------------------------------

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

__device__ __forceinline__ uint32_t compact3(uint64_t x) {
  x &= 0x1249249249249249ULL;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
  x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
  x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
  x = (x ^ (x >> 32)) & 0x1fffff;
  return (uint32_t)x;
}

// Box indices at levels up to 10 fit in 30 bits.  The index of box
// (i + di, j + dj, k + dk) follows from that of box (i, j, k) by one
// dilated addition per axis: the offset is spread onto its axis's bits
// (dilate_offset; a negative offset as its dilated two's complement) and
// added to the index with the other axes' bits set, so that carries pass
// over them.  Exact wherever the result lies inside the level.
constexpr uint32_t kMortonX = 0x49249249u, kMortonY = 0x92492492u, kMortonZ = 0x24924924u;
constexpr int kMortonAddMaxLevel = 10;

// offset d (|d| < 8: the list-4 offsets are within 3) on the axis of bit
// mask m, shift s (0, 1, 2)
__device__ __forceinline__ uint32_t dilate_offset(int d, int s, uint32_t m) {
  const uint32_t u = (uint32_t)(d < 0 ? -d : d);
  const uint32_t a = ((u & 1u) | ((u & 2u) << 2) | ((u & 4u) << 4)) << s;
  return d < 0 ? (0u - a) & m : a;
}

__device__ __forceinline__ uint32_t morton_add(uint32_t t, uint32_t ox, uint32_t oy,
                                               uint32_t oz) {
  return (((t | ~kMortonX) + ox) & kMortonX) | (((t | ~kMortonY) + oy) & kMortonY) |
         (((t | ~kMortonZ) + oz) & kMortonZ);
}

// Morton code of a point on the 2^20 grid, same arithmetic as the
// reference's wbfmm_point_index_3d (src/boxes.c:324-339)
__device__ __forceinline__ uint64_t point_code(double x, double y, double z,
                                               double x0, double y0,
                                               double z0, double D) {
  uint32_t i = (uint32_t)((x - x0) / D * 1048576.0);
  uint32_t j = (uint32_t)((y - y0) / D * 1048576.0);
  uint32_t k = (uint32_t)((z - z0) / D * 1048576.0);
  return spread3(i) | (spread3(j) << 1) | (spread3(k) << 2);
}

// element offset of coefficient c of box b in a level array.  On the
// device the levels are stored box-major (box b at b*C, C coefficients,
// reals for Laplace) so that one box's expansion is one contiguous run for
// the gathers of M2L; the reference's 8-sibling interleave
// (src/boxes.c:552-579, src/laplace.c:823-849) is restored when
// coefficients are copied to the host (wbfmm_tree_coefficients_download).
__device__ __forceinline__ size_t coef_at(uint64_t b, int c, int C) {
  return (size_t)b * (size_t)C + (size_t)c;
}

__device__ __forceinline__ int isqrt_deg(int r) {
  int n = (int)sqrt((double)r);
  while (n * n > r) n--;
  while ((n + 1) * (n + 1) <= r) n++;
  return n;
}

constexpr double kPi = 3.14159265358979323846;
constexpr int kWave = 64;

}  // namespace wbfmm
