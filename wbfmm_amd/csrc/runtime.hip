// runtime.hip -- stream, memory and timing services of the thin C ABI,
// plus the octree build kernels (Morton codes, stable sort, box ranges).
//
// Tree build restates src/boxes.c:118-277: the reference sorts an index
// array by Morton code with a stable qsort and then segments it level by
// level; here the codes are computed by one kernel, sorted with rocPRIM's
// stable LSD radix sort, and every level's [start,end) ranges are found by
// one kernel that marks segment boundaries in the sorted codes.

#include <cstdio>
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hip_common.h"

namespace wbfmm {

// The library stream, created by wbfmm_hip_init.  A thread may select
// another stream of its own (wbfmm_hip_set_stream); every launch, copy and
// stream-ordered allocation of that thread then goes there, which is how
// the serving engine runs independent solves side by side.
static hipStream_t g_stream = nullptr;
static thread_local hipStream_t t_stream = nullptr;
static std::mutex g_mu;
static thread_local char g_err[512] = "";
static bool g_timing = false;
static std::mutex g_timing_mu;
struct Ev {
  std::string name;
  hipEvent_t a, b;
};
static std::vector<Ev> g_events;
static std::map<std::string, std::pair<int, double>> g_totals;

hipStream_t stream() {
  if (t_stream) return t_stream;
  if (!g_stream) wbfmm_hip_init();
  return g_stream;
}

int fail(const char *what, hipError_t e) {
  snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
  fprintf(stderr, "wbfmm (HIP): %s\n", g_err);
  return WBFMM_HIP_ERR_RUNTIME;
}

int fail_msg(const char *msg, int code) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  fprintf(stderr, "wbfmm (HIP): %s\n", g_err);
  return code;
}

// Timing covers launches made on the library stream only; launches on a
// thread's own stream are not timed.
void timing_begin(const char *name) {
  if (!g_timing || t_stream) return;
  std::lock_guard<std::mutex> lk(g_timing_mu);
  Ev ev;
  ev.name = name;
  hipEventCreate(&ev.a);
  hipEventCreate(&ev.b);
  hipEventRecord(ev.a, g_stream);
  g_events.push_back(ev);
}

void timing_end(const char *name) {
  if (!g_timing || t_stream) return;
  (void)name;
  std::lock_guard<std::mutex> lk(g_timing_mu);
  hipEventRecord(g_events.back().b, g_stream);
}

static void timing_collect() {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  if (g_events.empty()) return;
  hipStreamSynchronize(g_stream);
  for (auto &ev : g_events) {
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ev.a, ev.b);
    auto &t = g_totals[ev.name];
    t.first += 1;
    t.second += ms;
    hipEventDestroy(ev.a);
    hipEventDestroy(ev.b);
  }
  g_events.clear();
}

// ---------------------------------------------------------------------
// tree kernels
// ---------------------------------------------------------------------

__global__ void k_morton(const char *pts, size_t pstr, unsigned n, double x0,
                         double y0, double z0, double D, double *xyz,
                         uint64_t *codes, uint32_t *idx, unsigned *nbad) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *x = (const double *)(pts + (size_t)i * pstr);
  double a = x[0], b = x[1], c = x[2];
  if (a <= x0 || a >= x0 + D || b <= y0 || b >= y0 + D || c <= z0 ||
      c >= z0 + D)
    atomicAdd(nbad, 1u);
  codes[i] = point_code(a, b, c, x0, y0, z0, D);
  idx[i] = i;
  xyz[4 * (size_t)i + 0] = a;
  xyz[4 * (size_t)i + 1] = b;
  xyz[4 * (size_t)i + 2] = c;
  xyz[4 * (size_t)i + 3] = 0.0;
}

__global__ void k_stage4(const char *pts, size_t pstr, unsigned n,
                         double *out) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *x = (const double *)(pts + (size_t)i * pstr);
  out[4 * (size_t)i + 0] = x[0];
  out[4 * (size_t)i + 1] = x[1];
  out[4 * (size_t)i + 2] = x[2];
  out[4 * (size_t)i + 3] = 0.0;
}

__global__ void k_gather4(const double *in, const uint32_t *idx, unsigned n,
                          double *out) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *s = in + 4 * (size_t)idx[i];
  double *d = out + 4 * (size_t)i;
  d[0] = s[0];
  d[1] = s[1];
  d[2] = s[2];
  d[3] = s[3];
}

// out record i = in record idx[i], records of rsize bytes moved as
// 4-byte words when rsize allows, else bytes
__global__ void k_gather_records4(const uint32_t *in, size_t rw, const uint32_t *idx,
                                  unsigned n, uint32_t *out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * rw) return;
  const size_t i = e / rw, w = e % rw;
  out[e] = in[(size_t)idx[i] * rw + w];
}

__global__ void k_gather_records1(const uint8_t *in, size_t rsize, const uint32_t *idx,
                                  unsigned n, uint8_t *out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * rsize) return;
  const size_t i = e / rsize, w = e % rsize;
  out[e] = in[(size_t)idx[i] * rsize + w];
}

__global__ void k_gather_strided(const double *src, int sstr, int ncomp,
                                 const uint32_t *idx, unsigned n,
                                 double *out) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncomp; c++)
    out[(size_t)i * ncomp + c] = src[(size_t)idx[i] * sstr + c];
}

__global__ void k_box_ranges(const uint64_t *codes, unsigned n, int shift,
                             uint32_t *start, uint32_t *end) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t b = codes[i] >> shift;
  if (i == 0 || (codes[i - 1] >> shift) != b) start[b] = i;
  if (i == n - 1 || (codes[i + 1] >> shift) != b) end[b] = i + 1;
}

__global__ void k_box_flags(const uint32_t *start, const uint32_t *end,
                            const uint32_t *flag, unsigned nb,
                            uint32_t *out) {
  unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  out[b] = flag ? (flag[b] != 0) : (end[b] > start[b]);
}

__global__ void k_box_scatter(const uint32_t *flags, const uint32_t *pos,
                              unsigned nb, uint32_t *list) {
  unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (flags[b]) list[pos[b]] = b;
}

__global__ void k_mark_boxes(const uint64_t *codes, unsigned n, int shift,
                             uint32_t *flag) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[codes[i] >> shift] = 1u;
}

// per-block minimum and maximum of the 3 coordinates: out[6 b + 0..2] min,
// out[6 b + 3..5] max
__global__ void k_bounds(const char *pts, size_t pstr, unsigned n,
                         double *out) {
  __shared__ double red[6][256];
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const double *x = (const double *)(pts + (size_t)i * pstr);
    for (int c = 0; c < 3; c++) {
      lo[c] = fmin(lo[c], x[c]);
      hi[c] = fmax(hi[c], x[c]);
    }
  }
  for (int c = 0; c < 3; c++) {
    red[c][threadIdx.x] = lo[c];
    red[3 + c][threadIdx.x] = hi[c];
  }
  __syncthreads();
  for (unsigned h = blockDim.x / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h)
      for (int c = 0; c < 3; c++) {
        red[c][threadIdx.x] = fmin(red[c][threadIdx.x], red[c][threadIdx.x + h]);
        red[3 + c][threadIdx.x] =
            fmax(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + h]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_accumulate(double *dst, int dstr, const double *src,
                             int sstr, int ncomp, unsigned n) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncomp; c++)
    dst[(size_t)i * dstr + c] += src[(size_t)i * sstr + c];
}

static inline unsigned nblk(size_t n, unsigned b) {
  return (unsigned)((n + b - 1) / b);
}


// allocations of this thread still to be refused (failure-handling tests)
static thread_local int t_fail_allocs = 0;

int dmalloc(void **p, size_t bytes) {
  *p = nullptr;
  if (t_fail_allocs > 0) {
    t_fail_allocs--;
    return fail_msg("device allocation refused (forced by wbfmm_hip_debug_fail_allocs)",
                    WBFMM_HIP_ERR_NOMEM);
  }
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMallocAsync(p, bytes, stream());
  if (e != hipSuccess) {
    *p = nullptr;
    fail("hipMallocAsync", e);
    return WBFMM_HIP_ERR_NOMEM;
  }
  return 0;
}

void dfree(void *p) {
  if (p) (void)hipFreeAsync(p, stream());
}

}  // namespace wbfmm

using namespace wbfmm;

extern "C" {

int wbfmm_hip_init(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_stream) return 0;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) {
    snprintf(g_err, sizeof(g_err), "no HIP device available (%s)",
             hipGetErrorString(e));
    fprintf(stderr, "wbfmm (HIP): %s\n", g_err);
    return WBFMM_HIP_ERR_RUNTIME;
  }
  // freed blocks stay in the default pool instead of going back to the
  // driver, so the per-solve allocations are served from the pool
  int dev = 0;
  WBFMM_CHECK(hipGetDevice(&dev));
  hipMemPool_t pool;
  WBFMM_CHECK(hipDeviceGetDefaultMemPool(&pool, dev));
  uint64_t keep = UINT64_MAX;
  WBFMM_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  // default priority, non-blocking with respect to the null stream
  WBFMM_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  return 0;
}

int wbfmm_hip_mem_info(size_t *free_bytes, size_t *total_bytes) {
  if (wbfmm_hip_init()) return WBFMM_HIP_ERR_RUNTIME;
  WBFMM_CHECK(hipMemGetInfo(free_bytes, total_bytes));
  return 0;
}

int wbfmm_hip_stream_create(void **s) {
  if (wbfmm_hip_init()) return WBFMM_HIP_ERR_RUNTIME;
  hipStream_t h = nullptr;
  WBFMM_CHECK(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = (void *)h;
  return 0;
}

int wbfmm_hip_stream_destroy(void *s) {
  if (!s) return 0;
  WBFMM_CHECK(hipStreamSynchronize((hipStream_t)s));
  WBFMM_CHECK(hipStreamDestroy((hipStream_t)s));
  return 0;
}

static thread_local int t_variant[WBFMM_VARIANT_COUNT] = {-1, -1};

int wbfmm_hip_set_variant(int which, int value) {
  // M2L: -1..2; field: -1..1
  const int vmax = which == WBFMM_VARIANT_FIELD ? 1 : 2;
  if (which < 0 || which >= WBFMM_VARIANT_COUNT || value < -1 || value > vmax)
    return fail_msg("unknown kernel variant", WBFMM_HIP_ERR_ARG);
  t_variant[which] = value;
  return 0;
}

int wbfmm_hip_get_variant(int which) {
  return which >= 0 && which < WBFMM_VARIANT_COUNT ? t_variant[which] : -1;
}

int wbfmm_hip_set_stream(void *s, void **prev) {
  if (prev) *prev = (void *)t_stream;
  t_stream = (hipStream_t)s;
  return 0;
}

void *wbfmm_hip_get_stream(void) { return (void *)stream(); }

const char *wbfmm_hip_last_error(void) { return g_err; }

int wbfmm_hip_debug_fail_allocs(int n) {
  t_fail_allocs = n > 0 ? n : 0;
  return 0;
}

int wbfmm_hip_malloc(void **p, size_t bytes) {
  if (wbfmm_hip_init()) return WBFMM_HIP_ERR_RUNTIME;
  return dmalloc(p, bytes);
}

int wbfmm_hip_free(void *p) {
  if (!p) return 0;
  WBFMM_CHECK(hipFreeAsync(p, stream()));
  return 0;
}

int wbfmm_hip_is_device(const void *p) {
  if (!p) return 0;
  if (wbfmm_hip_init()) return 0;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    hipGetLastError();  // clear the sticky error from an unknown pointer
    return 0;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

int wbfmm_hip_memcpy(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return 0;
  WBFMM_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, stream()));
  if (!wbfmm_hip_is_device(dst) || !wbfmm_hip_is_device(src))
    WBFMM_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int wbfmm_hip_memset(void *p, int v, size_t bytes) {
  if (bytes == 0) return 0;
  WBFMM_CHECK(hipMemsetAsync(p, v, bytes, stream()));
  return 0;
}

int wbfmm_hip_sync(void) {
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int wbfmm_hip_timing_enable(int on) {
  timing_collect();
  g_timing = (on != 0);
  return 0;
}

int wbfmm_hip_timing_reset(void) {
  timing_collect();
  g_totals.clear();
  return 0;
}

int wbfmm_hip_timing_query(const char *name, int *launches, double *ms) {
  timing_collect();
  auto it = g_totals.find(name);
  if (it == g_totals.end()) {
    *launches = 0;
    *ms = 0.0;
    return 0;
  }
  *launches = it->second.first;
  *ms = it->second.second;
  return 0;
}

int wbfmm_hip_morton(const void *pts, size_t pstr, unsigned n,
                     const double *x0, double D, double *xyz,
                     uint64_t *codes, uint32_t *idx, unsigned *nbad) {
  if (n == 0) {
    *nbad = 0;
    return 0;
  }
  const char *dpts = (const char *)pts;
  void *staged = nullptr;
  if (!wbfmm_hip_is_device(pts)) {
    // copy the strided host block as it lies
    size_t bytes = (size_t)(n - 1) * pstr + 3 * sizeof(double);
    if (wbfmm_hip_malloc(&staged, bytes)) return WBFMM_HIP_ERR_NOMEM;
    WBFMM_CHECK(hipMemcpyAsync(staged, pts, bytes, hipMemcpyHostToDevice,
                               stream()));
    dpts = (const char *)staged;
  }
  unsigned *dbad = nullptr;
  if (dmalloc((void **)&dbad, sizeof(unsigned))) return WBFMM_HIP_ERR_NOMEM;
  WBFMM_CHECK(hipMemsetAsync(dbad, 0, sizeof(unsigned), stream()));
  WBFMM_LAUNCH("tree_morton", k_morton, dim3(nblk(n, 256)), dim3(256), 0,
               dpts, pstr, n, x0[0], x0[1], x0[2], D, xyz, codes, idx, dbad);
  WBFMM_CHECK(hipMemcpyAsync(nbad, dbad, sizeof(unsigned),
                             hipMemcpyDeviceToHost, stream()));
  dfree(dbad);
  dfree(staged);
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  return 0;
}

int wbfmm_hip_stage4(const void *pts, size_t pstr, unsigned n, double *out) {
  if (n == 0) return 0;
  const char *dpts = (const char *)pts;
  void *staged = nullptr;
  if (!wbfmm_hip_is_device(pts)) {
    size_t bytes = (size_t)(n - 1) * pstr + 3 * sizeof(double);
    if (wbfmm_hip_malloc(&staged, bytes)) return WBFMM_HIP_ERR_NOMEM;
    WBFMM_CHECK(hipMemcpyAsync(staged, pts, bytes, hipMemcpyHostToDevice,
                               stream()));
    dpts = (const char *)staged;
  }
  WBFMM_LAUNCH("stage4", k_stage4, dim3(nblk(n, 256)), dim3(256), 0, dpts,
               pstr, n, out);
  if (staged) {
    dfree(staged);
    // the caller's host block may be reused once this returns
    WBFMM_CHECK(hipStreamSynchronize(stream()));
  }
  return 0;
}

int wbfmm_hip_sort(uint64_t *codes, uint32_t *idx, unsigned n) {
  if (n < 2) return 0;
  uint64_t *k2 = nullptr;
  uint32_t *v2 = nullptr;
  void *tmp = nullptr;
  if (dmalloc((void **)&k2, sizeof(uint64_t) * n) ||
      dmalloc((void **)&v2, sizeof(uint32_t) * n))
    return WBFMM_HIP_ERR_NOMEM;
  size_t bytes = 0;
  WBFMM_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, codes, k2, idx, v2,
                                        n, 0, 60, stream()));
  if (dmalloc(&tmp, bytes)) return WBFMM_HIP_ERR_NOMEM;
  timing_begin("tree_sort");
  WBFMM_CHECK(rocprim::radix_sort_pairs(tmp, bytes, codes, k2, idx, v2, n,
                                        0, 60, stream()));
  timing_end("tree_sort");
  WBFMM_CHECK(hipMemcpyAsync(codes, k2, sizeof(uint64_t) * n,
                             hipMemcpyDeviceToDevice, stream()));
  WBFMM_CHECK(hipMemcpyAsync(idx, v2, sizeof(uint32_t) * n,
                             hipMemcpyDeviceToDevice, stream()));
  dfree(tmp);
  dfree(k2);
  dfree(v2);
  return 0;
}

int wbfmm_hip_gather4(const double *in, const uint32_t *idx, unsigned n,
                      double *out) {
  if (n == 0) return 0;
  WBFMM_LAUNCH("gather4", k_gather4, dim3(nblk(n, 256)), dim3(256), 0, in,
               idx, n, out);
  return 0;
}

int wbfmm_hip_gather_strided(const double *src, int sstr, int ncomp,
                             const uint32_t *idx, unsigned n, double *out) {
  if (n == 0) return 0;
  if (sstr < ncomp) return WBFMM_HIP_ERR_ARG;
  const double *dsrc = src;
  void *staged = nullptr;
  if (!wbfmm_hip_is_device(src)) {
    size_t bytes = sizeof(double) * ((size_t)(n - 1) * sstr + ncomp);
    // the index range covers 0..n-1 (a permutation of the input)
    if (wbfmm_hip_malloc(&staged, bytes)) return WBFMM_HIP_ERR_NOMEM;
    WBFMM_CHECK(hipMemcpyAsync(staged, src, bytes, hipMemcpyHostToDevice,
                               stream()));
    dsrc = (const double *)staged;
  }
  WBFMM_LAUNCH("gather_strength", k_gather_strided, dim3(nblk(n, 256)),
               dim3(256), 0, dsrc, sstr, ncomp, idx, n, out);
  if (staged) {
    dfree(staged);
    WBFMM_CHECK(hipStreamSynchronize(stream()));
  }
  return 0;
}

int wbfmm_hip_gather_records(const void *in, size_t rsize, const uint32_t *idx, unsigned n,
                             void *out) {
  if (n == 0) return 0;
  if (rsize == 0) return WBFMM_HIP_ERR_ARG;
  if (rsize % 4 == 0 && (uintptr_t)in % 4 == 0 && (uintptr_t)out % 4 == 0) {
    const size_t rw = rsize / 4, tot = (size_t)n * rw;
    WBFMM_LAUNCH("gather_records", k_gather_records4, dim3((unsigned)((tot + 255) / 256)),
                 dim3(256), 0, (const uint32_t *)in, rw, idx, n, (uint32_t *)out);
  } else {
    const size_t tot = (size_t)n * rsize;
    WBFMM_LAUNCH("gather_records", k_gather_records1, dim3((unsigned)((tot + 255) / 256)),
                 dim3(256), 0, (const uint8_t *)in, rsize, idx, n, (uint8_t *)out);
  }
  return 0;
}

int wbfmm_hip_box_ranges(const uint64_t *codes, unsigned n, unsigned level,
                         uint32_t *start, uint32_t *end) {
  size_t nb = (size_t)1 << (3 * level);
  WBFMM_CHECK(hipMemsetAsync(start, 0, sizeof(uint32_t) * nb, stream()));
  WBFMM_CHECK(hipMemsetAsync(end, 0, sizeof(uint32_t) * nb, stream()));
  if (n == 0) return 0;
  WBFMM_LAUNCH("tree_box_ranges", k_box_ranges, dim3(nblk(n, 256)),
               dim3(256), 0, codes, n, (int)(3 * (20 - level)), start, end);
  return 0;
}

int wbfmm_hip_compact_boxes(const uint32_t *start, const uint32_t *end,
                            const uint32_t *flag, unsigned level,
                            uint32_t *list, unsigned *count) {
  unsigned nb = 1u << (3 * level);
  uint32_t *flags = nullptr, *pos = nullptr;
  void *tmp = nullptr;
  if (dmalloc((void **)&flags, sizeof(uint32_t) * nb) ||
      dmalloc((void **)&pos, sizeof(uint32_t) * nb))
    return WBFMM_HIP_ERR_NOMEM;
  WBFMM_LAUNCH("tree_box_flags", k_box_flags, dim3(nblk(nb, 256)), dim3(256),
               0, start, end, flag, nb, flags);
  size_t bytes = 0;
  WBFMM_CHECK(rocprim::exclusive_scan(nullptr, bytes, flags, pos, 0u, nb,
                                      rocprim::plus<uint32_t>(), stream()));
  if (dmalloc(&tmp, bytes)) return WBFMM_HIP_ERR_NOMEM;
  WBFMM_CHECK(rocprim::exclusive_scan(tmp, bytes, flags, pos, 0u, nb,
                                      rocprim::plus<uint32_t>(), stream()));
  WBFMM_LAUNCH("tree_box_scatter", k_box_scatter, dim3(nblk(nb, 256)),
               dim3(256), 0, flags, pos, nb, list);
  uint32_t lastf = 0, lastp = 0;
  WBFMM_CHECK(hipMemcpyAsync(&lastf, flags + nb - 1, 4, hipMemcpyDeviceToHost,
                             stream()));
  WBFMM_CHECK(hipMemcpyAsync(&lastp, pos + nb - 1, 4, hipMemcpyDeviceToHost,
                             stream()));
  dfree(tmp);
  dfree(flags);
  dfree(pos);
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  *count = lastf + lastp;
  return 0;
}

int wbfmm_hip_bounds(const void *pts, size_t pstr, unsigned n, double *lo,
                     double *hi) {
  for (int c = 0; c < 3; c++) {
    lo[c] = INFINITY;
    hi[c] = -INFINITY;
  }
  if (n == 0) return 0;
  if (!wbfmm_hip_is_device(pts)) {
    for (unsigned i = 0; i < n; i++) {
      const double *x = (const double *)((const char *)pts + (size_t)i * pstr);
      for (int c = 0; c < 3; c++) {
        lo[c] = fmin(lo[c], x[c]);
        hi[c] = fmax(hi[c], x[c]);
      }
    }
    return 0;
  }
  const unsigned nb = nblk(n, 256) < 512 ? nblk(n, 256) : 512;
  double *d = nullptr;
  if (dmalloc((void **)&d, sizeof(double) * 6 * nb)) return WBFMM_HIP_ERR_NOMEM;
  std::vector<double> h(6 * (size_t)nb);
  WBFMM_LAUNCH("bounds", k_bounds, dim3(nb), dim3(256), 0,
               (const char *)pts, pstr, n, d);
  WBFMM_CHECK(hipMemcpyAsync(h.data(), d, sizeof(double) * 6 * nb,
                             hipMemcpyDeviceToHost, stream()));
  dfree(d);
  WBFMM_CHECK(hipStreamSynchronize(stream()));
  for (unsigned b = 0; b < nb; b++)
    for (int c = 0; c < 3; c++) {
      lo[c] = fmin(lo[c], h[6 * b + c]);
      hi[c] = fmax(hi[c], h[6 * b + 3 + c]);
    }
  return 0;
}

int wbfmm_hip_accumulate(double *dst, int dstr, const double *src, int sstr,
                         int ncomp, unsigned n) {
  if (n == 0) return 0;
  if (dstr < ncomp || sstr < ncomp) return WBFMM_HIP_ERR_ARG;
  WBFMM_LAUNCH("accumulate", k_accumulate, dim3(nblk(n, 256)), dim3(256), 0,
               dst, dstr, src, sstr, ncomp, n);
  return 0;
}

int wbfmm_hip_mark_boxes(const uint64_t *codes, unsigned n, unsigned level,
                         uint32_t *flag) {
  if (n == 0) return 0;
  WBFMM_LAUNCH("tree_mark_boxes", k_mark_boxes, dim3(nblk(n, 256)),
               dim3(256), 0, codes, n, (int)(3 * (20 - level)), flag);
  return 0;
}

}  // extern "C"
