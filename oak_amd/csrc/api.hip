// C ABI (include/oakg.h): device context, staging of host arrays, the zone
// loop of locAnalysisIncrement (rrsqrt.F90:258-426) in chunks of zones, and
// the error contract (outputs untouched on error with host pointers).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/oakg.h"
#include "analysis.hpp"
#include "common.hpp"
#include "internal.hpp"
#include "interp.hpp"

#ifndef OAKG_PIPE_SLOTS
#define OAKG_PIPE_SLOTS 4  // at most this many zone chunks in flight on as many streams
#endif

namespace oakg {
// Bump allocator over a fixed device buffer (the no-allocation entry point's
// scratch array): buffers of one call, 256-byte aligned, released together.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* get(size_t bytes) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at > cap || bytes > cap - at) return nullptr;
    used = at + bytes;
    return base + at;
  }
  size_t left() const {
    const size_t at = (used + 255) & ~(size_t)255;
    return at < cap ? cap - at : 0;
  }
};
}  // namespace oakg

struct oakg_ctx {
  int device = 0;
  int precision = OAKG_FP64;
  hipStream_t stream = nullptr;
  std::string err;
  oakg::Workspace ws;
  unsigned* d_err = nullptr;
  oakg_stats stats{};
  std::vector<int64_t> counts;
  bool timing = false;
  oakg::PhaseTimer timer;
  size_t ws_budget = (size_t)8 << 30;  // chunk workspace of one zone loop (all slots)
  double jacobi_tol = 0;  // 0 = default for the precision
  int max_sweeps = 40;
  int eig_method = 1;     // 1 = tridiagonal QL (default), 0 = one-sided Jacobi
  // last call, for the lazily copied per-zone counts
  int64_t last_nz = 0;
  const int64_t* last_sel_ptr = nullptr;
  int64_t* d_stats = nullptr;  // device: error word + reduced statistics
  bool counts_ready = true;
  // internal streams of the chunk pipeline (default priority) and their events
  hipStream_t aux[OAKG_PIPE_SLOTS - 1] = {};
  hipEvent_t ev_fork = nullptr, ev_join[OAKG_PIPE_SLOTS - 1] = {};
  // tuning knobs (oakg_set_tuning): zone chunks in flight, zones per QL wave
  int tune_slots = 3;
  int tune_ql_zw = 16;
  int tune_order = 0;  // 1: zones ordered by observation count (k_zone_order)
  int tune_tri_split = 0;  // 1: N = 128 Householder reduction as two launches
  int tune_dual = 0;  // N = 128: zones with 1 <= p < tune_dual solved in observation space
  int tune_persist = 0;  // N = 128: the Gram as a persistent kernel (2 zone per workgroup, 1 per wave)
  // the zone loop captured in a HIP graph and replayed while its launches
  // (kernel arguments, buffers, streams) stay the same
  bool graph = false;
  hipStream_t cap_stream = nullptr;
  hipEvent_t ev_cap_in = nullptr, ev_cap_out = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<char> gkey;
  int64_t graph_replays = 0, graph_captures = 0;
  // events around each call's device work (its ms_total), created once
  hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
  // set only during a call of oakg_local_analysis_csr_noalloc: every buffer
  // of the call comes from this arena (the device's scratch array)
  oakg::Arena* arena = nullptr;
};

namespace oakg {

__global__ void k_fill_i64(int64_t* p, int64_t n, int64_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u32(unsigned* p, int64_t n, unsigned v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// Per-call statistics reduced on the device: [0] error bits, [1] zones with
// p > 0, [2] sum p, [3] max p, [4] max sweeps, [5] zones the QL eigensolver
// handed to the Jacobi fallback.  One block.
__global__ __launch_bounds__(1024) void k_call_stats(int64_t nz, const int64_t* __restrict__ sel_ptr,
                                                     const int32_t* __restrict__ sweeps,
                                                     const unsigned* __restrict__ err,
                                                     int64_t* __restrict__ out) {
  __shared__ int64_t sh[3][1024];
  int64_t na = 0, pm = 0, sm = 0;
  for (int64_t z = threadIdx.x; z < nz; z += blockDim.x) {
    const int64_t p = sel_ptr[z + 1] - sel_ptr[z];
    na += p > 0;
    pm = p > pm ? p : pm;
    if (sweeps) sm = sweeps[z] > sm ? sweeps[z] : sm;
  }
  const int t = threadIdx.x;
  sh[0][t] = na; sh[1][t] = pm; sh[2][t] = sm;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
      sh[0][t] += sh[0][t + s];
      sh[1][t] = sh[1][t] > sh[1][t + s] ? sh[1][t] : sh[1][t + s];
      sh[2][t] = sh[2][t] > sh[2][t + s] ? sh[2][t] : sh[2][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = (int64_t)err[0];
    out[5] = (int64_t)err[1];
    out[1] = sh[0][0];
    out[2] = nz > 0 ? sel_ptr[nz] - sel_ptr[0] : 0;
    out[3] = sh[1][0];
    out[4] = sh[2][0];
  }
}

void fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_i64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}
__global__ void k_fill_i32b(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_i32b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}
void fill_u32(unsigned* p, int64_t n, unsigned v, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}
__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// Order of the zone loop by observation count p: buckets of a quarter octave
// of p, the largest counts first, empty zones last; the ranks are dealt
// round-robin to the K chunks of the loop, chunk c holding positions
// c q + min(c, r) .. (+ q + (c < r)) with nz = K q + r, so that every chunk
// gets the same mix and the heaviest zones of each chunk start first.  One
// workgroup; the order inside a bucket is that of the LDS atomics, which
// changes the schedule and never the results (each zone is computed alone).
constexpr int kOrdBuckets = 64;
__device__ __forceinline__ int ord_bucket(int64_t p) {
  if (p <= 0) return kOrdBuckets - 1;
  const int e = 63 - __clzll((unsigned long long)p);  // floor(log2 p)
  const int f = e >= 2 ? (int)((p >> (e - 2)) & 3) : (int)((p << (2 - e)) & 3);
  return max(0, kOrdBuckets - 2 - (4 * e + f));  // p = 1 -> 62, p = 8000 -> 11
}
__global__ __launch_bounds__(1024) void k_zone_order(int64_t nz, int64_t K,
                                                     const int64_t* __restrict__ sel_ptr,
                                                     int32_t* __restrict__ zorder) {
  __shared__ unsigned long long cnt[kOrdBuckets];
  if (threadIdx.x < kOrdBuckets) cnt[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t z = threadIdx.x; z < nz; z += blockDim.x)
    atomicAdd(&cnt[ord_bucket(sel_ptr[z + 1] - sel_ptr[z])], 1ull);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long acc = 0;
    for (int b = 0; b < kOrdBuckets; ++b) {
      const unsigned long long k = cnt[b];
      cnt[b] = acc;  // first rank of the bucket
      acc += k;
    }
  }
  __syncthreads();
  const int64_t q = nz / K, r = nz % K;
  for (int64_t z = threadIdx.x; z < nz; z += blockDim.x) {
    const int64_t i = (int64_t)atomicAdd(&cnt[ord_bucket(sel_ptr[z + 1] - sel_ptr[z])], 1ull);
    const int64_t ch = i % K, j = i / K;
    zorder[ch * q + min(ch, r) + j] = (int32_t)z;
  }
}

// The two sides of the dual split (run_zones): position -> zone for the
// zones of each side, -1 at the other side's positions.
__global__ void k_dual_split(int64_t nz, const int32_t* __restrict__ zorder,
                             const int64_t* __restrict__ sel_ptr, int dual_lt,
                             int32_t* __restrict__ zord_p, int32_t* __restrict__ zord_d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nz) return;
  const int32_t z = zorder ? zorder[i] : (int32_t)i;
  const int64_t p = sel_ptr[z + 1] - sel_ptr[z];
  const bool dual = p >= 1 && p < dual_lt;
  zord_p[i] = dual ? -1 : z;
  zord_d[i] = dual ? z : -1;
}

void fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}
static void fill_f64(double* p, int64_t n, double v, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

}  // namespace oakg

using namespace oakg;

namespace {

struct Fail {
  int code;
};

void check(oakg_ctx* c, hipError_t e, const char* what) {
  if (e != hipSuccess) {
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    throw Fail{e == hipErrorOutOfMemory ? OAKG_ENOMEM : OAKG_EHIP};
  }
}

void arg_fail(oakg_ctx* c, const std::string& msg) {
  c->err = msg;
  throw Fail{OAKG_EARG};
}

// A device buffer of the call: from the context's grow-only workspace, or,
// during a no-allocation call, from the scratch arena (nullptr when full).
void* wsget(oakg_ctx* c, Workspace::Slot s, size_t bytes) {
  if (c->arena) return c->arena->get(bytes == 0 ? 8 : bytes);
  return c->ws.get(s, bytes);
}

int padded_n(int N) {
  if (N <= 16) return 16;
  if (N <= 32) return 32;
  if (N <= 64) return 64;
  if (N <= 128) return 128;
  if (N <= 256) return 256;
  return 0;
}

// Stages a host array into a workspace slot (or returns the device pointer).
template <typename T>
const T* stage(oakg_ctx* c, Workspace::Slot s, const T* p, int64_t count, bool dev) {
  if (!p || count <= 0) return p;
  if (dev) return p;
  T* d = (T*)wsget(c, s, sizeof(T) * count);
  if (!d) check(c, hipErrorOutOfMemory, "staging");
  check(c, hipMemcpyAsync(d, p, sizeof(T) * count, hipMemcpyHostToDevice, c->stream), "H2D");
  return d;
}

// Stages a column-major rows x cols block with leading dimension ld into a
// dense (ld = rows) device copy.
template <typename T>
const T* stage2d(oakg_ctx* c, Workspace::Slot s, const T* p, int64_t rows, int cols, int64_t ld,
                 bool dev, int64_t* ld_out) {
  if (dev) {
    *ld_out = ld;
    return p;
  }
  T* d = (T*)wsget(c, s, sizeof(T) * std::max<int64_t>(rows, 1) * cols);
  if (!d) check(c, hipErrorOutOfMemory, "staging");
  if (rows > 0 && ld == rows)
    check(c, hipMemcpyAsync(d, p, sizeof(T) * rows * cols, hipMemcpyHostToDevice, c->stream),
          "H2D 2D");
  else if (rows > 0)
    check(c,
          hipMemcpy2DAsync(d, sizeof(T) * rows, p, sizeof(T) * ld, sizeof(T) * rows, cols,
                           hipMemcpyHostToDevice, c->stream),
          "H2D 2D");
  *ld_out = rows;
  return d;
}

// Copies a dense rows x cols device block into a host block with leading
// dimension ld: one 1-D copy when the host block is dense too, else one copy
// per column.  (hipMemcpy2D from device to host has returned success without
// copying anything under rocprofv3's kernel tracing.)
template <typename T>
void copy_out_2d(oakg_ctx* c, T* dst, int64_t ld, const T* src, int64_t rows, int cols,
                 const char* what) {
  if (rows <= 0 || cols <= 0) return;
  if (ld == rows) {
    check(c, hipMemcpy(dst, src, sizeof(T) * rows * cols, hipMemcpyDeviceToHost), what);
    return;
  }
  for (int k = 0; k < cols; ++k)
    check(c, hipMemcpy(dst + (size_t)k * ld, src + (size_t)k * rows, sizeof(T) * rows,
                       hipMemcpyDeviceToHost),
          what);
}

struct Geo {
  ZonesDev zd;
  ObsDev od;
};

Geo stage_geo(oakg_ctx* c, const oakg_zones* zones, const oakg_obs* obs, bool dev, bool need_geo) {
  Geo g{};
  const int64_t nz = zones ? zones->nzones : 0;
  const int64_t m = obs->m;
  if (zones) {
    g.zd.nzones = nz;
    g.zd.metric = zones->metric;
    g.zd.loctype = zones->loctype;
    g.zd.zone_ptr = stage(c, Workspace::Z_PTR, zones->zone_ptr, nz + 1, dev);
    if (need_geo) {
      g.zd.zpos = stage(c, Workspace::Z_POS, zones->zpos, nz * 4, dev);
      g.zd.corr = stage(c, Workspace::Z_CORR, zones->corr_len, nz, dev);
      g.zd.maxlen = stage(c, Workspace::Z_MAXLEN, zones->max_len, nz, dev);
    }
  }
  g.od.m = m;
  if (need_geo) {
    g.od.x = stage(c, Workspace::O_X, obs->x, m, dev);
    g.od.y = stage(c, Workspace::O_Y, obs->y, m, dev);
    g.od.z = stage(c, Workspace::O_Z, obs->z, m, dev);
    g.od.t = stage(c, Workspace::O_T, obs->t, m, dev);
  }
  g.od.yo = stage(c, Workspace::O_YO, obs->yo, m, dev);
  g.od.rmse2 = stage(c, Workspace::O_RMSE2, obs->rmse2, m, dev);
  g.od.excluded = stage(c, Workspace::O_EXCL, obs->excluded, m, dev);
  return g;
}

void validate_zones_host(oakg_ctx* c, const oakg_zones* z, int64_t n_expect, bool need_geo) {
  if (!z || z->nzones < 0 || !z->zone_ptr) arg_fail(c, "zones: NULL or negative count");
  if (need_geo) {
    if (z->nzones > 0 && (!z->zpos || !z->corr_len || !z->max_len))
      arg_fail(c, "zones: zpos/corr_len/max_len required");
    if (z->metric < 0 || z->metric > 2) arg_fail(c, "zones: metric must be 0, 1 or 2");
    if (z->loctype < 1 || z->loctype > 3) arg_fail(c, "zones: loctype must be 1, 2 or 3");
  }
  (void)n_expect;
}

void validate_zone_ptr_host(oakg_ctx* c, int64_t nz, const int64_t* zp) {
  if (zp[0] != 0) arg_fail(c, "zone_ptr[0] must be 0");
  for (int64_t z = 0; z < nz; ++z)
    if (zp[z + 1] < zp[z]) arg_fail(c, "zone_ptr must be non-decreasing");
}

void validate_obs(oakg_ctx* c, const oakg_obs* o, bool need_geo, int loctype) {
  if (!o || o->m < 0) arg_fail(c, "obs: NULL or negative count");
  if (o->m > 0 && (!o->yo || !o->rmse2)) arg_fail(c, "obs: yo and rmse2 required");
  if (need_geo && o->m > 0) {
    if (loctype == 1 && (!o->x || !o->y)) arg_fail(c, "obs: x and y required for loctype 1");
    if (loctype == 2 && !o->z) arg_fail(c, "obs: z required for loctype 2");
    if (loctype == 3 && !o->t) arg_fail(c, "obs: t required for loctype 3");
  }
  if (o->m > INT32_MAX) arg_fail(c, "obs: m above 2^31-1");
}

void begin_call(oakg_ctx* c) {
  c->err.clear();
  check(c, hipSetDevice(c->device), "hipSetDevice");
  if (c->arena) {
    // no-allocation call: the error word and the statistics come first in
    // the scratch (NoallocCall swaps them in); no events are created
  } else {
    if (!c->d_err) {
      check(c, hipMalloc(&c->d_err, 2 * sizeof(unsigned)), "hipMalloc");
      check(c, hipMalloc(&c->d_stats, 8 * sizeof(int64_t)), "hipMalloc");
    }
    if (!c->ev_t0) check(c, hipEventCreate(&c->ev_t0), "event");
    if (!c->ev_t1) check(c, hipEventCreate(&c->ev_t1), "event");
  }
  fill_u32(c->d_err, 2, 0u, c->stream);  // [0] error bits, [1] eigensolver fallbacks
  memset(&c->stats, 0, sizeof c->stats);
  c->counts_ready = true;
  c->counts.clear();
  if (c->timing) c->timer.reset(c->stream);
}

// Ends a call: reduces the statistics on the device, brings them back with
// the error word in one copy, synchronizes, and turns device errors into
// return codes.  sel_ptr may be NULL (no zone loop in the call).
void finish_call(oakg_ctx* c, int64_t nz, const int64_t* sel_ptr, const int32_t* sweeps) {
  int64_t h[6] = {0, 0, 0, 0, 0, 0};
  hipLaunchKernelGGL(k_call_stats, dim3(1), dim3(1024), 0, c->stream, sel_ptr ? nz : 0,
                     sel_ptr, sweeps, c->d_err, c->d_stats);
  check(c, hipGetLastError(), "stats launch");
  check(c, hipMemcpyAsync(h, c->d_stats, sizeof h, hipMemcpyDeviceToHost, c->stream), "D2H stats");
  check(c, hipStreamSynchronize(c->stream), "sync");
  c->stats.zones_analysed = h[1];
  c->stats.obs_selected = h[2];
  c->stats.pmax = h[3];
  c->stats.sweeps_max = (int32_t)h[4];
  c->stats.eig_fallbacks = (int32_t)h[5];
  const unsigned herr = (unsigned)h[0];
  if (herr & OAKG_DERR_RMSE) arg_fail(c, "observation error variance rmse^2 <= 0 or not finite");
  if (herr & OAKG_DERR_IDX) arg_fail(c, "index out of range in H or in the selection");
  if (herr & OAKG_DERR_NOCONV) {
    c->err = "the QL iteration did not finish a zone (no fallback above N = 128)";
    throw Fail{OAKG_ECONV};
  }
  if (herr & OAKG_DERR_NAN) {
    c->err = "NaN in the analysis amplitudes (rrsqrt.F90:145-149)";
    throw Fail{OAKG_ENAN};
  }
}

// Records the phase times of the call and where its per-zone counts are;
// the counts are copied to the host only when asked for (oakg_get_counts).
void collect_stats(oakg_ctx* c, int64_t nz, const int64_t* d_sel_ptr, double ms_total) {
  c->last_nz = nz;
  c->last_sel_ptr = d_sel_ptr;
  c->counts_ready = false;
  c->stats.zones = nz;
  c->stats.ms_total = ms_total;
  if (c->timing) {
    double ms[PH_COUNT];
    c->timer.collect(ms);
    c->stats.ms_prep = ms[PH_PREP];
    c->stats.ms_select = ms[PH_SELECT];
    c->stats.ms_gram = ms[PH_GRAM];
    c->stats.ms_eig = ms[PH_EIG];
    c->stats.ms_tmat = ms[PH_TMAT];
    c->stats.ms_update = ms[PH_UPDATE];
  }
}

void materialize_counts(oakg_ctx* c) {
  if (c->counts_ready) return;
  const int64_t nz = c->last_nz;
  std::vector<int64_t> ptr(nz + 1);
  check(c, hipSetDevice(c->device), "hipSetDevice");
  check(c, hipMemcpy(ptr.data(), c->last_sel_ptr, sizeof(int64_t) * (nz + 1),
                     hipMemcpyDeviceToHost), "D2H counts");
  c->counts.resize(nz);
  for (int64_t z = 0; z < nz; ++z) c->counts[z] = ptr[z + 1] - ptr[z];
  c->counts_ready = true;
}

// Runs the chunked zone loop (gram, eigensolve, tmat, update) for all zones.
// Without phase timing, consecutive chunks rotate over the caller's stream
// and OAKG_PIPE_SLOTS - 1 internal streams, each with its own part of the
// chunk workspace, so that the tail of one chunk's kernels (the QL sweeps of
// its slowest zones, above all) overlaps the next chunks' work.  The streams are
// joined by events before the call continues.  With phase timing the chunks
// run in order on the caller's stream (the timer's events are on that stream).
void run_zones(oakg_ctx* c, int NP, ChunkArgs base, int64_t nz, int32_t* sweeps_all) {
  // QL rotation log: records of 16 B per zone, a multiple of 4 (the log is
  // interleaved in groups of 4 records over the 64 zones of a QL wave)
  // (about 1.1 N^2 records for typical zones; 2 N^2 + 8 N leaves room for slow ones)
  const int64_t cap = (((int64_t)2 * NP * NP + 8 * NP) + 3) & ~(int64_t)3;
  const bool ql = c->eig_method == 1 || NP == 256;  // N > 128: QL only
  // zones with 1 <= p < dual_lt in observation space (dual.hip): N > 64, the
  // analysis modes (not the likelihood), not in a no-allocation call
  const int dual_lt = (NP == 128 && base.mode != 2 && !c->arena && nz > 0) ? c->tune_dual : 0;
  constexpr int PD = 64;
  const size_t per_zone = sizeof(double) * ((size_t)(NP == 256 ? 3 : 2) * NP * NP + 6 * NP) +
                          (ql ? (size_t)cap * 16 + 4 : 0) +
                          (dual_lt ? sizeof(double) * ((size_t)2 * PD * PD + 4 * PD) : 0);
  const size_t budget = c->ws_budget;
  const int nslot = (!c->timing && !c->arena && nz >= 4096) ? c->tune_slots : 1;
  int64_t zc = std::max<int64_t>(1, (int64_t)(budget / (per_zone * nslot)));
  zc = std::min<int64_t>(zc, 1 << 20);
  if (nslot > 1) zc = std::min<int64_t>(zc, (nz + nslot - 1) / nslot);  // a chunk per slot
  zc = std::min<int64_t>(zc, nz);
  zc = (zc + 63) & ~(int64_t)63;  // slots start on 64-zone groups of the rotation log
  if (c->arena) {
    // chunks of what is left of the scratch (7 buffers of up to 255 bytes of
    // alignment), in whole 64-zone groups of the rotation log
    const size_t left = c->arena->left();
    const int64_t fit = left > 8 * 256 ? (int64_t)((left - 8 * 256) / per_zone) : 0;
    zc = std::min<int64_t>(zc, fit & ~(int64_t)63);
    if (zc < 64) {
      c->err = "the no-allocation scratch is too small for one group of 64 zones at this N";
      throw Fail{OAKG_EARG};
    }
  }
  const int64_t zs = zc * nslot;
  double* mat = (double*)wsget(c, Workspace::MAT, sizeof(double) * (size_t)NP * NP * zs);
  double* ymat = (double*)wsget(c, Workspace::YMAT, sizeof(double) * (size_t)NP * NP * zs);
  double* vec = (double*)wsget(c, Workspace::VEC, sizeof(double) * (size_t)4 * NP * zs);
  double* umat = nullptr;  // NP = 256: Q, U, Y (the reflectors stay in ymat)
  if (NP == 256) umat = (double*)wsget(c, Workspace::UMAT, sizeof(double) * (size_t)NP * NP * zs);
  if (!mat || !ymat || !vec || (NP == 256 && !umat)) check(c, hipErrorOutOfMemory, "workspace");
  double *dvec = nullptr, *evec = nullptr;
  char* rlog = nullptr;
  int32_t* tst = nullptr;
  if (ql) {
    dvec = (double*)wsget(c, Workspace::DVEC, sizeof(double) * (size_t)NP * zs);
    evec = (double*)wsget(c, Workspace::EVEC, sizeof(double) * (size_t)NP * zs);
    rlog = (char*)wsget(c, Workspace::RLOG, (size_t)16 * cap * zs);
    tst = (int32_t*)wsget(c, Workspace::TQL_STATUS, sizeof(int32_t) * zs);
    if (!dvec || !evec || !rlog || !tst) check(c, hipErrorOutOfMemory, "workspace");
  }
  // zone order (not in a no-allocation call: its scratch is sized without it)
  const bool order = c->tune_order == 1 && !c->arena && nz > 1 && nz <= INT32_MAX;
  const int64_t nchunk = nz > 0 ? (nz + zc - 1) / zc : 0;
  int32_t* zorder = nullptr;
  if (order) {
    zorder = (int32_t*)wsget(c, Workspace::ZORDER, sizeof(int32_t) * (size_t)nz);
    if (!zorder) check(c, hipErrorOutOfMemory, "workspace");
  }
  // the dual split: the loop's order with -1 at the zones of the other side,
  // and the NP = 64 buffers of the dual side
  int32_t *zord_p = nullptr, *zord_d = nullptr;
  double *mat_d = nullptr, *ymat_d = nullptr, *vec_d = nullptr;
  if (dual_lt) {
    if (nz > INT32_MAX) arg_fail(c, "more than 2^31-1 zones");
    zord_p = (int32_t*)wsget(c, Workspace::ZORDER_P, sizeof(int32_t) * (size_t)nz);
    zord_d = (int32_t*)wsget(c, Workspace::ZORDER_D, sizeof(int32_t) * (size_t)nz);
    mat_d = (double*)wsget(c, Workspace::MAT_D, sizeof(double) * (size_t)PD * PD * zs);
    ymat_d = (double*)wsget(c, Workspace::YMAT_D, sizeof(double) * (size_t)PD * PD * zs);
    vec_d = (double*)wsget(c, Workspace::VEC_D, sizeof(double) * (size_t)4 * PD * zs);
    if (!zord_p || !zord_d || !mat_d || !ymat_d || !vec_d) check(c, hipErrorOutOfMemory, "workspace");
  }
  // work counters of the persistent Gram, one per slot on its own 256 B
  const int persist_item = c->tune_persist == 2 ? 1 : 0;
  const int persist_grid =
      NP == 128 && c->tune_persist != 0 && !c->arena && nz > 0 ? persist_grid128() : 0;
  const bool persist = persist_grid > 0;
  unsigned* wq = nullptr;
  if (persist) {
    wq = (unsigned*)wsget(c, Workspace::PERSIST_WQ, (size_t)256 * OAKG_PIPE_SLOTS);
    if (!wq) check(c, hipErrorOutOfMemory, "workspace");
  }
  // streams and events of the pipeline, created before any capture
  if (nslot > 1) {
    if (!c->ev_fork) check(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "event");
    for (int i = 1; i < nslot; ++i) {
      if (!c->aux[i - 1])
        check(c, hipStreamCreateWithFlags(&c->aux[i - 1], hipStreamNonBlocking), "stream");
      if (!c->ev_join[i - 1])
        check(c, hipEventCreateWithFlags(&c->ev_join[i - 1], hipEventDisableTiming), "event");
    }
  }
  // HIP graph: the launches below depend only on these values; the same key
  // replays the captured graph (on the context's capture stream, ordered
  // after the caller's stream and joined back into it)
  const bool use_graph = c->graph && !c->timing && !c->arena;
  std::vector<char> key;
  if (use_graph) {
    struct Key {
      ChunkArgs base;
      int64_t nz, zc, cap;
      int NP, nslot, ql, ql_zw, order, tri_split, dual_lt, persist;
      const void* buf[16];
      hipStream_t stream;
    } kk;
    memset(&kk, 0, sizeof kk);
    kk.base = base;
    kk.nz = nz;
    kk.zc = zc;
    kk.cap = cap;
    kk.NP = NP;
    kk.nslot = nslot;
    kk.ql = ql;
    kk.ql_zw = c->tune_ql_zw;
    kk.order = order;
    kk.tri_split = c->tune_tri_split;
    kk.dual_lt = dual_lt;
    kk.persist = persist ? 1 + persist_item : 0;
    const void* bufs[16] = {mat,  ymat,   vec,    dvec,   evec,  rlog,  tst,   sweeps_all,
                            zorder, zord_p, zord_d, mat_d, ymat_d, vec_d, wq,  umat};
    memcpy(kk.buf, bufs, sizeof bufs);
    kk.stream = c->stream;
    key.assign((const char*)&kk, (const char*)&kk + sizeof kk);
    if (!c->cap_stream)
      check(c, hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking), "stream");
    if (!c->ev_cap_in) check(c, hipEventCreateWithFlags(&c->ev_cap_in, hipEventDisableTiming), "event");
    if (!c->ev_cap_out) check(c, hipEventCreateWithFlags(&c->ev_cap_out, hipEventDisableTiming), "event");
    check(c, hipEventRecord(c->ev_cap_in, c->stream), "event");
    check(c, hipStreamWaitEvent(c->cap_stream, c->ev_cap_in, 0), "stream wait");
    if (c->gexec && key == c->gkey) {
      check(c, hipGraphLaunch(c->gexec, c->cap_stream), "hipGraphLaunch");
      ++c->graph_replays;
      check(c, hipEventRecord(c->ev_cap_out, c->cap_stream), "event");
      check(c, hipStreamWaitEvent(c->stream, c->ev_cap_out, 0), "stream wait");
      return;
    }
    check(c, hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal), "capture");
  }
  hipStream_t s0 = use_graph ? c->cap_stream : c->stream;
  // ends a capture left open by an error below (the stream stays usable)
  struct CaptureGuard {
    hipStream_t s;
    bool on;
    ~CaptureGuard() {
      if (on) {
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      }
    }
  } guard{s0, use_graph};
  if (ql) fill_i32(sweeps_all, nz, 0, s0);
  if (order) {
    hipLaunchKernelGGL(k_zone_order, dim3(1), dim3(1024), 0, s0, nz, nchunk, base.sel_ptr, zorder);
    check(c, hipGetLastError(), "zone order");
  }
  if (dual_lt) {
    hipLaunchKernelGGL(k_dual_split, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s0, nz, zorder,
                       base.sel_ptr, dual_lt, zord_p, zord_d);
    check(c, hipGetLastError(), "dual split");
  }
  hipStream_t S[OAKG_PIPE_SLOTS];
  for (int i = 0; i < OAKG_PIPE_SLOTS; ++i) S[i] = s0;
  if (nslot > 1) {
    check(c, hipEventRecord(c->ev_fork, S[0]), "event");
    for (int i = 1; i < nslot; ++i) {
      S[i] = c->aux[i - 1];
      check(c, hipStreamWaitEvent(S[i], c->ev_fork, 0), "stream wait");
    }
  }
  // chunk k: positions z0 .. z0 + nzc - 1 of the loop (with an order, the
  // layout k_zone_order deals: nz = nchunk q + r, at most zc zones each)
  const int64_t oq = nchunk ? nz / nchunk : 0, orr = nchunk ? nz % nchunk : 0;
  for (int64_t k = 0; k < nchunk; ++k) {
    const int sl = (int)(k % nslot);
    const size_t zo = (size_t)sl * zc;  // first zone of the slot's workspace
    const int64_t z0 = order ? k * oq + std::min(k, orr) : k * zc;
    ChunkArgs A = base;
    A.z0 = z0;
    A.zorder = zorder;
    A.nzc = (int)(order ? oq + (k < orr ? 1 : 0) : std::min<int64_t>(zc, nz - z0));
    A.mat = mat + zo * NP * NP;
    A.ymat = ymat + zo * NP * NP;
    A.vec = vec + zo * 4 * NP;
    A.umat = umat ? umat + zo * NP * NP : nullptr;
    A.sweeps = sweeps_all + z0;
    A.eig_method = ql ? 1 : 0;
    A.dvec = ql ? dvec + zo * NP : nullptr;
    A.evec = ql ? evec + zo * NP : nullptr;
    A.rlog = ql ? (void*)(rlog + (size_t)16 * cap * zo) : nullptr;
    A.rlog_cap = cap;
    A.tql_status = ql ? tst + zo : nullptr;
    A.ql_zw = c->tune_ql_zw;
    A.tri_split = c->tune_tri_split;
    A.persist_wq = persist ? wq + (size_t)64 * sl : nullptr;
    A.persist_item = persist_item;
    A.persist_grid = persist_grid;
    PhaseTimer* tm = c->timing ? &c->timer : nullptr;
    if (dual_lt) {
      // ensemble-space zones, then observation-space zones (writing their
      // transform inputs into the same slots), then the transform and update
      // of every zone of the chunk
      ChunkArgs Ap = A;
      Ap.zorder = zord_p;
      launch_solve128(Ap, S[sl], tm);
      ChunkArgs Ad = A;
      Ad.zorder = zord_d;
      Ad.mat = mat_d + zo * PD * PD;
      Ad.ymat = ymat_d + zo * PD * PD;
      Ad.vec = vec_d + zo * 4 * PD;
      Ad.ampl = nullptr;
      Ad.N = PD;
      launch_dual128(Ad, A.N, A.ymat, A.vec, A.ampl, S[sl], tm);
      launch_apply128(A, S[sl], tm);
    } else {
      launch_chunk(NP, A, S[sl], tm);
    }
    check(c, hipGetLastError(), "chunk launch");
  }
  for (int i = 1; i < nslot; ++i) {
    check(c, hipEventRecord(c->ev_join[i - 1], S[i]), "event");
    check(c, hipStreamWaitEvent(S[0], c->ev_join[i - 1], 0), "stream wait");
  }
  if (use_graph) {
    guard.on = false;
    hipGraph_t g = nullptr;
    check(c, hipStreamEndCapture(c->cap_stream, &g), "end capture");
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    check(c, e, "hipGraphInstantiate");
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = ex;
    c->gkey = key;
    ++c->graph_captures;
    check(c, hipGraphLaunch(c->gexec, c->cap_stream), "hipGraphLaunch");
    check(c, hipEventRecord(c->ev_cap_out, c->cap_stream), "event");
    check(c, hipStreamWaitEvent(c->stream, c->ev_cap_out, 0), "stream wait");
  }
}

template <typename F>
int guarded(oakg_ctx* c, F f) {
  if (!c) return OAKG_EARG;
  try {
    return f();
  } catch (const Fail& e) {
    return e.code;
  } catch (...) {
    c->err = "unexpected host exception";
    return OAKG_EHIP;
  }
}

// Common body of the two locAnalysis entry points.
int local_analysis_impl(oakg_ctx* c, const oakg_zones* zones, int64_t nzones_csr,
                        const int64_t* zone_ptr_csr, const int64_t* sel_ptr_in,
                        const int32_t* sel_idx_in, const double* sel_w_in, const oakg_obs* obs,
                        int32_t N, const double* xf, const double* Hxf, const double* Sf,
                        int64_t ldSf, const double* HSf, int64_t ldHSf, double* xa, double* Sa,
                        int64_t ldSa, double* ampl, int32_t localise, uint32_t flags) {
  const bool dev = flags & OAKG_DEVICE_PTRS;
  const bool csr = sel_ptr_in != nullptr;
  if (N < 1 || N > OAKG_MAX_N) arg_fail(c, "N out of range");
  const int NP = padded_n(N);
  if (!NP) {
    c->err = "N above 256 is not supported by this build";
    return OAKG_EUNSUP;
  }
  if (!csr) validate_zones_host(c, zones, 0, true);
  validate_obs(c, obs, !csr, zones ? zones->loctype : 1);
  const int64_t nz = csr ? nzones_csr : zones->nzones;
  const int64_t* zp_host = csr ? zone_ptr_csr : zones->zone_ptr;
  if (nz < 0 || !zp_host) arg_fail(c, "zone_ptr required");
  if (!xf || !Sf || !xa || !Sa || (obs->m > 0 && (!Hxf || !HSf)))
    arg_fail(c, "xf, Sf, xa, Sa, Hxf and HSf are required");
  int64_t n = 0;
  if (!dev) {
    validate_zone_ptr_host(c, nz, zp_host);
    n = zp_host[nz];
  } else {
    int64_t last = 0;
    if (nz > 0)
      check(c, hipMemcpy(&last, zp_host + nz, sizeof(int64_t), hipMemcpyDeviceToHost), "D2H n");
    n = last;
  }
  const int64_t m = obs->m;
  if (ldSf < n || ldSa < n || (m > 0 && ldHSf < m)) arg_fail(c, "leading dimension too small");
  begin_call(c);
  if (c->timing) c->timer.mark(PH_PREP);
  if (!c->arena) check(c, hipEventRecord(c->ev_t0, c->stream), "event");

  oakg_zones zz{};
  if (csr) {
    zz.nzones = nz;
    zz.zone_ptr = zone_ptr_csr;
  } else {
    zz = *zones;
  }
  Geo g = stage_geo(c, &zz, obs, dev, !csr);
  int64_t ld_sf = 0, ld_hs = 0;
  const double* dSf = stage2d(c, Workspace::IN_S, Sf, n, N, ldSf, dev, &ld_sf);
  const double* dHSf = stage2d(c, Workspace::IN_HSF, HSf, m, N, ldHSf, dev, &ld_hs);
  const double* dxf = stage(c, Workspace::IN_XF, xf, n, dev);
  const double* dHxf = stage(c, Workspace::IN_HXF, Hxf, m, dev);

  double* HS = (double*)wsget(c, Workspace::HS, sizeof(double) * std::max<int64_t>(m, 1) * NP);
  double* innov = (double*)wsget(c, Workspace::INNOV, sizeof(double) * std::max<int64_t>(m, 1));
  double* rho = (double*)wsget(c, Workspace::RHO, sizeof(double) * std::max<int64_t>(m, 1));
  if (!HS || !innov || !rho) check(c, hipErrorOutOfMemory, "workspace");
  PrepLocArgs P{m, N, dHSf, ld_hs, dHxf, g.od.yo, g.od.rmse2, g.od.excluded, HS, innov, rho,
                c->d_err, 1.0};
  launch_prep_loc(NP, P, c->stream);
  check(c, hipGetLastError(), "prep");

  if (c->timing) c->timer.mark(PH_SELECT);
  SelDev sel{};
  int64_t total = 0;
  if (csr) {
    // caller-supplied selection
    const int64_t* dptr = stage(c, Workspace::U_SEL_PTR, sel_ptr_in, nz + 1, dev);
    int64_t tot = 0;
    if (!dev) {
      if (sel_ptr_in[0] != 0) arg_fail(c, "sel_ptr[0] must be 0");
      for (int64_t z = 0; z < nz; ++z)
        if (sel_ptr_in[z + 1] < sel_ptr_in[z]) arg_fail(c, "sel_ptr must be non-decreasing");
      tot = sel_ptr_in[nz];
      for (int64_t e = 0; e < tot; ++e)
        if (sel_idx_in[e] < 0 || sel_idx_in[e] >= m) arg_fail(c, "sel_idx out of range");
      if (!localise)
        for (int64_t z = 0; z < nz; ++z) {
          int64_t pz = sel_ptr_in[z + 1] - sel_ptr_in[z];
          if (pz != 0 && pz != m) arg_fail(c, "localise_obs = 0 needs all or no observations per zone");
        }
    } else if (nz > 0) {
      check(c, hipMemcpy(&tot, sel_ptr_in + nz, sizeof(int64_t), hipMemcpyDeviceToHost), "D2H");
    }
    sel.ptr = dptr;
    sel.idx = stage(c, Workspace::U_SEL_IDX, sel_idx_in, tot, dev);
    sel.w = stage(c, Workspace::U_SEL_W, sel_w_in, tot, dev);
    total = tot;
  } else {
    check(c, run_selection(c->ws, g.zd, g.od, localise, c->stream, c->d_err, sel, &total),
          "selection");
  }

  double* d_ampl = nullptr;
  if (ampl) {
    d_ampl = dev ? ampl : (double*)wsget(c, Workspace::OUT_AMPL, sizeof(double) * (size_t)N * std::max<int64_t>(nz, 1));
    if (!d_ampl) check(c, hipErrorOutOfMemory, "workspace");
    fill_f64(d_ampl, (int64_t)N * nz, 0.0, c->stream);
  }
  double* dxa = dev ? xa : (double*)wsget(c, Workspace::OUT_XA, sizeof(double) * std::max<int64_t>(n, 1));
  double* dSa = dev ? Sa : (double*)wsget(c, Workspace::OUT_S, sizeof(double) * std::max<int64_t>(n, 1) * N);
  const int64_t ld_sa = dev ? ldSa : n;
  if (!dxa || !dSa) check(c, hipErrorOutOfMemory, "workspace");
  int32_t* sweeps = (int32_t*)wsget(c, Workspace::SWEEPS, sizeof(int32_t) * std::max<int64_t>(nz, 1));
  if (!sweeps) check(c, hipErrorOutOfMemory, "workspace");

  ChunkArgs A{};
  A.N = N;
  A.mode = 0;
  A.m = m;
  A.zone_ptr = g.zd.zone_ptr;
  A.sel_ptr = sel.ptr;
  A.sel_idx = sel.idx;
  A.sel_w = sel.w;
  A.rho = rho;
  A.innov = innov;
  A.HS = HS;
  A.ampl = (ampl && !localise) ? d_ampl : nullptr;
  A.err = c->d_err;
  A.tol = c->jacobi_tol > 0 ? c->jacobi_tol : 4.0 * NP * 2.220446049250313e-16;
  A.max_sweeps = c->max_sweeps;
  A.X0 = dxf;
  A.S = dSf;
  A.ldS = ld_sf;
  A.xa = dxa;
  A.Out = dSa;
  A.ldO = ld_sa;
  A.scaling = 1.0;
  A.infl = 1.0;
  A.maxcorr = HUGE_VAL;
  run_zones(c, NP, A, nz, sweeps);
  if (c->timing) c->timer.mark(PH_END);
  if (!c->arena) check(c, hipEventRecord(c->ev_t1, c->stream), "event");
  finish_call(c, nz, sel.ptr, sweeps);
  float ms = 0;
  if (!c->arena) (void)hipEventElapsedTime(&ms, c->ev_t0, c->ev_t1);
  collect_stats(c, nz, sel.ptr, ms);
  (void)total;
  if (!dev) {
    check(c, hipMemcpy(xa, dxa, sizeof(double) * n, hipMemcpyDeviceToHost), "D2H xa");
    copy_out_2d(c, Sa, ldSa, dSa, n, N, "D2H Sa");
    if (ampl)
      check(c, hipMemcpy(ampl, d_ampl, sizeof(double) * (size_t)N * nz, hipMemcpyDeviceToHost),
            "D2H ampl");
  }
  return OAKG_OK;
}

// Innovation negative log-likelihood of the local analysis and its gradient
// with respect to log sigma (per observation), L (per zone) and the
// inflation of the anomalies (see nll.hip).  Geometry-mode selection.
int local_nll_impl(oakg_ctx* c, const oakg_zones* zones, const oakg_obs* obs, int32_t N,
                   const double* Hxf, const double* HSf, int64_t ldHSf, double inflation,
                   double* nll, double* grad_log_sigma, double* grad_corr_len,
                   double* grad_inflation, double* nll_zone, uint32_t flags) {
  const bool dev = flags & OAKG_DEVICE_PTRS;
  if (N < 2 || N > OAKG_MAX_N) arg_fail(c, "N out of range (2..256)");
  const int NP = padded_n(N);
  if (!NP || NP > 128) {
    c->err = "the likelihood supports N up to 128";
    return OAKG_EUNSUP;
  }
  validate_zones_host(c, zones, 0, true);
  validate_obs(c, obs, true, zones->loctype);
  const int64_t nz = zones->nzones, m = obs->m;
  if (!nll || !grad_log_sigma || !grad_corr_len || !grad_inflation)
    arg_fail(c, "nll, grad_log_sigma, grad_corr_len and grad_inflation are required");
  if (m > 0 && (!Hxf || !HSf)) arg_fail(c, "Hxf and HSf are required");
  if (m > 0 && ldHSf < m) arg_fail(c, "leading dimension too small");
  if (!(inflation > 0) || !std::isfinite(inflation)) arg_fail(c, "inflation must be > 0");
  if (!dev) validate_zone_ptr_host(c, nz, zones->zone_ptr);
  begin_call(c);
  if (c->timing) c->timer.mark(PH_PREP);
  check(c, hipEventRecord(c->ev_t0, c->stream), "event");
  Geo g = stage_geo(c, zones, obs, dev, true);
  int64_t ld_hs = 0;
  const double* dHSf = stage2d(c, Workspace::IN_HSF, HSf, m, N, ldHSf, dev, &ld_hs);
  const double* dHxf = stage(c, Workspace::IN_HXF, Hxf, m, dev);
  double* HS = (double*)wsget(c, Workspace::HS, sizeof(double) * std::max<int64_t>(m, 1) * NP);
  double* innov = (double*)wsget(c, Workspace::INNOV, sizeof(double) * std::max<int64_t>(m, 1));
  double* rho = (double*)wsget(c, Workspace::RHO, sizeof(double) * std::max<int64_t>(m, 1));
  double* zbuf = (double*)wsget(c, Workspace::NLL_Z, sizeof(double) * (3 * std::max<int64_t>(nz, 1) + 2));
  double* gth = dev ? grad_log_sigma
                    : (double*)wsget(c, Workspace::NLL_GTH, sizeof(double) * std::max<int64_t>(m, 1));
  if (!HS || !innov || !rho || !zbuf || !gth) check(c, hipErrorOutOfMemory, "workspace");
  double* Jz = zbuf;
  double* gLz = zbuf + nz;
  double* gAz = zbuf + 2 * nz;
  double* tot = zbuf + 3 * nz;
  fill_f64(gth, m, 0.0, c->stream);
  PrepLocArgs P{m, N, dHSf, ld_hs, dHxf, g.od.yo, g.od.rmse2, g.od.excluded, HS, innov, rho,
                c->d_err, inflation};
  launch_prep_loc(NP, P, c->stream);
  check(c, hipGetLastError(), "prep");
  if (c->timing) c->timer.mark(PH_SELECT);
  SelDev sel{};
  int64_t total = 0;
  check(c, run_selection(c->ws, g.zd, g.od, 1, c->stream, c->d_err, sel, &total), "selection");
  int32_t* sweeps = (int32_t*)wsget(c, Workspace::SWEEPS, sizeof(int32_t) * std::max<int64_t>(nz, 1));
  if (!sweeps) check(c, hipErrorOutOfMemory, "workspace");
  ChunkArgs A{};
  A.N = N;
  A.mode = 2;
  A.m = m;
  A.zone_ptr = g.zd.zone_ptr;
  A.sel_ptr = sel.ptr;
  A.sel_idx = sel.idx;
  A.sel_w = sel.w;
  A.rho = rho;
  A.innov = innov;
  A.HS = HS;
  A.err = c->d_err;
  A.tol = c->jacobi_tol > 0 ? c->jacobi_tol : 4.0 * NP * 2.220446049250313e-16;
  A.max_sweeps = c->max_sweeps;
  A.infl = inflation;
  A.corr = g.zd.corr;
  A.nll_J = Jz;
  A.nll_gL = gLz;
  A.nll_gA = gAz;
  A.nll_gth = gth;
  run_zones(c, NP, A, nz, sweeps);
  launch_nll_totals(nz, Jz, gAz, tot, c->stream);
  check(c, hipGetLastError(), "totals");
  if (c->timing) c->timer.mark(PH_END);
  check(c, hipEventRecord(c->ev_t1, c->stream), "event");
  finish_call(c, nz, sel.ptr, sweeps);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev_t0, c->ev_t1);
  collect_stats(c, nz, sel.ptr, ms);
  const hipMemcpyKind k = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  check(c, hipMemcpy(nll, tot, sizeof(double), k), "copy nll");
  check(c, hipMemcpy(grad_inflation, tot + 1, sizeof(double), k), "copy grad");
  if (nz > 0) check(c, hipMemcpy(grad_corr_len, gLz, sizeof(double) * nz, k), "copy grad");
  if (nll_zone && nz > 0) check(c, hipMemcpy(nll_zone, Jz, sizeof(double) * nz, k), "copy nll");
  if (!dev && m > 0)
    check(c, hipMemcpy(grad_log_sigma, gth, sizeof(double) * m, hipMemcpyDeviceToHost), "D2H");
  (void)total;
  return OAKG_OK;
}

// Ensemble analysis of oak_assim (assimilation.F90:3083-3134, 3235, 3299-3320):
// T = double runs the FP64 path, T = float the mixed-precision path.
template <typename T>
int ensemble_impl(oakg_ctx* c, const oakg_zones* zones, const oakg_obs* obs, const oakg_csr* H,
                  const T* Hshift, int32_t N, const T* Ef, int64_t ldEf, T* Ea, int64_t ldEa,
                  double inflation, double max_correction, const double* maxc_elem,
                  uint32_t flags) {
  constexpr bool mixed = std::is_same<T, float>::value;
  return guarded(c, [&] {
    const bool dev = flags & OAKG_DEVICE_PTRS;
    if (!zones || !obs || !H || !Ef || !Ea) arg_fail(c, "zones, obs, H, Ef and Ea are required");
    if (N < 2 || N > OAKG_MAX_N) arg_fail(c, "N out of range (2..256)");
    const int NP = padded_n(N);
    if (!NP) {
      c->err = "N above 256 is not supported by this build";
      return OAKG_EUNSUP;
    }
    validate_zones_host(c, zones, 0, true);
    validate_obs(c, obs, true, zones->loctype);
    const int64_t nz = zones->nzones, m = obs->m;
    if (H->m != m) arg_fail(c, "H->m must equal obs->m");
    if (m > 0 && (!H->row_ptr || (H->nnz > 0 && (!H->col || !H->val)) || !Hshift))
      arg_fail(c, "H arrays and Hshift are required");
    int64_t n = 0;
    if (!dev) {
      validate_zone_ptr_host(c, nz, zones->zone_ptr);
      n = zones->zone_ptr[nz];
      if (H->row_ptr[0] != 0 || H->row_ptr[m] != H->nnz) arg_fail(c, "H->row_ptr inconsistent");
      for (int64_t e = 0; e < H->nnz; ++e)
        if (H->col[e] < 0 || H->col[e] >= n) arg_fail(c, "H column out of range");
    } else if (nz > 0) {
      check(c, hipMemcpy(&n, zones->zone_ptr + nz, sizeof(int64_t), hipMemcpyDeviceToHost), "D2H n");
    }
    if (H->n != n) arg_fail(c, "H->n must equal the state size");
    if (ldEf < n || ldEa < n) arg_fail(c, "leading dimension too small");
    if (!(inflation > 0) || !(max_correction >= 0))
      arg_fail(c, "inflation must be > 0 and max_correction >= 0");
    if (maxc_elem && !dev)
      for (int64_t i = 0; i < n; ++i)
        if (!(maxc_elem[i] >= 0)) arg_fail(c, "max_correction_elem: every entry must be >= 0");
    begin_call(c);
    if (c->timing) c->timer.mark(PH_PREP);
    check(c, hipEventRecord(c->ev_t0, c->stream), "event");
    Geo g = stage_geo(c, zones, obs, dev, true);
    int64_t ld_e = 0;
    const T* dE = stage2d(c, Workspace::IN_S, Ef, n, N, ldEf, dev, &ld_e);
    const int64_t* drp = stage(c, Workspace::H_RP, H->row_ptr, m + 1, dev);
    const int32_t* dcol = stage(c, Workspace::H_COL, H->col, H->nnz, dev);
    const double* dval = stage(c, Workspace::H_VAL, H->val, H->nnz, dev);
    const T* dHs = stage(c, Workspace::H_SHIFT, Hshift, m, dev);
    const double* dmc = stage(c, Workspace::MAXCORR, maxc_elem, n, dev);
    double* xbar = (double*)wsget(c, Workspace::XBAR, sizeof(double) * std::max<int64_t>(n, 1));
    double* HEc = (double*)wsget(c, Workspace::HEC, sizeof(double) * std::max<int64_t>(m, 1) * N);
    double* HS = nullptr;
    float* HS32 = nullptr;
    if (mixed)
      HS32 = (float*)wsget(c, Workspace::HS32, sizeof(float) * std::max<int64_t>(m, 1) * NP);
    else
      HS = (double*)wsget(c, Workspace::HS, sizeof(double) * std::max<int64_t>(m, 1) * NP);
    double* innov = (double*)wsget(c, Workspace::INNOV, sizeof(double) * std::max<int64_t>(m, 1));
    double* rho = (double*)wsget(c, Workspace::RHO, sizeof(double) * std::max<int64_t>(m, 1));
    if (!xbar || !HEc || !(mixed ? (void*)HS32 : (void*)HS) || !innov || !rho)
      check(c, hipErrorOutOfMemory, "workspace");
    const double scaling = sqrt((double)N - 1.0);
    PrepEnsArgs P{};
    P.n = n;
    P.m = m;
    P.N = N;
    P.ldE = ld_e;
    P.row_ptr = drp;
    P.col = dcol;
    P.val = dval;
    P.yo = g.od.yo;
    P.rmse2 = g.od.rmse2;
    P.excl = g.od.excluded;
    P.scaling = scaling;
    P.xbar = xbar;
    P.HEc = HEc;
    P.HS = HS;
    P.HS32 = HS32;
    P.innov = innov;
    P.rho = rho;
    P.err = c->d_err;
    if (mixed) {
      P.E32 = (const float*)dE;
      P.Hshift32 = (const float*)dHs;
    } else {
      P.E = (const double*)dE;
      P.Hshift = (const double*)dHs;
    }
    launch_prep_ens(NP, P, c->stream);
    check(c, hipGetLastError(), "prep");
    if (c->timing) c->timer.mark(PH_SELECT);
    SelDev sel{};
    int64_t total = 0;
    check(c, run_selection(c->ws, g.zd, g.od, 1, c->stream, c->d_err, sel, &total), "selection");
    T* dEa = dev ? Ea : (T*)wsget(c, Workspace::OUT_S, sizeof(T) * std::max<int64_t>(n, 1) * N);
    const int64_t ld_ea = dev ? ldEa : n;
    int32_t* sweeps = (int32_t*)wsget(c, Workspace::SWEEPS, sizeof(int32_t) * std::max<int64_t>(nz, 1));
    if (!dEa || !sweeps) check(c, hipErrorOutOfMemory, "workspace");
    ChunkArgs A{};
    A.N = N;
    A.mode = 1;
    A.m = m;
    A.zone_ptr = g.zd.zone_ptr;
    A.sel_ptr = sel.ptr;
    A.sel_idx = sel.idx;
    A.sel_w = sel.w;
    A.rho = rho;
    A.innov = innov;
    A.HS = HS;
    A.ampl = nullptr;
    A.err = c->d_err;
    A.tol = c->jacobi_tol > 0 ? c->jacobi_tol : 4.0 * NP * 2.220446049250313e-16;
    A.max_sweeps = c->max_sweeps;
    A.X0 = xbar;
    A.ldS = ld_e;
    A.xa = nullptr;
    A.ldO = ld_ea;
    A.prec = mixed ? 1 : 0;
    if (mixed) {
      A.HS32 = HS32;
      A.S32 = (const float*)dE;
      A.Out32 = (float*)dEa;
    } else {
      A.S = (const double*)dE;
      A.Out = (double*)dEa;
    }
    A.scaling = scaling;
    A.infl = inflation;
    A.maxcorr = max_correction;
    A.maxcorr_v = dmc;
    run_zones(c, NP, A, nz, sweeps);
    if (c->timing) c->timer.mark(PH_END);
    check(c, hipEventRecord(c->ev_t1, c->stream), "event");
    finish_call(c, nz, sel.ptr, sweeps);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev_t0, c->ev_t1);
    collect_stats(c, nz, sel.ptr, ms);
    if (!dev) copy_out_2d(c, Ea, ldEa, (const T*)dEa, n, N, "D2H Ea");
    return OAKG_OK;
    });
}

// Statistics of report (assimilation.F90:3584-3628) over the observation
// ranges [rp[r], rp[r+1]) (host array), from d, rmse2, excl and HS already on
// the device (A's layout); writes nranges x 5 values to stats.
void run_report(oakg_ctx* c, int NP, ReportArgs A, int32_t nranges, const int64_t* rp,
                double* stats, bool dev) {
  std::vector<RepPiece> pcs;
  std::vector<int64_t> pbeg(nranges + 1);
  for (int r = 0; r < nranges; ++r) {
    pbeg[r] = (int64_t)pcs.size();
    for (int64_t a = rp[r]; a < rp[r + 1]; a += kRepRows)
      pcs.push_back(RepPiece{a, std::min<int64_t>(a + kRepRows, rp[r + 1])});
  }
  pbeg[nranges] = (int64_t)pcs.size();
  const int64_t np = (int64_t)pcs.size();
  RepPiece* dp = (RepPiece*)wsget(c, Workspace::REP_PIECES, sizeof(RepPiece) * std::max<int64_t>(np, 1));
  int64_t* db = (int64_t*)wsget(c, Workspace::REP_PBEG, sizeof(int64_t) * (nranges + 1));
  double* part = (double*)wsget(c, Workspace::REP_PART,
                                    sizeof(double) * (4 + NP) * std::max<int64_t>(np, 1));
  double* dst = dev ? stats : (double*)wsget(c, Workspace::REP_STATS, sizeof(double) * 5 * nranges);
  if (!dp || !db || !part || !dst) check(c, hipErrorOutOfMemory, "workspace");
  if (np > 0)
    check(c, hipMemcpyAsync(dp, pcs.data(), sizeof(RepPiece) * np, hipMemcpyHostToDevice, c->stream),
          "H2D pieces");
  check(c, hipMemcpyAsync(db, pbeg.data(), sizeof(int64_t) * (nranges + 1), hipMemcpyHostToDevice,
                          c->stream),
        "H2D pieces");
  A.npieces = np;
  A.nranges = nranges;
  A.pieces = dp;
  A.pbeg = db;
  A.part = part;
  A.stats = dst;
  A.err = c->d_err;
  launch_report(NP, A, c->stream);
  check(c, hipGetLastError(), "report launch");
  finish_call(c, 0, nullptr, nullptr);  // synchronizes (the host vectors above stay alive)
  if (!dev)
    check(c, hipMemcpy(stats, dst, sizeof(double) * 5 * nranges, hipMemcpyDeviceToHost), "D2H stats");
}

// Observation ranges of a report call as a host array: range_ptr (nranges+1
// offsets, a device array with dev) or one range [0, m).
std::vector<int64_t> report_ranges(oakg_ctx* c, int64_t m, int32_t nranges,
                                   const int64_t* range_ptr, bool dev) {
  std::vector<int64_t> rp;
  if (!range_ptr) {
    if (nranges != 1) arg_fail(c, "nranges must be 1 without range_ptr");
    rp = {0, m};
    return rp;
  }
  if (nranges < 1) arg_fail(c, "nranges must be at least 1");
  rp.resize(nranges + 1);
  if (dev)
    check(c, hipMemcpy(rp.data(), range_ptr, sizeof(int64_t) * (nranges + 1), hipMemcpyDeviceToHost),
          "D2H range_ptr");
  else
    memcpy(rp.data(), range_ptr, sizeof(int64_t) * (nranges + 1));
  if (rp[0] < 0 || rp[nranges] > m) arg_fail(c, "range_ptr outside [0, m]");
  for (int r = 0; r < nranges; ++r)
    if (rp[r + 1] < rp[r]) arg_fail(c, "range_ptr must be non-decreasing");
  return rp;
}

int report_impl(oakg_ctx* c, int64_t m, int32_t N, const double* yo_Hx, const double* HS,
                int64_t ldHS, const double* rmse2, const uint8_t* excluded, int32_t nranges,
                const int64_t* range_ptr, double* stats, uint32_t flags) {
  const bool dev = flags & OAKG_DEVICE_PTRS;
  if (m < 0 || N < 1) arg_fail(c, "m must be >= 0 and N >= 1");
  const int NP = padded_n(N);
  if (!NP) {
    c->err = "N above 256 is not supported by this build";
    return OAKG_EUNSUP;
  }
  if (!stats) arg_fail(c, "stats is required");
  if (m > 0 && (!yo_Hx || !HS || !rmse2)) arg_fail(c, "yo_Hx, HS and rmse2 are required");
  if (m > 0 && ldHS < m) arg_fail(c, "ldHS must be at least m");
  if (m > INT32_MAX) arg_fail(c, "m above 2^31-1");
  begin_call(c);
  const std::vector<int64_t> rp = report_ranges(c, m, nranges, range_ptr, dev);
  int64_t ld = ldHS;
  ReportArgs A{};
  A.N = N;
  A.layout = 0;
  A.d = stage(c, Workspace::REP_D, yo_Hx, m, dev);
  A.rmse2 = stage(c, Workspace::O_RMSE2, rmse2, m, dev);
  A.excl = stage(c, Workspace::O_EXCL, excluded, m, dev);
  A.HS = m > 0 ? stage2d(c, Workspace::REP_HS, HS, m, N, ldHS, dev, &ld) : nullptr;
  A.ldHS = ld;
  run_report(c, NP, A, nranges, rp.data(), stats, dev);
  return OAKG_OK;
}

// The forecast (E = Ef) or analysis (E = Ea) report of Assim's ensemble
// path: HE = H E + Hshift, Hx = mean(HE, 2), HS = (HE - Hx)/sqrt(N-1), d = yo - Hx
// (assimilation.F90:3112-3134, 3325-3357), by the kernels of the analysis.
template <typename T>
int ensemble_report_impl(oakg_ctx* c, const oakg_obs* obs, const oakg_csr* H, const T* Hshift,
                         int32_t N, const T* E, int64_t ldE, int32_t nranges,
                         const int64_t* range_ptr, double* stats, uint32_t flags) {
  constexpr bool mixed = std::is_same<T, float>::value;
  return guarded(c, [&] {
    const bool dev = flags & OAKG_DEVICE_PTRS;
    if (!obs || !H || !E || !stats) arg_fail(c, "obs, H, E and stats are required");
    if (N < 2 || N > OAKG_MAX_N) arg_fail(c, "N out of range (2..256)");
    const int NP = padded_n(N);
    validate_obs(c, obs, false, 1);
    const int64_t m = obs->m, n = H->n;
    if (H->m != m) arg_fail(c, "H->m must equal obs->m");
    if (n < 1) arg_fail(c, "H->n must be >= 1");
    if (m > 0 && (!H->row_ptr || (H->nnz > 0 && (!H->col || !H->val)) || !Hshift))
      arg_fail(c, "H arrays and Hshift are required");
    if (!dev && m > 0) {
      if (H->row_ptr[0] != 0 || H->row_ptr[m] != H->nnz) arg_fail(c, "H->row_ptr inconsistent");
      for (int64_t e = 0; e < H->nnz; ++e)
        if (H->col[e] < 0 || H->col[e] >= n) arg_fail(c, "H column out of range");
    }
    if (ldE < std::max<int64_t>(n, 1)) arg_fail(c, "leading dimension too small");
    begin_call(c);
    const std::vector<int64_t> rp = report_ranges(c, m, nranges, range_ptr, dev);
    Geo g = stage_geo(c, nullptr, obs, dev, false);
    int64_t ld_e = 0;
    const T* dE = stage2d(c, Workspace::IN_S, E, n, N, ldE, dev, &ld_e);
    const int64_t* drp = stage(c, Workspace::H_RP, H->row_ptr, m + 1, dev);
    const int32_t* dcol = stage(c, Workspace::H_COL, H->col, H->nnz, dev);
    const double* dval = stage(c, Workspace::H_VAL, H->val, H->nnz, dev);
    const T* dHs = stage(c, Workspace::H_SHIFT, Hshift, m, dev);
    double* HEc = (double*)wsget(c, Workspace::HEC, sizeof(double) * std::max<int64_t>(m, 1) * N);
    double* HS = nullptr;
    float* HS32 = nullptr;
    if (mixed)
      HS32 = (float*)wsget(c, Workspace::HS32, sizeof(float) * std::max<int64_t>(m, 1) * NP);
    else
      HS = (double*)wsget(c, Workspace::HS, sizeof(double) * std::max<int64_t>(m, 1) * NP);
    double* innov = (double*)wsget(c, Workspace::INNOV, sizeof(double) * std::max<int64_t>(m, 1));
    double* rho = (double*)wsget(c, Workspace::RHO, sizeof(double) * std::max<int64_t>(m, 1));
    if (!HEc || !(mixed ? (void*)HS32 : (void*)HS) || !innov || !rho)
      check(c, hipErrorOutOfMemory, "workspace");
    PrepEnsArgs P{};
    P.n = n;
    P.m = m;
    P.N = N;
    P.ldE = ld_e;
    P.row_ptr = drp;
    P.col = dcol;
    P.val = dval;
    P.yo = g.od.yo;
    P.rmse2 = g.od.rmse2;
    P.excl = g.od.excluded;
    P.scaling = sqrt((double)N - 1.0);
    P.xbar = nullptr;  // the state mean is not needed
    P.HEc = HEc;
    P.HS = HS;
    P.HS32 = HS32;
    P.innov = innov;
    P.rho = rho;
    P.err = c->d_err;
    if (mixed) {
      P.E32 = (const float*)dE;
      P.Hshift32 = (const float*)dHs;
    } else {
      P.E = (const double*)dE;
      P.Hshift = (const double*)dHs;
    }
    launch_prep_ens(NP, P, c->stream);
    check(c, hipGetLastError(), "prep");
    ReportArgs A{};
    A.N = N;
    A.layout = 1;
    A.d = innov;
    A.rmse2 = g.od.rmse2;
    A.excl = g.od.excluded;
    A.HS = HS;
    A.HS32 = HS32;
    A.ldHS = NP;
    run_report(c, NP, A, nranges, rp.data(), stats, dev);
    return OAKG_OK;
  });
}

// genObservationOper + packSparseMatrix on the GPU (interp.hip): the box
// trees and simplex tables are built on the host from the node coordinates
// (copied back first when they are device arrays), everything per
// observation runs on the device.
int obs_operator_impl(oakg_ctx* c, int32_t nvar, const oakg_grid* grids, int64_t m,
                      const int32_t* obs_var, const double* ox, const double* oy, const double* oz,
                      const double* ot, int64_t* row_ptr, int32_t* col, double* val, int64_t cap,
                      uint8_t* valid, int64_t* nnz, uint32_t flags) {
  const bool dev = flags & OAKG_DEVICE_PTRS;
  if (nvar < 0 || (nvar > 0 && !grids)) arg_fail(c, "grids required");
  if (m < 0 || m > INT32_MAX) arg_fail(c, "m out of range");
  if (!row_ptr || !valid || !nnz || (m > 0 && (!obs_var || !ox))) arg_fail(c, "row_ptr, valid, nnz, obs_var and ox are required");
  if (cap > 0 && (!col || !val)) arg_fail(c, "col and val are required");
  int nmax = 1;
  std::vector<int64_t> nodes(nvar);
  for (int v = 0; v < nvar; ++v) {
    const oakg_grid& g = grids[v];
    if (g.ndim < 1 || g.ndim > 4) arg_fail(c, "grid ndim must be 1..4");
    int64_t tot = 1;
    for (int d = 0; d < g.ndim; ++d) {
      if (g.shape[d] < 2) {
        c->err = "a grid dimension of one node makes every simplex degenerate (OAK's SVD branch): not supported";
        throw Fail{OAKG_EUNSUP};
      }
      if (!g.coord[d]) arg_fail(c, "grid coordinates required");
      tot *= g.shape[d];
    }
    if (tot > INT32_MAX) arg_fail(c, "grid above 2^31 nodes");
    nodes[v] = tot;
    nmax = std::max(nmax, (int)g.ndim);
  }
  const double* pos[4] = {ox, oy, oz, ot};
  for (int v = 0; v < nvar; ++v)
    for (int d = 0; d < grids[v].ndim; ++d)
      if (m > 0 && !pos[d]) arg_fail(c, "observation coordinate missing for a grid's dimension");
  begin_call(c);
  const int stride = 1 << nmax;
  // host side: trees and simplex tables, concatenated over the variables
  std::vector<BoxDev> boxes;
  std::vector<int64_t> box0(nvar), tet0(nvar);
  std::vector<double> tets;
  for (int v = 0; v < nvar; ++v) {
    const oakg_grid& g = grids[v];
    GridHost gh{};
    gh.n = g.ndim;
    std::vector<std::vector<double>> hc(g.ndim);
    for (int d = 0; d < 4; ++d) gh.shape[d] = d < g.ndim ? g.shape[d] : 1;
    for (int d = 0; d < g.ndim; ++d) {
      if (dev) {
        hc[d].resize(nodes[v]);
        check(c, hipMemcpy(hc[d].data(), g.coord[d], sizeof(double) * nodes[v], hipMemcpyDeviceToHost),
              "D2H grid");
        gh.coord[d] = hc[d].data();
      } else {
        gh.coord[d] = g.coord[d];
      }
    }
    // each variable's tree keeps its own box numbering; its launch gets the
    // address of its first box
    std::vector<BoxDev> b = build_box_tree(gh);
    box0[v] = (int64_t)boxes.size();
    boxes.insert(boxes.end(), b.begin(), b.end());
    if ((int64_t)boxes.size() > INT32_MAX) arg_fail(c, "grid trees above 2^31 boxes");
    std::vector<double> t = tet_table(g.ndim);
    tet0[v] = (int64_t)tets.size();
    tets.insert(tets.end(), t.begin(), t.end());
  }
  BoxDev* dbox = (BoxDev*)wsget(c, Workspace::OP_BOXES, sizeof(BoxDev) * std::max<size_t>(boxes.size(), 1));
  double* dtet = (double*)wsget(c, Workspace::OP_TET, sizeof(double) * std::max<size_t>(tets.size(), 1));
  if (!dbox || !dtet) check(c, hipErrorOutOfMemory, "workspace");
  if (!boxes.empty())
    check(c, hipMemcpyAsync(dbox, boxes.data(), sizeof(BoxDev) * boxes.size(), hipMemcpyHostToDevice,
                            c->stream), "H2D boxes");
  if (!tets.empty())
    check(c, hipMemcpyAsync(dtet, tets.data(), sizeof(double) * tets.size(), hipMemcpyHostToDevice,
                            c->stream), "H2D tetrahedra");
  ObsPosDev o{};
  o.var = stage(c, Workspace::OP_VAR, obs_var, m, dev);
  o.x = stage(c, Workspace::OP_X, ox, m, dev);
  o.y = stage(c, Workspace::OP_Y, oy, m, dev);
  o.z = stage(c, Workspace::OP_Z, oz, m, dev);
  o.t = stage(c, Workspace::OP_T, ot, m, dev);
  const int64_t mm = std::max<int64_t>(m, 1);
  int32_t* slotc = (int32_t*)wsget(c, Workspace::OP_SLOTC, sizeof(int32_t) * mm * stride);
  double* slotw = (double*)wsget(c, Workspace::OP_SLOTW, sizeof(double) * mm * stride);
  int64_t* cnt = (int64_t*)wsget(c, Workspace::OP_CNT, sizeof(int64_t) * (mm + 1));
  int64_t* drp = dev ? row_ptr : (int64_t*)wsget(c, Workspace::OP_RP, sizeof(int64_t) * (mm + 1));
  uint8_t* dvalid = dev ? valid : (uint8_t*)wsget(c, Workspace::OP_VALID, mm);
  unsigned* status = (unsigned*)wsget(c, Workspace::OP_STATUS, sizeof(unsigned));
  if (!slotc || !slotw || !cnt || !drp || !dvalid || !status) check(c, hipErrorOutOfMemory, "workspace");
  fill_i32(slotc, m * stride, -1, c->stream);
  fill_i64(cnt, m + 1, 0, c->stream);
  fill_u32(status, 1, 0u, c->stream);
  fill_u8(dvalid, m, 0, c->stream);
  // host grids: staged one after the other into three buffers of the call
  double* gc = nullptr;
  uint8_t* gm = nullptr;
  int64_t* gs = nullptr;
  if (!dev) {
    int64_t nc = 0, nm = 0, ns = 0;
    for (int v = 0; v < nvar; ++v) {
      nc += nodes[v] * grids[v].ndim;
      if (grids[v].masked) nm += nodes[v];
      if (grids[v].state_index) ns += nodes[v];
    }
    gc = (double*)wsget(c, Workspace::OP_GC, sizeof(double) * std::max<int64_t>(nc, 1));
    gm = (uint8_t*)wsget(c, Workspace::OP_GM, std::max<int64_t>(nm, 1));
    gs = (int64_t*)wsget(c, Workspace::OP_GS, sizeof(int64_t) * std::max<int64_t>(ns, 1));
    if (!gc || !gm || !gs) check(c, hipErrorOutOfMemory, "workspace");
  }
  // per variable: its grid arrays on the device and one launch over the observations
  for (int v = 0; v < nvar; ++v) {
    const oakg_grid& g = grids[v];
    GridDev gd{};
    gd.n = g.ndim;
    for (int d = 0; d < 4; ++d) gd.shape[d] = d < g.ndim ? g.shape[d] : 1;
    if (dev) {
      for (int d = 0; d < g.ndim; ++d) gd.coord[d] = g.coord[d];
      gd.mask = g.masked;
      gd.state = g.state_index;
    } else {
      for (int d = 0; d < g.ndim; ++d) {
        check(c, hipMemcpyAsync(gc, g.coord[d], sizeof(double) * nodes[v], hipMemcpyHostToDevice,
                                c->stream), "H2D grid");
        gd.coord[d] = gc;
        gc += nodes[v];
      }
      if (g.masked) {
        check(c, hipMemcpyAsync(gm, g.masked, nodes[v], hipMemcpyHostToDevice, c->stream), "H2D mask");
        gd.mask = gm;
        gm += nodes[v];
      }
      if (g.state_index) {
        check(c, hipMemcpyAsync(gs, g.state_index, sizeof(int64_t) * nodes[v], hipMemcpyHostToDevice,
                                c->stream), "H2D state index");
        gd.state = gs;
        gs += nodes[v];
      }
    }
    if (m > 0)
      launch_obs_interp(g.ndim, m, v, gd, dbox + box0[v], dtet + tet0[v], tet_count(g.ndim), o, stride,
                        slotc, slotw, cnt, dvalid, status, c->stream);
    check(c, hipGetLastError(), "interp launch");
  }
  check(c, scan_counts(c->ws, cnt, drp, m, c->stream), "scan");
  int64_t tot = 0;
  unsigned hst = 0;
  check(c, hipMemcpyAsync(&tot, drp + m, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream), "D2H nnz");
  check(c, hipMemcpyAsync(&hst, status, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream), "D2H status");
  check(c, hipStreamSynchronize(c->stream), "sync");
  if (hst & 1u) {
    c->err = "an observation's cell has degenerate simplices (|det| <= 1e-8, OAK's SVD branch): not supported";
    return OAKG_EUNSUP;
  }
  if (hst & 2u) {
    c->err = "grid search deeper than 48 levels";
    return OAKG_EUNSUP;
  }
  *nnz = tot;
  if (tot > cap) {
    c->err = "cap below the number of entries of H (returned in *nnz)";
    return OAKG_EARG;
  }
  int32_t* dcol = dev ? col : (int32_t*)wsget(c, Workspace::OP_COL, sizeof(int32_t) * std::max<int64_t>(tot, 1));
  double* dval = dev ? val : (double*)wsget(c, Workspace::OP_VAL, sizeof(double) * std::max<int64_t>(tot, 1));
  if (!dcol || !dval) check(c, hipErrorOutOfMemory, "workspace");
  launch_obs_compact(m, stride, drp, slotc, slotw, dcol, dval, c->stream);
  check(c, hipGetLastError(), "compact launch");
  check(c, hipStreamSynchronize(c->stream), "sync");
  if (!dev) {
    check(c, hipMemcpy(row_ptr, drp, sizeof(int64_t) * (m + 1), hipMemcpyDeviceToHost), "D2H row_ptr");
    if (m > 0) check(c, hipMemcpy(valid, dvalid, (size_t)m, hipMemcpyDeviceToHost), "D2H valid");
    if (tot > 0) {
      check(c, hipMemcpy(col, dcol, sizeof(int32_t) * tot, hipMemcpyDeviceToHost), "D2H col");
      check(c, hipMemcpy(val, dval, sizeof(double) * tot, hipMemcpyDeviceToHost), "D2H val");
    }
  }
  return OAKG_OK;
}

// Gather (load) or scatter (save) of ensemble rows between the model's full
// field vectors and the zone-permuted state.
int permute_rows_impl(oakg_ctx* c, bool gather, int64_t n, int32_t N, const int64_t* idx,
                      int64_t full_len, const void* in, int64_t ldi, void* out, int64_t ldo,
                      int32_t elem, uint32_t flags) {
  const bool dev = flags & OAKG_DEVICE_PTRS;
  if (elem != 4 && elem != 8) arg_fail(c, "elem_size must be 4 or 8");
  if (n < 0 || N < 0 || full_len < 0) arg_fail(c, "negative size");
  if (n > 0 && N > 0 && (!idx || !in || !out)) arg_fail(c, "idx, in and out are required");
  const int64_t rin = gather ? full_len : n, rout = gather ? n : full_len;
  if (N > 1 && (ldi < rin || ldo < rout)) arg_fail(c, "leading dimension too small");
  if (N > 65535) arg_fail(c, "N above 65535");
  if (!dev)
    for (int64_t l = 0; l < n; ++l)
      if (idx[l] < 0 || idx[l] >= full_len) arg_fail(c, "idx out of [0, full_len)");
  begin_call(c);
  if (n == 0 || N == 0) return OAKG_OK;
  const int64_t li = N > 1 ? ldi : rin, lo = N > 1 ? ldo : rout;
  const int64_t* didx = stage(c, Workspace::PR_IDX, idx, n, dev);
  const void* din = in;
  void* dout = out;
  int64_t dli = li, dlo = lo;
  if (!dev) {
    char* bi = (char*)wsget(c, Workspace::PR_IN, (size_t)elem * rin * N);
    char* bo = (char*)wsget(c, Workspace::PR_OUT, (size_t)elem * rout * N);
    if (!bi || !bo) check(c, hipErrorOutOfMemory, "workspace");
    for (int k = 0; k < N; ++k) {
      check(c, hipMemcpyAsync(bi + (size_t)elem * rin * k, (const char*)in + (size_t)elem * li * k,
                              (size_t)elem * rin, hipMemcpyHostToDevice, c->stream), "H2D");
      if (!gather)  // the scatter leaves untouched rows as they are
        check(c, hipMemcpyAsync(bo + (size_t)elem * rout * k, (const char*)out + (size_t)elem * lo * k,
                                (size_t)elem * rout, hipMemcpyHostToDevice, c->stream), "H2D");
    }
    din = bi;
    dout = bo;
    dli = rin;
    dlo = rout;
  }
  launch_permute_rows(gather, elem, n, N, didx, din, dli, dout, dlo, full_len, c->d_err, c->stream);
  check(c, hipGetLastError(), "permute launch");
  finish_call(c, 0, nullptr, nullptr);
  if (!dev)
    for (int k = 0; k < N; ++k)
      check(c, hipMemcpy((char*)out + (size_t)elem * lo * k, (const char*)dout + (size_t)elem * rout * k,
                         (size_t)elem * rout, hipMemcpyDeviceToHost), "D2H");
  return OAKG_OK;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point: one fixed array per device,
// reserved when the code object is loaded and held until the process ends.
// Its device address is looked up once per device (with that device
// current); host code uses only that address, never the array's name.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = OAKG_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The scratch address on the context's device (set up on the first call
// there).  The device must be current.
void* scratch_on_device(oakg_ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices) arg_fail(c, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void* ptr = nullptr;
    check(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) check(c, hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[c->device] = ptr;
  }
  return g_scratch_ptr[c->device];
}

// Bytes of scratch a no-allocation call needs besides its zone chunks.
size_t noalloc_fixed_bytes(int64_t m, int NP, int64_t nz) {
  const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t mm = (size_t)std::max<int64_t>(m, 1);
  return al(256) + al(sizeof(double) * mm * NP) + 2 * al(sizeof(double) * mm) +
         al(sizeof(int32_t) * (size_t)std::max<int64_t>(nz, 1));
}

// Bytes of the zone-chunk workspace per zone (as run_zones sizes it).
size_t noalloc_zone_bytes(int NP, bool ql) {
  const int64_t cap = (((int64_t)2 * NP * NP + 8 * NP) + 3) & ~(int64_t)3;
  return sizeof(double) * ((size_t)(NP == 256 ? 3 : 2) * NP * NP + 6 * NP) +
         (ql || NP == 256 ? (size_t)cap * 16 + 4 : 0);
}

// For the duration of a no-allocation call: the arena over the scratch, the
// error word and statistics at its head, no phase timing.
struct NoallocCall {
  oakg_ctx* c;
  Arena arena;
  unsigned* save_err;
  int64_t* save_stats;
  bool save_timing;
  NoallocCall(oakg_ctx* cc, void* base) : c(cc) {
    arena.base = (char*)base;
    arena.cap = kScratchBytes;
    save_err = c->d_err;
    save_stats = c->d_stats;
    save_timing = c->timing;
    char* head = (char*)arena.get(256);
    c->d_err = (unsigned*)head;               // [0] error bits, [1] eigensolver fallbacks
    c->d_stats = (int64_t*)(head + 64);       // 8 statistics
    c->timing = false;
    c->arena = &arena;
  }
  ~NoallocCall() {
    c->arena = nullptr;
    c->d_err = save_err;
    c->d_stats = save_stats;
    c->timing = save_timing;
  }
};

}  // namespace

extern "C" {

int oakg_version(void) { return 10000; }

int oakg_create(oakg_ctx** out, int device, int precision) {
  if (!out) return OAKG_EARG;
  *out = nullptr;
  if (precision != OAKG_FP64 && precision != OAKG_MIXED) return OAKG_EARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return OAKG_EHIP;
  if (hipSetDevice(device) != hipSuccess) return OAKG_EHIP;
  oakg_ctx* c = new (std::nothrow) oakg_ctx;
  if (!c) return OAKG_ENOMEM;
  c->device = device;
  c->precision = precision;
  // chunk workspace: a quarter of the device memory, between 8 and 64 GiB
  // (larger chunks leave fewer part-filled rounds of workgroups at the end
  // of each kernel; 64 GiB holds config 2's 64 800 zones of N = 128)
  // (OAKG_WORKSPACE_BYTES in the environment sets it instead, at least 1 MiB)
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    c->ws_budget = std::min<size_t>(std::max<size_t>(prop.totalGlobalMem / 4, (size_t)8 << 30),
                                    (size_t)64 << 30);
  if (const char* s = getenv("OAKG_WORKSPACE_BYTES")) {
    const long long v = atoll(s);
    if (v > 0) c->ws_budget = std::max<size_t>((size_t)v, (size_t)1 << 20);
  }
  *out = c;
  return OAKG_OK;
}

int oakg_destroy(oakg_ctx* c) {
  if (!c) return OAKG_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < OAKG_PIPE_SLOTS - 1; ++i) {
    if (c->aux[i]) {
      (void)hipStreamSynchronize(c->aux[i]);
      (void)hipStreamDestroy(c->aux[i]);
    }
    if (c->ev_join[i]) (void)hipEventDestroy(c->ev_join[i]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->cap_stream) {
    (void)hipStreamSynchronize(c->cap_stream);
    (void)hipStreamDestroy(c->cap_stream);
  }
  if (c->ev_cap_in) (void)hipEventDestroy(c->ev_cap_in);
  if (c->ev_cap_out) (void)hipEventDestroy(c->ev_cap_out);
  if (c->ev_t0) (void)hipEventDestroy(c->ev_t0);
  if (c->ev_t1) (void)hipEventDestroy(c->ev_t1);
  if (c->d_err) (void)hipFree(c->d_err);
  if (c->d_stats) (void)hipFree(c->d_stats);
  c->ws.release();
  delete c;
  return OAKG_OK;
}

const char* oakg_last_error(const oakg_ctx* c) { return c ? c->err.c_str() : "NULL context"; }

int oakg_set_stream(oakg_ctx* c, void* stream) {
  if (!c) return OAKG_EARG;
  c->stream = (hipStream_t)stream;
  return OAKG_OK;
}

int oakg_set_timing(oakg_ctx* c, int on) {
  if (!c) return OAKG_EARG;
  c->timing = on != 0;
  return OAKG_OK;
}

int oakg_set_eigensolver(oakg_ctx* c, int method) {
  if (!c) return OAKG_EARG;
  if (method != OAKG_EIG_JACOBI && method != OAKG_EIG_QL) {
    c->err = "oakg_set_eigensolver: unknown method";
    return OAKG_EARG;
  }
  c->eig_method = method;
  return OAKG_OK;
}

int oakg_set_tuning(oakg_ctx* c, int key, int value) {
  if (!c) return OAKG_EARG;
  switch (key) {
    case OAKG_TUNE_SLOTS:
      if (value < 1 || value > OAKG_PIPE_SLOTS) break;
      c->tune_slots = value;
      return OAKG_OK;
    case OAKG_TUNE_QL_ZONES:
      if (value != 4 && value != 8 && value != 16 && value != 32) break;
      c->tune_ql_zw = value;
      return OAKG_OK;
    case OAKG_TUNE_ORDER:
      if (value != 0 && value != 1) break;
      c->tune_order = value;
      return OAKG_OK;
    case OAKG_TUNE_TRIDIAG_SPLIT:
      if (value < 0 || value > 4) break;
      c->tune_tri_split = value;
      return OAKG_OK;
    case OAKG_TUNE_DUAL:
      if (value != 0 && value != 16 && value != 32 && value != 64) break;
      c->tune_dual = value;
      return OAKG_OK;
    case OAKG_TUNE_PERSIST:
      if (value < 0 || value > 2) break;
      c->tune_persist = value;
      return OAKG_OK;
    case OAKG_TUNE_GRAPH:
      if (value != 0 && value != 1) break;
      c->graph = value != 0;
      if (!c->graph && c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
        c->gkey.clear();
      }
      return OAKG_OK;
    default:
      break;
  }
  c->err = "oakg_set_tuning: unknown key or value out of range";
  return OAKG_EARG;
}

int oakg_get_tuning(oakg_ctx* c, int key, int64_t* value) {
  if (!c || !value) return OAKG_EARG;
  switch (key) {
    case OAKG_TUNE_SLOTS: *value = c->tune_slots; return OAKG_OK;
    case OAKG_TUNE_QL_ZONES: *value = c->tune_ql_zw; return OAKG_OK;
    case OAKG_TUNE_GRAPH: *value = c->graph ? 1 : 0; return OAKG_OK;
    case OAKG_TUNE_ORDER: *value = c->tune_order; return OAKG_OK;
    case OAKG_TUNE_TRIDIAG_SPLIT: *value = c->tune_tri_split; return OAKG_OK;
    case OAKG_TUNE_DUAL: *value = c->tune_dual; return OAKG_OK;
    case OAKG_TUNE_PERSIST: *value = c->tune_persist; return OAKG_OK;
    case OAKG_TUNE_GRAPH_REPLAYS: *value = c->graph_replays; return OAKG_OK;
    case OAKG_TUNE_GRAPH_CAPTURES: *value = c->graph_captures; return OAKG_OK;
    default: break;
  }
  c->err = "oakg_get_tuning: unknown key";
  return OAKG_EARG;
}

int oakg_get_stats(oakg_ctx* c, oakg_stats* out) {
  if (!c || !out) return OAKG_EARG;
  *out = c->stats;
  return OAKG_OK;
}

int oakg_get_counts(oakg_ctx* c, int64_t* counts, int64_t nzones) {
  if (!c || !counts) return OAKG_EARG;
  return guarded(c, [&] {
    materialize_counts(c);
    if (nzones != (int64_t)c->counts.size()) arg_fail(c, "nzones differs from the last call");
    memcpy(counts, c->counts.data(), sizeof(int64_t) * nzones);
    return OAKG_OK;
  });
}

int oakg_local_analysis(oakg_ctx* c, const oakg_zones* zones, const oakg_obs* obs, int32_t N,
                        const double* xf, const double* Hxf, const double* Sf, int64_t ldSf,
                        const double* HSf, int64_t ldHSf, double* xa, double* Sa, int64_t ldSa,
                        double* amplitudes, int32_t localise_obs, uint32_t flags) {
  return guarded(c, [&] {
    if (!zones || !obs) arg_fail(c, "zones and obs are required");
    return local_analysis_impl(c, zones, 0, nullptr, nullptr, nullptr, nullptr, obs, N, xf, Hxf,
                               Sf, ldSf, HSf, ldHSf, xa, Sa, ldSa, amplitudes, localise_obs, flags);
  });
}

int oakg_local_analysis_csr(oakg_ctx* c, int64_t nzones, const int64_t* zone_ptr,
                            const int64_t* sel_ptr, const int32_t* sel_idx, const double* sel_w,
                            const oakg_obs* obs, int32_t N, const double* xf, const double* Hxf,
                            const double* Sf, int64_t ldSf, const double* HSf, int64_t ldHSf,
                            double* xa, double* Sa, int64_t ldSa, double* amplitudes,
                            int32_t localise_obs, uint32_t flags) {
  return guarded(c, [&] {
    if (!obs || !sel_ptr || !zone_ptr) arg_fail(c, "obs, zone_ptr and sel_ptr are required");
    return local_analysis_impl(c, nullptr, nzones, zone_ptr, sel_ptr, sel_idx, sel_w, obs, N, xf,
                               Hxf, Sf, ldSf, HSf, ldHSf, xa, Sa, ldSa, amplitudes, localise_obs,
                               flags);
  });
}

int oakg_local_analysis_csr_noalloc(oakg_ctx* c, int64_t nzones, const int64_t* zone_ptr,
                                    const int64_t* sel_ptr, const int32_t* sel_idx,
                                    const double* sel_w, const oakg_obs* obs, int32_t N,
                                    const double* xf, const double* Hxf, const double* Sf,
                                    int64_t ldSf, const double* HSf, int64_t ldHSf, double* xa,
                                    double* Sa, int64_t ldSa, double* amplitudes,
                                    int32_t localise_obs, uint32_t flags) {
  return guarded(c, [&] {
    if (!(flags & OAKG_DEVICE_PTRS))
      arg_fail(c, "oakg_local_analysis_csr_noalloc takes device pointers only (OAKG_DEVICE_PTRS)");
    if (!obs || !sel_ptr || !zone_ptr) arg_fail(c, "obs, zone_ptr and sel_ptr are required");
    if (N < 1 || N > OAKG_MAX_N) arg_fail(c, "N out of range");
    const int NP = padded_n(N);
    if (nzones < 0 || obs->m < 0) arg_fail(c, "negative nzones or m");
    const size_t need = noalloc_fixed_bytes(obs->m, NP, nzones) +
                        64 * noalloc_zone_bytes(NP, c->eig_method == 1) + 8 * 256;
    if (need > kScratchBytes)
      arg_fail(c, "oakg_local_analysis_csr_noalloc: m = " + std::to_string(obs->m) + ", N = " +
                      std::to_string(N) + " and " + std::to_string(nzones) + " zones need " +
                      std::to_string(need >> 20) + " MiB of scratch, above its " +
                      std::to_string(kScratchBytes >> 20) + " MiB");
    check(c, hipSetDevice(c->device), "hipSetDevice");
    NoallocCall guard(c, scratch_on_device(c));
    return local_analysis_impl(c, nullptr, nzones, zone_ptr, sel_ptr, sel_idx, sel_w, obs, N, xf,
                               Hxf, Sf, ldSf, HSf, ldHSf, xa, Sa, ldSa, amplitudes, localise_obs,
                               flags);
  });
}

int oakg_select_observations(oakg_ctx* c, const oakg_zones* zones, const oakg_obs* obs,
                             int32_t localise_obs, int64_t* sel_ptr, int32_t* sel_idx,
                             double* sel_w, int64_t cap, int64_t* total, uint32_t flags) {
  return guarded(c, [&] {
    const bool dev = flags & OAKG_DEVICE_PTRS;
    validate_zones_host(c, zones, 0, true);
    validate_obs(c, obs, true, zones->loctype);
    if (!sel_ptr || !total) arg_fail(c, "sel_ptr and total are required");
    begin_call(c);
    Geo g = stage_geo(c, zones, obs, dev, true);
    SelDev sel{};
    int64_t tot = 0;
    check(c, run_selection(c->ws, g.zd, g.od, localise_obs, c->stream, c->d_err, sel, &tot),
          "selection");
    finish_call(c, zones->nzones, sel.ptr, nullptr);
    const int64_t nz = zones->nzones;
    hipMemcpyKind k = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    check(c, hipMemcpy(sel_ptr, sel.ptr, sizeof(int64_t) * (nz + 1), k), "copy sel_ptr");
    const int64_t ncopy = std::min(tot, cap);
    if (sel_idx && ncopy > 0) check(c, hipMemcpy(sel_idx, sel.idx, sizeof(int32_t) * ncopy, k), "copy idx");
    if (sel_w && ncopy > 0) check(c, hipMemcpy(sel_w, sel.w, sizeof(double) * ncopy, k), "copy w");
    *total = tot;
    return OAKG_OK;
  });
}

int oakg_ensemble_analysis(oakg_ctx* c, const oakg_zones* zones, const oakg_obs* obs,
                           const oakg_csr* H, const double* Hshift, int32_t N, const double* Ef,
                           int64_t ldEf, double* Ea, int64_t ldEa, double inflation,
                           double max_correction, const double* max_correction_elem,
                           uint32_t flags) {
  return ensemble_impl<double>(c, zones, obs, H, Hshift, N, Ef, ldEf, Ea, ldEa, inflation,
                               max_correction, max_correction_elem, flags);
}

int oakg_ensemble_analysis_f32(oakg_ctx* c, const oakg_zones* zones, const oakg_obs* obs,
                               const oakg_csr* H, const float* Hshift, int32_t N, const float* Ef,
                               int64_t ldEf, float* Ea, int64_t ldEa, double inflation,
                               double max_correction, const double* max_correction_elem,
                               uint32_t flags) {
  return ensemble_impl<float>(c, zones, obs, H, Hshift, N, Ef, ldEf, Ea, ldEa, inflation,
                              max_correction, max_correction_elem, flags);
}

int oakg_report(oakg_ctx* c, int64_t m, int32_t N, const double* yo_Hx, const double* HS,
                int64_t ldHS, const double* rmse2, const uint8_t* excluded, int32_t nranges,
                const int64_t* range_ptr, double* stats, uint32_t flags) {
  return guarded(c, [&] {
    return report_impl(c, m, N, yo_Hx, HS, ldHS, rmse2, excluded, nranges, range_ptr, stats,
                       flags);
  });
}

int oakg_ensemble_report(oakg_ctx* c, const oakg_obs* obs, const oakg_csr* H,
                         const double* Hshift, int32_t N, const double* E, int64_t ldE,
                         int32_t nranges, const int64_t* range_ptr, double* stats,
                         uint32_t flags) {
  return ensemble_report_impl<double>(c, obs, H, Hshift, N, E, ldE, nranges, range_ptr, stats,
                                      flags);
}

int oakg_ensemble_report_f32(oakg_ctx* c, const oakg_obs* obs, const oakg_csr* H,
                             const float* Hshift, int32_t N, const float* E, int64_t ldE,
                             int32_t nranges, const int64_t* range_ptr, double* stats,
                             uint32_t flags) {
  return ensemble_report_impl<float>(c, obs, H, Hshift, N, E, ldE, nranges, range_ptr, stats,
                                     flags);
}

int oakg_local_nll(oakg_ctx* c, const oakg_zones* zones, const oakg_obs* obs, int32_t N,
                   const double* Hxf, const double* HSf, int64_t ldHSf, double inflation,
                   double* nll, double* grad_log_sigma, double* grad_corr_len,
                   double* grad_inflation, double* nll_zone, uint32_t flags) {
  return guarded(c, [&] {
    if (!zones || !obs) arg_fail(c, "zones and obs are required");
    return local_nll_impl(c, zones, obs, N, Hxf, HSf, ldHSf, inflation, nll, grad_log_sigma,
                          grad_corr_len, grad_inflation, nll_zone, flags);
  });
}

int oakg_obs_operator(oakg_ctx* c, int32_t nvar, const oakg_grid* grids, int64_t m,
                      const int32_t* obs_var, const double* ox, const double* oy,
                      const double* oz, const double* ot, int64_t* row_ptr, int32_t* col,
                      double* val, int64_t cap, uint8_t* valid, int64_t* nnz, uint32_t flags) {
  return guarded(c, [&] {
    return obs_operator_impl(c, nvar, grids, m, obs_var, ox, oy, oz, ot, row_ptr, col, val, cap,
                             valid, nnz, flags);
  });
}

int oakg_gather_rows(oakg_ctx* c, int64_t n, int32_t N, const int64_t* idx, int64_t full_len,
                     const void* in, int64_t ldi, void* out, int64_t ldo, int32_t elem_size,
                     uint32_t flags) {
  return guarded(c, [&] {
    return permute_rows_impl(c, true, n, N, idx, full_len, in, ldi, out, ldo, elem_size, flags);
  });
}

int oakg_scatter_rows(oakg_ctx* c, int64_t n, int32_t N, const int64_t* idx, int64_t full_len,
                      const void* in, int64_t ldi, void* out, int64_t ldo, int32_t elem_size,
                      uint32_t flags) {
  return guarded(c, [&] {
    return permute_rows_impl(c, false, n, N, idx, full_len, in, ldi, out, ldo, elem_size, flags);
  });
}

int oakg_coo_to_csr(int64_t m, int64_t nnz, const int32_t* row, const int32_t* col,
                    const double* val, int64_t* row_ptr, int32_t* col_out, double* val_out) {
  if (m < 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!row || !col || !val || !col_out || !val_out)))
    return OAKG_EARG;
  for (int64_t i = 0; i <= m; ++i) row_ptr[i] = 0;
  for (int64_t e = 0; e < nnz; ++e) {
    if (row[e] < 0 || row[e] >= m) return OAKG_EARG;
    row_ptr[row[e] + 1]++;
  }
  for (int64_t i = 0; i < m; ++i) row_ptr[i + 1] += row_ptr[i];
  std::vector<int64_t> pos(row_ptr, row_ptr + m);
  for (int64_t e = 0; e < nnz; ++e) {
    int64_t p = pos[row[e]]++;
    col_out[p] = col[e];
    val_out[p] = val[e];
  }
  return OAKG_OK;
}

int64_t oakg_partition(int64_t n, const int64_t* partition, int64_t* zone_size,
                       int64_t* zone_index, int64_t* inv_zone_index) {
  if (n < 0 || (n > 0 && (!partition || !zone_size || !zone_index || !inv_zone_index))) return -1;
  std::vector<int64_t> u(partition, partition + n);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  const int64_t nz = (int64_t)u.size();
  std::vector<int64_t> lab(n);
  for (int64_t i = 0; i < n; ++i)
    lab[i] = std::lower_bound(u.begin(), u.end(), partition[i]) - u.begin();
  for (int64_t z = 0; z < nz; ++z) zone_size[z] = 0;
  for (int64_t i = 0; i < n; ++i) zone_size[lab[i]]++;
  std::vector<int64_t> ci(nz);
  int64_t acc = 0;
  for (int64_t z = 0; z < nz; ++z) { ci[z] = acc; acc += zone_size[z]; }
  for (int64_t i = 0; i < n; ++i) {
    zone_index[ci[lab[i]]] = i;
    inv_zone_index[i] = ci[lab[i]];
    ci[lab[i]]++;
  }
  return nz;
}

}  // extern "C"
