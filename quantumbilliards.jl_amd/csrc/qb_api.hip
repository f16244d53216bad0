// C ABI of libqbhip (declared in include/qb.h): contexts, resident geometry,
// construct_boundary_matrices!, and the Beyn window (assembly + LU of every
// contour node, then the moment solves).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "qb_internal.h"
#include "qb_linalg.h"

// ---------------------------------------------------------------------------
// errors and allocation
// ---------------------------------------------------------------------------
int qb_fail(qb_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

// Device memory of a context.  Freed blocks stay in a per-context cache keyed
// by exact size (windows of one size allocate the same list of buffers), up
// to cache_limit bytes; an allocation that fails releases the cache and
// retries once.  `bytes` counts memory handed out plus cached.
// Device memory is allocated and released in stream order on the context's
// stream (hipMallocAsync / hipFreeAsync): hipFree waits for every kernel on
// the device, so with several contexts on one GPU (the serving loop's slots)
// a temporary freed by one context would wait for the other's long LU.
static void cache_release(qb_ctx* ctx) {
  for (auto& e : ctx->cache) {
    (void)hipFreeAsync(e.second, ctx->stream);
    ctx->bytes -= (int64_t)e.first;
  }
  ctx->cache.clear();
  ctx->cache_bytes = 0;
}

int qb_dev_alloc(qb_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return QB_OK;
  for (size_t i = 0; i < ctx->cache.size(); ++i) {
    if (ctx->cache[i].first == bytes) {
      *p = ctx->cache[i].second;
      ctx->cache_bytes -= (int64_t)bytes;
      ctx->cache.erase(ctx->cache.begin() + i);
      return QB_OK;
    }
  }
  hipError_t e = hipMallocAsync(p, bytes, ctx->stream);
  if (e != hipSuccess && !ctx->cache.empty()) {
    (void)hipGetLastError();
    cache_release(ctx);
    // the released blocks return to the pool once this stream reaches them
    (void)hipStreamSynchronize(ctx->stream);
    e = hipMallocAsync(p, bytes, ctx->stream);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return qb_fail(ctx, QB_E_NOMEM,
                   "hipMallocAsync of " + std::to_string(bytes) + " bytes failed: " +
                       hipGetErrorString(e));
  }
  ctx->bytes += (int64_t)bytes;
  return QB_OK;
}

void qb_dev_free(qb_ctx* ctx, void* p, size_t bytes) {
  if (!p) return;
  if (bytes > 0 && ctx->cache_bytes + (int64_t)bytes <= ctx->cache_limit) {
    // the block may still be in use by queued work on this context's
    // streams; reuse only happens through this context, whose later
    // launches are ordered after that work
    ctx->cache.push_back({bytes, p});
    ctx->cache_bytes += (int64_t)bytes;
    return;
  }
  (void)hipFreeAsync(p, ctx->stream);
  ctx->bytes -= (int64_t)bytes;
}

// ---------------------------------------------------------------------------
// launch accounting
// ---------------------------------------------------------------------------
static void prof_reduce(qb_ctx* ctx) {
  if (ctx->prof_recs.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream_aux) (void)hipStreamSynchronize(ctx->stream_aux);
  for (auto& r : ctx->prof_recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      ctx->prof_ms[r.kind] += ms;
      ctx->prof_n[r.kind] += 1;
      ctx->prof_flops[r.kind] += r.flops;
      ctx->prof_bytes[r.kind] += r.bytes;
    }
    ctx->prof_pool.push_back(r.a);
    ctx->prof_pool.push_back(r.b);
  }
  ctx->prof_recs.clear();
}

static hipEvent_t prof_event(qb_ctx* ctx) {
  if (!ctx->prof_pool.empty()) {
    hipEvent_t e = ctx->prof_pool.back();
    ctx->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

int qb_prof_begin(qb_ctx* ctx, int kind, double flops, double bytes) {
  if (!ctx->prof) return -1;
  if (ctx->prof_recs.size() >= 8192) prof_reduce(ctx);
  QbProfRec r{prof_event(ctx), prof_event(ctx), kind, flops, bytes};
  (void)hipEventRecord(r.a, ctx->stream);
  ctx->prof_recs.push_back(r);
  return (int)ctx->prof_recs.size() - 1;
}

void qb_prof_end(qb_ctx* ctx, int slot) {
  if (slot < 0) return;
  (void)hipEventRecord(ctx->prof_recs[slot].b, ctx->stream);
}

// ---------------------------------------------------------------------------
// device array of qb_assemble_noalloc
// ---------------------------------------------------------------------------
// The Chebyshev tables of a no-allocation assembly live in this fixed array;
// every GPU the library runs on holds its own copy until the process ends.
// Host code never passes the array's name to a kernel (in host code it is a
// stand-in, not a device address): qb_ctx_create looks its device address up
// once per device (hipGetSymbolAddress, with the device current) and every
// use goes through that address.
static __device__ __align__(256) unsigned char scratch[QB_NOALLOC_SCRATCH_BYTES];
#define QB_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[QB_MAX_DEVICES] = {nullptr};
static hipError_t g_scratch_err[QB_MAX_DEVICES] = {hipSuccess};

// With `device` current: the address of `scratch` on it, looked up once.
static void scratch_setup(int device) {
  if (device < 0 || device >= QB_MAX_DEVICES) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess) g_scratch_ptr[device] = p;
}

static double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

extern "C" {

int qb_ctx_create(int device, qb_ctx** out) {
  if (!out) return QB_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return QB_E_HIP;
  if (device < 0 || device >= n) return QB_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return QB_E_HIP;
  qb_ctx* c = new qb_ctx();
  c->device = device;
  for (int t = 0; t < 2; ++t)
    for (int k = 0; k < QB_GEMM_CLASSES; ++k) c->gemm_pick[t][k] = c->gemm_pick_conc[t][k] = -1;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream_aux, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_main, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_aux, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_first, hipEventDisableTiming) != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream_aux) (void)hipStreamDestroy(c->stream_aux);
    delete c;
    return QB_E_HIP;
  }
  scratch_setup(device);
  *out = c;
  return QB_OK;
}

void qb_ctx_destroy(qb_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  prof_reduce(ctx);
  for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
  if (ctx->stream_aux) {
    (void)hipStreamSynchronize(ctx->stream_aux);
    (void)hipStreamDestroy(ctx->stream_aux);
  }
  (void)hipStreamSynchronize(ctx->stream);
  qb_graphs_clear(ctx);
  cache_release(ctx);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->ev_main) (void)hipEventDestroy(ctx->ev_main);
  if (ctx->ev_aux) (void)hipEventDestroy(ctx->ev_aux);
  if (ctx->ev_first) (void)hipEventDestroy(ctx->ev_first);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* qb_last_error(const qb_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t qb_ctx_device_bytes(const qb_ctx* ctx) { return ctx ? ctx->bytes : 0; }

int qb_ctx_set_graphs(qb_ctx* ctx, int32_t on) {
  if (!ctx) return QB_E_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->graphs = on != 0;
  if (!ctx->graphs) qb_graphs_clear(ctx);
  return QB_OK;
}

int qb_ctx_mem_info(qb_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes) {
  if (!ctx || !free_bytes || !total_bytes) return QB_E_ARG;
  (void)hipSetDevice(ctx->device);
  size_t f = 0, t = 0;
  hipError_t e = hipMemGetInfo(&f, &t);
  if (e != hipSuccess) return qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  *free_bytes = (int64_t)f;
  *total_bytes = (int64_t)t;
  return QB_OK;
}

int qb_host_alloc(qb_ctx* ctx, int64_t bytes, void** out) {
  if (!ctx) return QB_E_ARG;
  if (!out || bytes < 0) return qb_fail(ctx, QB_E_ARG, "null output or negative size");
  *out = nullptr;
  if (bytes == 0) return QB_OK;
  (void)hipSetDevice(ctx->device);
  hipError_t e = hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *out = nullptr;
    return qb_fail(ctx, QB_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  return QB_OK;
}

void qb_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int qb_ctx_set_cache_limit(qb_ctx* ctx, int64_t bytes) {
  if (!ctx || bytes < 0) return QB_E_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->cache_limit = bytes;
  if (ctx->cache_bytes > bytes) cache_release(ctx);
  return QB_OK;
}

int qb_ctx_set_profiling(qb_ctx* ctx, int32_t enable) {
  if (!ctx) return QB_E_ARG;
  (void)hipSetDevice(ctx->device);
  prof_reduce(ctx);
  for (int k = 0; k < QB_K_COUNT; ++k) {
    ctx->prof_ms[k] = 0;
    ctx->prof_n[k] = 0;
    ctx->prof_flops[k] = 0;
    ctx->prof_bytes[k] = 0;
    ctx->prof_issued[k] = 0;
  }
  ctx->prof = enable != 0;
  return QB_OK;
}

int qb_ctx_kernel_stats(qb_ctx* ctx, int32_t kind, double* total_ms, int64_t* launches,
                        double* flops, double* bytes) {
  if (!ctx) return QB_E_ARG;
  if (kind < 0 || kind >= QB_K_COUNT) return qb_fail(ctx, QB_E_ARG, "unknown kernel class");
  (void)hipSetDevice(ctx->device);
  prof_reduce(ctx);
  if (total_ms) *total_ms = ctx->prof_ms[kind];
  if (launches) *launches = ctx->prof_n[kind];
  if (flops) *flops = ctx->prof_flops[kind];
  if (bytes) *bytes = ctx->prof_bytes[kind];
  return QB_OK;
}

int qb_ctx_kernel_issued(qb_ctx* ctx, int32_t kind, double* issued) {
  if (!ctx) return QB_E_ARG;
  if (kind < 0 || kind >= QB_K_COUNT) return qb_fail(ctx, QB_E_ARG, "unknown kernel class");
  if (issued) *issued = ctx->prof_issued[kind];
  return QB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
// Kress R first column and log term per component: one thread per (comp, d).
// Even N_c = 2n: R_d = -(4 pi / N_c) sum_{m=1}^{n-1} cos(2 pi m d / N_c) / m
//                      - (4 pi / N_c^2) (-1)^d;
// odd N_c = 2n+1: no Nyquist term, m = 1..n (boundarypoints.jl:310-344).
// lt_d = log(4 sin^2(pi d / N_c)), lt_0 = 0 (boundarypoints.jl:878-879).
__global__ void k_kress_cols(int64_t N, int ncomp, DevGeom g, double* rk, double* lt) {
  int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= N) return;
  int c = 0;
  while (c + 1 < ncomp && gi >= g.off[c + 1]) ++c;
  int64_t o = g.off[c], nc = g.off[c + 1] - o;
  int64_t d = gi - o;
  int64_t mmax = (nc % 2 == 0) ? nc / 2 - 1 : (nc - 1) / 2;
  double acc = 0.0;
  for (int64_t m = 1; m <= mmax; ++m) {
    int64_t md = (m * d) % nc;
    acc += cospi(2.0 * (double)md / (double)nc) / (double)m;
  }
  const double two_pi = 6.283185307179586476925;
  double R = -(2.0 * two_pi / (double)nc) * acc;
  if (nc % 2 == 0) R -= (2.0 * two_pi / ((double)nc * (double)nc)) * ((d & 1) ? -1.0 : 1.0);
  rk[gi] = R;
  if (d == 0) {
    lt[gi] = 0.0;
  } else {
    double s = sinpi((double)d / (double)nc);
    lt[gi] = log(4.0 * s * s);
  }
}

// Padded range of pairwise distances (cheb-CFIE:374-418, cheb-DLP:135-147):
// same-component pairs skip i == j, all pairs skip r <= eps.  A block takes
// 256 targets i against the sources j = blockIdx.y, blockIdx.y + gridDim.y,
// ..., reduces in registers and LDS, and issues one atomic pair, so at most
// 256 x ceil(N/256) atomics meet on the two addresses (an atomic pair per
// wavefront and source column would be N^2/64 of them, serialised).
// Non-negative doubles order like their bit patterns, so the unsigned 64-bit
// atomics take the min and max of the distances.
__global__ __launch_bounds__(256) void k_rrange(DevGeom g, int only_first,
                                                unsigned long long* mm) {
  const int64_t N = only_first ? g.off[1] : g.N;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rmn = 1e300, rmx = 0.0;
  if (i < N) {
    const double xi = g.x[i], yi = g.y[i];
    for (int64_t j = blockIdx.y; j < N; j += gridDim.y) {
      if (j == i) continue;
      const double r = hypot(xi - g.x[j], yi - g.y[j]);
      if (r > 2.220446049250313e-16) {
        rmn = fmin(rmn, r);
        rmx = fmax(rmx, r);
      }
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    rmn = fmin(rmn, __shfl_xor(rmn, off));
    rmx = fmax(rmx, __shfl_xor(rmx, off));
  }
  __shared__ double smn[4], smx[4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smn[w] = rmn;
    smx[w] = rmx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; ++q) {
      rmn = fmin(rmn, smn[q]);
      rmx = fmax(rmx, smx[q]);
    }
    atomicMin(&mm[0], (unsigned long long)__double_as_longlong(rmn));
    atomicMax(&mm[1], (unsigned long long)__double_as_longlong(rmx));
  }
}

static double bits2d(unsigned long long b) {
  double v;
  memcpy(&v, &b, sizeof(v));
  return v;
}

static int rrange(qb_ctx* ctx, const DevGeom& d, int only_first, double* rmin, double* rmax) {
  unsigned long long* mm = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&mm, 2 * sizeof(unsigned long long));
  if (rc) return rc;
  double big = 1e300;
  unsigned long long init[2];
  memcpy(&init[0], &big, 8);
  init[1] = 0ull;
  QB_HIP(ctx, hipMemcpyAsync(mm, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
  int64_t N = only_first ? d.off[1] : d.N;
  // k_rrange reduces over exactly 4 wavefronts of 256 threads per block
  dim3 grid((unsigned)((N + 255) / 256), (unsigned)std::min<int64_t>(N, 256));
  hipLaunchKernelGGL(k_rrange, grid, dim3(256), 0, ctx->stream, d, only_first, mm);
  QB_HIP(ctx, hipGetLastError());
  unsigned long long h[2];
  QB_HIP(ctx, hipMemcpyAsync(h, mm, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  QB_HIP(ctx, hipStreamSynchronize(ctx->stream));
  qb_dev_free(ctx, mm, 2 * sizeof(unsigned long long));
  *rmin = 0.95 * bits2d(h[0]);
  *rmax = 1.05 * bits2d(h[1]);
  return QB_OK;
}

extern "C" int qb_geom_create(qb_ctx* ctx, int32_t ncomp, const int64_t* n_per_comp,
                              const double* x, const double* y, const double* dx, const double* dy,
                              const double* ddx, const double* ddy, const double* ts,
                              const double* ws, qb_geom** out) {
  if (!ctx || !out) return QB_E_ARG;
  *out = nullptr;
  if (ncomp < 1 || ncomp > QB_MAX_COMP)
    return qb_fail(ctx, QB_E_ARG, "ncomp must be in [1, 8]");
  if (!n_per_comp || !x || !y || !dx || !dy || !ddx || !ddy || !ts || !ws)
    return qb_fail(ctx, QB_E_ARG, "null geometry array");
  int64_t N = 0;
  for (int c = 0; c < ncomp; ++c) {
    if (n_per_comp[c] < 4) return qb_fail(ctx, QB_E_ARG, "each component needs >= 4 nodes");
    N += n_per_comp[c];
  }
  if (N > 65535) return qb_fail(ctx, QB_E_ARG, "N must be <= 65535");
  // uniform Kress grid per component (ts spaced +-2pi/N_c)
  int64_t o = 0;
  for (int c = 0; c < ncomp; ++c) {
    int64_t nc = n_per_comp[c];
    double h = 2.0 * M_PI / (double)nc;
    for (int64_t i = 0; i + 1 < nc; ++i) {
      double d = fabs(ts[o + i + 1] - ts[o + i]);
      if (fabs(d - h) > 1e-9 * (1.0 + h))
        return qb_fail(ctx, QB_E_ARG,
                       "component " + std::to_string(c) +
                           ": ts is not the uniform Kress grid (graded/corner variants are not "
                           "supported)");
    }
    o += nc;
  }
  (void)hipSetDevice(ctx->device);
  qb_geom* g = new qb_geom();
  g->ctx = ctx;
  // host-side per-node quantities (O(N))
  std::vector<double> speed(N), kappa(N);
  std::vector<int32_t> comp(N);
  o = 0;
  for (int c = 0; c < ncomp; ++c) {
    for (int64_t i = 0; i < n_per_comp[c]; ++i) comp[o + i] = c;
    o += n_per_comp[c];
  }
  for (int64_t i = 0; i < N; ++i) {
    speed[i] = sqrt(dx[i] * dx[i] + dy[i] * dy[i]);
    double num = -(dx[i] * ddy[i] - dy[i] * ddx[i]);
    double den = dx[i] * dx[i] + dy[i] * dy[i];
    kappa[i] = (1.0 / (2.0 * M_PI)) * (num / den);
  }
  // one allocation: 9 double arrays + int32 comp
  size_t nd = 9 * (size_t)N;
  size_t bytes = nd * sizeof(double) + (size_t)N * sizeof(int32_t);
  int rc = qb_dev_alloc(ctx, &g->buf, bytes);
  if (rc) {
    delete g;
    return rc;
  }
  g->bytes = (int64_t)bytes;
  double* b = (double*)g->buf;
  double* dX = b;
  double* dY = b + N;
  double* dDX = b + 2 * N;
  double* dDY = b + 3 * N;
  double* dSp = b + 4 * N;
  double* dW = b + 5 * N;
  double* dK = b + 6 * N;
  double* dRk = b + 7 * N;
  double* dLt = b + 8 * N;
  int32_t* dC = (int32_t*)(b + 9 * N);
  auto up = [&](double* d, const double* h) {
    return hipMemcpyAsync(d, h, N * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  };
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = up(dX, x);
  if (e == hipSuccess) e = up(dY, y);
  if (e == hipSuccess) e = up(dDX, dx);
  if (e == hipSuccess) e = up(dDY, dy);
  if (e == hipSuccess) e = up(dSp, speed.data());
  if (e == hipSuccess) e = up(dW, ws);
  if (e == hipSuccess) e = up(dK, kappa.data());
  if (e == hipSuccess)
    e = hipMemcpyAsync(dC, comp.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) {
    qb_geom_destroy(g);
    return qb_fail(ctx, QB_E_HIP, std::string("geometry upload: ") + hipGetErrorString(e));
  }
  DevGeom& d = g->d;
  d.N = N;
  d.ncomp = ncomp;
  d.off[0] = 0;
  for (int c = 0; c < ncomp; ++c) d.off[c + 1] = d.off[c] + n_per_comp[c];
  for (int c = ncomp + 1; c <= QB_MAX_COMP; ++c) d.off[c] = N;
  d.x = dX;
  d.y = dY;
  d.dx = dDX;
  d.dy = dDY;
  d.speed = dSp;
  d.ws = dW;
  d.kappa = dK;
  d.comp = dC;
  d.rk = dRk;
  d.lt = dLt;
  hipLaunchKernelGGL(k_kress_cols, dim3((unsigned)((N + 127) / 128)), dim3(128), 0, ctx->stream,
                     N, ncomp, d, dRk, dLt);
  e = hipGetLastError();
  if (e != hipSuccess) {
    qb_geom_destroy(g);
    return qb_fail(ctx, QB_E_HIP, std::string("kress columns: ") + hipGetErrorString(e));
  }
  rc = rrange(ctx, d, 0, &g->rmin_geom, &g->rmax_geom);
  if (!rc) rc = rrange(ctx, d, 1, &g->rmin_dlp, &g->rmax_dlp);
  if (rc) {
    qb_geom_destroy(g);
    return rc;
  }
  *out = g;
  return QB_OK;
}

extern "C" void qb_geom_destroy(qb_geom* g) {
  if (!g) return;
  qb_dev_free(g->ctx, g->sym_buf, (size_t)g->sym_bytes);
  qb_dev_free(g->ctx, g->buf, (size_t)g->bytes);
  delete g;
}

extern "C" int64_t qb_geom_size(const qb_geom* g) { return g ? g->d.N : 0; }

extern "C" int64_t qb_geom_dim(const qb_geom* g) { return g ? qb_dim(g) : 0; }

// Orbit tables of a symmetry reduction.  The checks are those the reference
// makes while building its orbit maps (_build_periodic_orbit_maps,
// symmetry.jl:300-323: no collision, every node covered) plus the ones its
// assembly relies on: the representative is image 0 with weight 1, and no
// orbit meets a node twice (a node on a reflection axis,
// _has_node_on_reflection_axis, symmetry.jl:326-345).
extern "C" int32_t qb_geom_irreps(const qb_geom* g) { return g ? qb_nirr(g) : 0; }

extern "C" int qb_geom_set_orbits(qb_ctx* ctx, qb_geom* g, int64_t m, int32_t G,
                                  const int64_t* img, const double* scale) {
  return qb_geom_set_orbits_multi(ctx, g, m, G, 1, img, scale);
}

extern "C" int qb_geom_set_orbits_multi(qb_ctx* ctx, qb_geom* g, int64_t m, int32_t G,
                                        int32_t nirr, const int64_t* img, const double* scale) {
  if (!ctx) return QB_E_ARG;
  if (!g) return qb_fail(ctx, QB_E_ARG, "null geometry");
  if (g->ctx != ctx) return qb_fail(ctx, QB_E_ARG, "geometry belongs to another context");
  (void)hipSetDevice(ctx->device);
  if (m == 0) {
    qb_dev_free(ctx, g->sym_buf, (size_t)g->sym_bytes);
    g->sym_buf = nullptr;
    g->sym_bytes = 0;
    g->sym = DevSym{};
    return QB_OK;
  }
  if (!img || !scale) return qb_fail(ctx, QB_E_ARG, "null orbit table");
  if (nirr < 1 || nirr > QB_MAX_IRREPS)
    return qb_fail(ctx, QB_E_ARG, "nirr must be in [1, 16]");
  const int64_t N = g->d.N;
  if (G < 2 || G > 64) return qb_fail(ctx, QB_E_ARG, "orbit size G must be in [2, 64]");
  if (m < 1 || m * (int64_t)G != N)
    return qb_fail(ctx, QB_E_ARG, "orbits must partition the nodes: m * G != N");
  std::vector<int32_t> im((size_t)m * G);
  std::vector<char> seen((size_t)N, 0);
  for (int64_t e = 0; e < nirr; ++e)
    for (int64_t b = 0; b < m; ++b) {
      const double* sc = scale + 2 * (e * m + b) * G;
      if (sc[0] != 1.0 || sc[1] != 0.0)
        return qb_fail(ctx, QB_E_ARG, "orbit " + std::to_string(b) +
                                          ": the representative (image 0) must have weight 1");
    }
  for (int64_t b = 0; b < m; ++b) {
    for (int l = 0; l < G; ++l) {
      const int64_t q = img[b * G + l];
      if (q < 0 || q >= N)
        return qb_fail(ctx, QB_E_ARG, "orbit " + std::to_string(b) + ": node index out of range");
      if (seen[q])
        return qb_fail(ctx, QB_E_ARG,
                       "node " + std::to_string(q) +
                           " appears twice in the orbits (collision, or a node on a symmetry axis)");
      seen[q] = 1;
      im[b * G + l] = (int32_t)q;
    }
  }
  const size_t bi = im.size() * sizeof(int32_t), bs = (size_t)nirr * im.size() * sizeof(cd);
  const size_t bytes = bs + ((bi + 255) & ~(size_t)255);
  void* buf = nullptr;
  int rc = qb_dev_alloc(ctx, &buf, bytes);
  if (rc) return rc;
  cd* ds = (cd*)buf;
  int32_t* di = (int32_t*)((char*)buf + bs);
  hipError_t e = hipMemcpyAsync(ds, scale, bs, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(di, im.data(), bi, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    qb_dev_free(ctx, buf, bytes);
    return qb_fail(ctx, QB_E_HIP, std::string("orbit upload: ") + hipGetErrorString(e));
  }
  DevSym sym;
  sym.m = m;
  sym.G = G;
  sym.nirr = nirr;
  sym.img = di;
  sym.scale = ds;
  // Pair sharing (k_assemble_red_pair): find, for each image l, the image
  // inv[l] whose distances mirror those of l on a few sample pairs, then
  // verify the identity on every (a, b, l) on the device.
  {
    std::vector<double> hx(N), hy(N);
    e = hipMemcpyAsync(hx.data(), g->d.x, N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hy.data(), g->d.y, N * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      qb_dev_free(ctx, buf, bytes);
      return qb_fail(ctx, QB_E_HIP, std::string("orbit check: ") + hipGetErrorString(e));
    }
    auto dist = [&](int64_t p, int64_t q) { return hypot(hx[p] - hx[q], hy[p] - hy[q]); };
    double scale = 0.0;
    for (int64_t q = 0; q < N; ++q) scale = std::max(scale, std::max(fabs(hx[q]), fabs(hy[q])));
    // a few ulps of the coordinate scale: the rounding of any node distance
    const double floor_abs = 16.0 * 2.220446049250313e-16 * scale;
    const int64_t samp[4] = {0, m / 3, (2 * m) / 3, m - 1};
    bool found_all = true;
    for (int l = 0; l < G && found_all; ++l) {
      int pick = -1;
      for (int c = 0; c < G && pick < 0; ++c) {
        bool okc = true;
        for (int sa = 0; sa < 4 && okc; ++sa)
          for (int sb = 0; sb < 4 && okc; ++sb) {
            const int64_t a = samp[sa], b = samp[sb];
            const double r1 = dist(im[a * G], im[b * G + l]);
            const double r2 = dist(im[b * G], im[a * G + c]);
            okc = fabs(r1 - r2) <= 1e-13 * r1 + floor_abs;
          }
        if (okc) pick = c;
      }
      if (pick < 0) found_all = false;
      else sym.inv[l] = (int8_t)pick;
    }
    // the mirror entry sums over l' = inv[l] for every l: inv must be a
    // permutation of the images
    if (found_all) {
      std::vector<char> hit(G, 0);
      for (int l = 0; l < G; ++l) hit[sym.inv[l]] = 1;
      for (int l = 0; l < G; ++l) found_all = found_all && hit[l];
    }
    int ok = 0;
    if (found_all) {
      rc = qb_orbit_pairs_ok(ctx, g->d, sym, floor_abs, &ok);
      if (rc) {
        qb_dev_free(ctx, buf, bytes);
        return rc;
      }
    }
    sym.pair = ok != 0 && getenv("QB_SYM_NO_PAIR") == nullptr;
    if (getenv("QB_SYM_TRACE")) {
      fprintf(stderr, "[qb] orbits m=%lld G=%d pair=%d inv=", (long long)m, G, (int)sym.pair);
      for (int l = 0; l < G; ++l) fprintf(stderr, "%d%s", found_all ? sym.inv[l] : -1, l + 1 < G ? "," : "\n");
    }
  }
  qb_dev_free(ctx, g->sym_buf, (size_t)g->sym_bytes);
  g->sym_buf = buf;
  g->sym_bytes = (int64_t)bytes;
  g->sym = sym;
  return QB_OK;
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
static void op_range(const qb_geom* g, int op, double* rmin, double* rmax) {
  if (op == QB_DLP_KRESS) {
    *rmin = g->rmin_dlp;
    *rmax = g->rmax_dlp;
  } else {
    *rmin = g->rmin_geom;
    *rmax = g->rmax_geom;
  }
}

static int check_cheb(qb_ctx* ctx, const qb_cheb_params* c) {
  if (!c) return qb_fail(ctx, QB_E_ARG, "null Chebyshev parameters");
  if (c->n_panels_h < 1 || c->n_panels_j < 1 || c->M_h < 2 || c->M_j < 2 || c->M_h > 16 ||
      c->M_j > 16)
    return qb_fail(ctx, QB_E_ARG, "Chebyshev parameters out of range (panels >= 1, 2 <= M <= 16)");
  return QB_OK;
}

extern "C" int qb_assemble(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* z, int64_t nz,
                           const qb_cheb_params* cheb, double* out, int32_t flags,
                           double* rminmax) {
  if (!ctx) return QB_E_ARG;
  if (!g || !z || !out || nz < 0) return qb_fail(ctx, QB_E_ARG, "null argument");
  if (g->ctx != ctx) return qb_fail(ctx, QB_E_ARG, "geometry belongs to another context");
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS) return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (op == QB_DLP_KRESS && g->d.ncomp != 1)
    return qb_fail(ctx, QB_E_ARG, "DLP_kress expects one boundary component");
  int rc = check_cheb(ctx, cheb);
  if (rc) return rc;
  if (nz == 0) return QB_OK;
  (void)hipSetDevice(ctx->device);
  double rmin, rmax;
  op_range(g, op, &rmin, &rmax);
  if (rminmax) {
    rminmax[0] = rmin;
    rminmax[1] = rmax;
  }
  int64_t N = qb_dim(g);
  cd* zd = nullptr;
  rc = qb_dev_alloc(ctx, (void**)&zd, nz * sizeof(cd));
  if (rc) return rc;
  DevTables t;
  cd* od = nullptr;
  size_t obytes = (size_t)nz * qb_nirr(g) * N * N * sizeof(cd);
  bool on_dev = (flags & QB_OUT_DEVICE) != 0;
  hipError_t e = hipMemcpyAsync(zd, z, nz * sizeof(cd), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) rc = qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  // the conjugate-pair kernel of the Beyn windows reads envelope tables of
  // its own (QB_ASM_TABLES=1: the uniform Hankel tables, as the windows then)
  const bool pairs = (flags & QB_ASM_CONJ_PAIRS) != 0;
  const bool env = pairs && getenv("QB_ASM_TABLES") == nullptr;
  if (!rc && !env) rc = qb_tables_build(ctx, (const double*)zd, nz, rmin, rmax, *cheb, true, &t);
  if (!rc) {
    if (on_dev)
      od = (cd*)out;
    else
      rc = qb_dev_alloc(ctx, (void**)&od, obytes);
  }
  if (!rc) {
    if (pairs) {
      rc = env ? qb_assemble_env_dev(ctx, g, op, (const cd*)z, zd, nz, rmin, rmax, od, nullptr)
               : qb_assemble_conj_dev(ctx, g, op, (const cd*)z, zd, nz, t, od, nullptr);
      if (rc == 1)
        rc = qb_fail(ctx, QB_E_ARG,
                     "QB_ASM_CONJ_PAIRS needs nodes closed under conjugation and no symmetry "
                     "orbits (with QB_ASM_TABLES: tables with M_h = M_j in {5, 7})");
    } else {
      rc = qb_assemble_dev(ctx, g, op, zd, nz, t, od, nullptr);
    }
  }
  if (!rc && !on_dev) {
    e = hipMemcpyAsync(out, od, obytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) rc = qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  }
  if (!rc) {
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  } else {
    (void)hipStreamSynchronize(ctx->stream);
  }
  if (!on_dev && od) qb_dev_free(ctx, od, obytes);
  qb_tables_free(ctx, &t);
  qb_dev_free(ctx, zd, nz * sizeof(cd));
  return rc;
}

extern "C" int qb_assemble_noalloc(qb_ctx* ctx, const qb_geom* g, int32_t op,
                                   const double* z_dev, int64_t nz, const qb_cheb_params* cheb,
                                   double* out_dev, int32_t flags, double* rminmax) {
  if (!ctx) return QB_E_ARG;
  if (!g || !z_dev || !out_dev || nz < 0) return qb_fail(ctx, QB_E_ARG, "null argument");
  if (g->ctx != ctx) return qb_fail(ctx, QB_E_ARG, "geometry belongs to another context");
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS) return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (op == QB_DLP_KRESS && g->d.ncomp != 1)
    return qb_fail(ctx, QB_E_ARG, "DLP_kress expects one boundary component");
  if (flags != 0)
    return qb_fail(ctx, QB_E_ARG,
                   "qb_assemble_noalloc takes no flags (device in and out, J tables)");
  int rc = check_cheb(ctx, cheb);
  if (rc) return rc;
  const size_t need = qb_tables_bytes(nz, *cheb, nullptr);
  if (need > (size_t)QB_NOALLOC_SCRATCH_BYTES)
    return qb_fail(ctx, QB_E_ARG,
                   "qb_assemble_noalloc: the tables of " + std::to_string(nz) +
                       " wavenumbers need " + std::to_string(need) +
                       " bytes, more than the fixed device array of " +
                       std::to_string((long long)QB_NOALLOC_SCRATCH_BYTES) +
                       " bytes; assemble fewer wavenumbers per call");
  if (ctx->device < 0 || ctx->device >= QB_MAX_DEVICES)
    return qb_fail(ctx, QB_E_ARG, "device index out of range for the no-allocation array");
  void* mem = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    mem = g_scratch_ptr[ctx->device];
    if (!mem)
      return qb_fail(ctx, QB_E_HIP,
                     std::string("no device address for the no-allocation array: ") +
                         hipGetErrorString(g_scratch_err[ctx->device]));
  }
  double rmin, rmax;
  op_range(g, op, &rmin, &rmax);
  if (rminmax) {
    rminmax[0] = rmin;
    rminmax[1] = rmax;
  }
  if (nz == 0) return QB_OK;
  (void)hipSetDevice(ctx->device);
  DevTables t;
  rc = qb_tables_build_in(ctx, z_dev, nz, rmin, rmax, *cheb, mem,
                          (size_t)QB_NOALLOC_SCRATCH_BYTES, &t);
  if (!rc) rc = qb_assemble_dev(ctx, g, op, (const cd*)z_dev, nz, t, (cd*)out_dev, nullptr);
  // complete on return, as qb_assemble is
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (!rc && e != hipSuccess) rc = qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  return rc;
}

extern "C" int qb_assemble_deriv(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* z,
                                 int64_t nz, const qb_cheb_params* cheb, double* out,
                                 double* dout, double* ddout, int32_t flags, double* rminmax) {
  if (!ctx) return QB_E_ARG;
  if (!g || !z || !out || !dout || !ddout || nz < 0) return qb_fail(ctx, QB_E_ARG, "null argument");
  if (g->ctx != ctx) return qb_fail(ctx, QB_E_ARG, "geometry belongs to another context");
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS) return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (op == QB_DLP_KRESS && g->d.ncomp != 1)
    return qb_fail(ctx, QB_E_ARG, "DLP_kress expects one boundary component");
  int rc = check_cheb(ctx, cheb);
  if (rc) return rc;
  if ((flags & QB_DERIV_REFERENCE_SIGN) && (flags & QB_DERIV_MATRIX_SIGN))
    return qb_fail(ctx, QB_E_ARG,
                   "QB_DERIV_REFERENCE_SIGN and QB_DERIV_MATRIX_SIGN exclude each other");
  if (nz == 0) return QB_OK;
  (void)hipSetDevice(ctx->device);
  double rmin, rmax;
  op_range(g, op, &rmin, &rmax);
  if (rminmax) {
    rminmax[0] = rmin;
    rminmax[1] = rmax;
  }
  const int64_t N = qb_dim(g);
  const size_t obytes = (size_t)nz * N * N * sizeof(cd);
  const bool on_dev = (flags & QB_OUT_DEVICE) != 0;
  cd* zd = nullptr;
  cd* od[3] = {nullptr, nullptr, nullptr};
  double* host[3] = {out, dout, ddout};
  DevTables t;
  rc = qb_dev_alloc(ctx, (void**)&zd, nz * sizeof(cd));
  if (!rc && hipMemcpyAsync(zd, z, nz * sizeof(cd), hipMemcpyHostToDevice, ctx->stream) !=
                 hipSuccess)
    rc = qb_fail(ctx, QB_E_HIP, "upload failed");
  if (!rc) rc = qb_tables_build(ctx, (const double*)zd, nz, rmin, rmax, *cheb, true, &t);
  for (int m = 0; m < 3 && !rc; ++m) {
    if (on_dev)
      od[m] = (cd*)host[m];
    else
      rc = qb_dev_alloc(ctx, (void**)&od[m], obytes);
  }
  if (!rc)
    rc = qb_assemble_deriv_dev(ctx, g, op, zd, nz, t, od[0], od[1], od[2],
                               (flags & QB_DERIV_MATRIX_SIGN) == 0);
  for (int m = 0; m < 3 && !rc && !on_dev; ++m)
    if (hipMemcpyAsync(host[m], od[m], obytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
      rc = qb_fail(ctx, QB_E_HIP, "download failed");
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc)
    rc = qb_fail(ctx, QB_E_HIP, "assembly failed");
  if (!on_dev)
    for (int m = 0; m < 3; ++m) qb_dev_free(ctx, od[m], obytes);
  qb_tables_free(ctx, &t);
  qb_dev_free(ctx, zd, nz * sizeof(cd));
  return rc;
}

// ---------------------------------------------------------------------------
// residual check of candidate eigenpairs
// ---------------------------------------------------------------------------
extern "C" int qb_residuals(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* lam,
                            int64_t nl, const qb_cheb_params* cheb, const double* phi,
                            int32_t batch, double* res2, double* resn1, double* resninf,
                            double* opn1, double* opninf) {
  if (!ctx) return QB_E_ARG;
  if (!g || (nl > 0 && (!lam || !phi)) || nl < 0 || batch < 0)
    return qb_fail(ctx, QB_E_ARG, "null argument or negative count");
  if (g->ctx != ctx) return qb_fail(ctx, QB_E_ARG, "geometry belongs to another context");
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS) return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (op == QB_DLP_KRESS && g->d.ncomp != 1)
    return qb_fail(ctx, QB_E_ARG, "DLP_kress expects one boundary component");
  if (qb_nirr(g) != 1) return qb_fail(ctx, QB_E_ARG, "the residual check takes one irrep");
  int rc = check_cheb(ctx, cheb);
  if (rc) return rc;
  if (nl == 0) return QB_OK;
  (void)hipSetDevice(ctx->device);
  double rmin, rmax;
  op_range(g, op, &rmin, &rmax);
  const int64_t N = qb_dim(g);
  const size_t per = (size_t)N * N * sizeof(cd);
  int64_t G = batch > 0 ? batch : (int64_t)std::max<size_t>(1, ((size_t)8 << 30) / per);
  G = std::min<int64_t>(G, nl);
  const int nsplit = 8;
  cd *zd = nullptr, *Td = nullptr, *phid = nullptr, *part = nullptr;
  double *rowabs = nullptr, *colsum = nullptr;
  const size_t b_z = G * sizeof(cd), b_T = G * per, b_phi = (size_t)G * N * sizeof(cd);
  const size_t b_part = (size_t)nsplit * G * N * sizeof(cd);
  const size_t b_row = (size_t)nsplit * G * N * sizeof(double), b_col = (size_t)G * N * sizeof(double);
  rc = qb_dev_alloc(ctx, (void**)&zd, b_z);
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&Td, b_T);
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&phid, b_phi);
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&part, b_part);
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&rowabs, b_row);
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&colsum, b_col);
  std::vector<cd> part_h;
  std::vector<double> row_h, col_h;
  for (int64_t q0 = 0; !rc && q0 < nl; q0 += G) {
    const int64_t nb = std::min<int64_t>(G, nl - q0);
    hipError_t e = hipMemcpyAsync(zd, lam + 2 * q0, nb * sizeof(cd), hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(phid, phi + 2 * q0 * N, (size_t)nb * N * sizeof(cd),
                         hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      rc = qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
      break;
    }
    DevTables t;
    rc = qb_tables_build(ctx, (const double*)zd, nb, rmin, rmax, *cheb, true, &t);
    if (!rc) rc = qb_assemble_dev(ctx, g, op, zd, nb, t, Td, nullptr);
    if (!rc) rc = qb_gemv_norms(ctx, Td, phid, N, (int)nb, nsplit, part, rowabs, colsum);
    part_h.resize((size_t)nsplit * nb * N);
    row_h.resize((size_t)nsplit * nb * N);
    col_h.resize((size_t)nb * N);
    if (!rc) {
      e = hipMemcpyAsync(part_h.data(), part, part_h.size() * sizeof(cd), hipMemcpyDeviceToHost,
                         ctx->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(row_h.data(), rowabs, row_h.size() * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(col_h.data(), colsum, col_h.size() * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) rc = qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
    } else {
      (void)hipStreamSynchronize(ctx->stream);
    }
    qb_tables_free(ctx, &t);
    if (rc) break;
    for (int64_t b = 0; b < nb; ++b) {
      double s2 = 0, s1 = 0, smx = 0, rmx = 0, cmx = 0;
      for (int64_t i = 0; i < N; ++i) {
        double yr = 0, yi = 0, ra = 0;
        for (int s = 0; s < nsplit; ++s) {
          const size_t o = ((size_t)s * nb + b) * N + i;
          yr += part_h[o].re;
          yi += part_h[o].im;
          ra += row_h[o];
        }
        const double a = hypot(yr, yi);
        s2 += a * a;
        s1 += a;
        smx = std::max(smx, a);
        rmx = std::max(rmx, ra);
        cmx = std::max(cmx, col_h[(size_t)b * N + i]);
      }
      if (res2) res2[q0 + b] = sqrt(s2);
      if (resn1) resn1[q0 + b] = s1;
      if (resninf) resninf[q0 + b] = smx;
      if (opn1) opn1[q0 + b] = cmx;
      if (opninf) opninf[q0 + b] = rmx;
    }
  }
  qb_dev_free(ctx, colsum, b_col);
  qb_dev_free(ctx, rowabs, b_row);
  qb_dev_free(ctx, part, b_part);
  qb_dev_free(ctx, phid, b_phi);
  qb_dev_free(ctx, Td, b_T);
  qb_dev_free(ctx, zd, b_z);
  return rc;
}

// ---------------------------------------------------------------------------
// Beyn window
// ---------------------------------------------------------------------------
struct qb_window {
  qb_ctx* ctx = nullptr;
  const qb_geom* g = nullptr;
  // nq: matrices of the window (the batch of every launch) = nc contour
  // nodes x nwin contours x nirr irreps assembled together; matrix
  // (e * nwin + c) * nc + j for irrep e, contour c, node j.  A moment block
  // is one (irrep, contour) pair: block e * nwin + c.
  int op = 0, nq = 0, precision = 0;
  int nc = 0, nirr = 1, nwin = 1;
  int64_t N = 0;
  double k0re = 0, k0im = 0, R = 0;
  std::vector<cd> zj, wj, wzj;  // host copies
  cd* z_dev = nullptr;
  cd* w_dev = nullptr;   // [w_j | w_j z_j]
  cd* T64 = nullptr;     // FP64 factors (QB_FP64) or the FP64 matrices (QB_MIXED)
  cf* LU32 = nullptr;    // complex64 factors (QB_MIXED)
  int* ipiv = nullptr;   // [nq][N]
  int* perm = nullptr;   // [nq][N]
  int* info = nullptr;   // [nq]
  double* tnorm = nullptr;  // [nq] ||T_j||_inf (QB_MIXED)
  double* scratch_d = nullptr;  // [3 nq] per-node norms
  std::vector<int> fb_nodes;    // nodes refactored in FP64
  cd* fb_lu = nullptr;          // [fb][N][N]
  int* fb_ipiv = nullptr;
  int* fb_perm = nullptr;
  size_t fb_cap = 0;
  // residues of T for the int8 (Ozaki) refinement residual (QB_MIXED):
  // [nq][16][2][Np][Np] int8 and the row exponents [nq][N]
  signed char* ozA = nullptr;
  int* ozE = nullptr;
  cd* qr = nullptr;      // [A0 | A1] after qb_window_moments_qr (N x 2r)
  cd* qr_tau = nullptr;  // [r]
  size_t qr_cap = 0;     // elements of qr
  int qr_rcap = 0, qr_r = 0;
  qb_stats st{};
};

// The int8 (Ozaki) residual of the mixed-precision refinement: taken by
// QB_MIXED_INT8RES for N that its exponent range covers (else, and with
// QB_MIXED, the residual is a complex128 GEMM on T64).
static bool oz_enabled(int precision, int64_t N) {
  return precision == QB_MIXED_INT8RES && N <= 65535;
}
// refinement nodes whose residual runs in one int8 launch (bounds the step buffers)
#define QB_OZ_SLOTS 16

static size_t win_bytes_core(int64_t N, int nq, int precision) {
  size_t NN = (size_t)N * N;
  size_t b = (size_t)nq * NN * sizeof(cd);
  if (precision == QB_MIXED || precision == QB_MIXED_INT8RES) {
    b += (size_t)nq * NN * sizeof(cf);
    if (oz_enabled(precision, N)) b += qb_oz_window_bytes(N, nq);
  }
  b += 2 * (size_t)nq * N * sizeof(int) + (size_t)nq * sizeof(int);
  return b;
}

extern "C" int64_t qb_window_footprint(int64_t N, int32_t nq, int32_t precision, int32_t r) {
  size_t b = win_bytes_core(N, nq, precision);
  size_t nv = (size_t)N * r;
  // moment work buffers: X and R (complex128) per node, a complex64 copy
  // per node in mixed mode, V, A0, A1, and the QR block [A0 | A1] that
  // qb_window_moments_qr keeps
  b += 2 * (size_t)nq * nv * sizeof(cd) + 5 * nv * sizeof(cd);
  if (precision == QB_MIXED || precision == QB_MIXED_INT8RES) {
    b += (size_t)nq * nv * sizeof(cf);
    if (oz_enabled(precision, N))
      b += qb_oz_step_bytes(N, r, nq < QB_OZ_SLOTS ? nq : QB_OZ_SLOTS);
  }
  // Chebyshev tables, bounded by 1.2e5 panels of degree 7 per node
  b += (size_t)nq * (32u << 20);
  return (int64_t)b;
}

static void win_free(qb_window* w) {
  qb_ctx* ctx = w->ctx;
  size_t NN = (size_t)w->N * w->N;
  qb_dev_free(ctx, w->z_dev, (size_t)w->nwin * w->nc * sizeof(cd));
  qb_dev_free(ctx, w->w_dev, 2 * (size_t)w->nwin * w->nc * sizeof(cd));
  qb_dev_free(ctx, w->T64, (size_t)w->nq * NN * sizeof(cd));
  qb_dev_free(ctx, w->LU32, (size_t)w->nq * NN * sizeof(cf));
  qb_dev_free(ctx, w->ipiv, (size_t)w->nq * w->N * sizeof(int));
  qb_dev_free(ctx, w->perm, (size_t)w->nq * w->N * sizeof(int));
  qb_dev_free(ctx, w->info, w->nq * sizeof(int));
  qb_dev_free(ctx, w->tnorm, w->nq * sizeof(double));
  qb_dev_free(ctx, w->scratch_d, 3 * w->nq * sizeof(double));
  qb_dev_free(ctx, w->fb_lu, w->fb_cap * NN * sizeof(cd));
  qb_dev_free(ctx, w->fb_ipiv, w->fb_cap * w->N * sizeof(int));
  qb_dev_free(ctx, w->fb_perm, w->fb_cap * w->N * sizeof(int));
  qb_dev_free(ctx, w->qr, w->qr_cap * sizeof(cd));
  qb_dev_free(ctx, w->qr_tau, (size_t)w->qr_rcap * sizeof(cd));
  if (w->ozA) {
    qb_dev_free(ctx, w->ozA, qb_oz_window_bytes(w->N, w->nq) - (size_t)w->nq * w->N * sizeof(int));
    qb_dev_free(ctx, w->ozE, (size_t)w->nq * w->N * sizeof(int));
  }
}

extern "C" void qb_window_destroy(qb_window* w) {
  if (!w) return;
  (void)hipSetDevice(w->ctx->device);
  (void)hipStreamSynchronize(w->ctx->stream);
  win_free(w);
  delete w;
}

// on error: keep the status in the caller's `rc` and jump to its cleanup
#define QB_TRY(x)      \
  do {                 \
    int _rc = (x);     \
    if (_rc) {         \
      rc = _rc;        \
      goto fail;       \
    }                  \
  } while (0)

// nwin contours k0[c] + R[c] e^{i 2 pi j / nq} on one geometry, batched.
static int win_create(qb_ctx* ctx, const qb_geom* g, int32_t op, int32_t nwin, const double* k0,
                      const double* Rs, int32_t nq, const qb_cheb_params* cheb,
                      int32_t precision, qb_window** out, qb_stats* stats) {
  if (!ctx) return QB_E_ARG;
  if (!g || !out || !k0 || !Rs) return qb_fail(ctx, QB_E_ARG, "null argument");
  *out = nullptr;
  if (nwin < 1 || nwin > 4096) return qb_fail(ctx, QB_E_ARG, "nwin must be in [1, 4096]");
  for (int c = 0; c < nwin; ++c)
    if (!(Rs[c] > 0.0)) return qb_fail(ctx, QB_E_ARG, "contour radius must be > 0");
  if (nq >= 1 && (int64_t)nq * nwin * qb_nirr(g) > 4096)
    return qb_fail(ctx, QB_E_ARG, "nq x nwin x irreps must be <= 4096");
  if (g->ctx != ctx) return qb_fail(ctx, QB_E_ARG, "geometry belongs to another context");
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS) return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (op == QB_DLP_KRESS && g->d.ncomp != 1)
    return qb_fail(ctx, QB_E_ARG, "DLP_kress expects one boundary component");
  if (nq < 1 || nq > 4096) return qb_fail(ctx, QB_E_ARG, "nq must be in [1, 4096]");
  if (precision != QB_FP64 && precision != QB_MIXED && precision != QB_MIXED_INT8RES)
    return qb_fail(ctx, QB_E_ARG, "precision must be QB_FP64, QB_MIXED or QB_MIXED_INT8RES");
  int rc = check_cheb(ctx, cheb);
  if (rc) return rc;
  // the int8 residual of QB_MIXED_INT8RES, or the complex128 GEMM of QB_MIXED
  const bool oz = oz_enabled(precision, qb_dim(g));
  if (precision == QB_MIXED_INT8RES) precision = QB_MIXED;
  (void)hipSetDevice(ctx->device);
  qb_window* w = new qb_window();
  w->ctx = ctx;
  w->g = g;
  w->op = op;
  w->nc = nq;
  w->nwin = nwin;
  w->nirr = qb_nirr(g);
  w->nq = nq * nwin * w->nirr;
  w->precision = precision;
  w->N = qb_dim(g);
  w->k0re = k0[0];
  w->k0im = k0[1];
  w->R = Rs[0];
  const int64_t N = w->N;
  const size_t NN = (size_t)N * N;
  const int nc = w->nc;     // contour nodes per contour
  const int nz = nc * nwin;  // wavenumbers assembled; nq from here on is the batch nz x nirr
  nq = w->nq;
  double t0 = now_ms(), t1, t2;
  bool first_done = false;  // first LU panel factored during the assembly
  DevTables tab;
  std::vector<cd> wbuf(2 * nz);
  double rmin, rmax;
  op_range(g, op, &rmin, &rmax);
  // contour nodes and weights of every contour (beyn.jl:258-261)
  w->zj.resize(nz);
  w->wj.resize(nz);
  w->wzj.resize(nz);
  for (int cw = 0; cw < nwin; ++cw) {
    const double R = Rs[cw];
    for (int j = 0; j < nc; ++j) {
      long double th =
          2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)nc;
      double c = (double)cosl(th), s = (double)sinl(th);
      cd e = mk(c, s);
      const int q = cw * nc + j;
      w->zj[q] = mk(k0[2 * cw] + R * c, k0[2 * cw + 1] + R * s);
      w->wj[q] = (R / (double)nc) * e;
      w->wzj[q] = w->wj[q] * w->zj[q];
      wbuf[q] = w->wj[q];
      wbuf[nz + q] = w->wzj[q];
    }
  }
  QB_TRY(qb_dev_alloc(ctx, (void**)&w->z_dev, nz * sizeof(cd)));
  QB_TRY(qb_dev_alloc(ctx, (void**)&w->w_dev, 2 * nz * sizeof(cd)));
  QB_TRY(qb_dev_alloc(ctx, (void**)&w->T64, (size_t)nq * NN * sizeof(cd)));
  if (precision == QB_MIXED) {
    QB_TRY(qb_dev_alloc(ctx, (void**)&w->LU32, (size_t)nq * NN * sizeof(cf)));
    QB_TRY(qb_dev_alloc(ctx, (void**)&w->tnorm, nq * sizeof(double)));
  }
  QB_TRY(qb_dev_alloc(ctx, (void**)&w->ipiv, (size_t)nq * N * sizeof(int)));
  QB_TRY(qb_dev_alloc(ctx, (void**)&w->perm, (size_t)nq * N * sizeof(int)));
  QB_TRY(qb_dev_alloc(ctx, (void**)&w->info, nq * sizeof(int)));
  QB_TRY(qb_dev_alloc(ctx, (void**)&w->scratch_d, 3 * nq * sizeof(double)));
  if (hipMemcpyAsync(w->z_dev, w->zj.data(), nz * sizeof(cd), hipMemcpyHostToDevice,
                     ctx->stream) != hipSuccess ||
      hipMemcpyAsync(w->w_dev, wbuf.data(), 2 * nz * sizeof(cd), hipMemcpyHostToDevice,
                     ctx->stream) != hipSuccess ||
      hipMemsetAsync(w->info, 0, nq * sizeof(int), ctx->stream) != hipSuccess) {
    rc = qb_fail(ctx, QB_E_HIP, "window upload failed");
    goto fail;
  }
  // tables + assembly (construct_boundary_matrices!, beyn.jl:264)
  {
    // With real k0 the contour nodes come in conjugate pairs, assembled
    // together from envelope tables of H_n e^{-ikr} held in LDS (no uniform
    // tables).  QB_ASM_TABLES=1: the same pairs from the uniform Hankel
    // tables; QB_ASM_NO_CONJ=1: every node from its own H and J tables.
    cf* o32 = precision == QB_MIXED ? w->LU32 : nullptr;
    const bool no_conj = getenv("QB_ASM_NO_CONJ") != nullptr;
    int crc = 1;
    // The first LU panel (columns [0, NB) of every node, complete after the
    // assembly's first launch) is factored on the second stream while the
    // rest of the matrices assemble; QB_FIRST_PANEL_AHEAD=0 turns this off.
    const char* fpa = getenv("QB_FIRST_PANEL_AHEAD");
    const bool ahead = !(fpa && fpa[0] == '0') && !ctx->graphs;
    const int nb = qb_lu_nb((int)N);
    auto first_panel = [&]() -> int {
      const int r = precision == QB_FP64
                        ? qb_getrf_first_panel<cd>(ctx, w->T64, (int)N, N, (int64_t)NN, w->ipiv,
                                                   N, w->info, nq)
                        : qb_getrf_first_panel<cf>(ctx, w->LU32, (int)N, N, (int64_t)NN,
                                                   w->ipiv, N, w->info, nq);
      if (r == QB_OK) first_done = true;
      return r;
    };
    if (!no_conj && !getenv("QB_ASM_TABLES"))
      crc = qb_assemble_env_dev(ctx, g, op, w->zj.data(), w->z_dev, nz, rmin, rmax, w->T64, o32,
                                ahead ? nb : 0,
                                ahead ? std::function<int()>(first_panel) : std::function<int()>());
    if (crc == 1) {
      QB_TRY(qb_tables_build(ctx, (const double*)w->z_dev, nz, rmin, rmax, *cheb, true, &tab));
      crc = no_conj ? 1 : qb_assemble_conj_dev(ctx, g, op, w->zj.data(), w->z_dev, nz, tab,
                                               w->T64, o32);
      if (crc == 1) crc = qb_assemble_dev(ctx, g, op, w->z_dev, nz, tab, w->T64, o32);
    }
    QB_TRY(crc);
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    rc = qb_fail(ctx, QB_E_HIP, "assembly failed");
    goto fail;
  }
  qb_tables_free(ctx, &tab);
  if (oz) {
    // residues of every T(z_j) for the int8 refinement residual
    const size_t ebytes = (size_t)nq * N * sizeof(int);
    QB_TRY(qb_dev_alloc(ctx, (void**)&w->ozA, qb_oz_window_bytes(N, nq) - ebytes));
    QB_TRY(qb_dev_alloc(ctx, (void**)&w->ozE, ebytes));
    QB_TRY(qb_oz_prepare(ctx, w->T64, N, nq, w->ozA, w->ozE));
  }
  t1 = now_ms();
  // LU of every node (beyn.jl:270-277)
  if (precision == QB_FP64) {
    QB_TRY(qb_getrf_graph<cd>(ctx, w->T64, (int)N, N, (int64_t)NN, w->ipiv, N, w->info, nq,
                              first_done));
  } else {
    QB_TRY(qb_node_rownorm(ctx, w->T64, N, nq, w->tnorm));
    QB_TRY(qb_getrf_graph<cf>(ctx, w->LU32, (int)N, N, (int64_t)NN, w->ipiv, N, w->info, nq,
                              first_done));
  }
  QB_TRY(qb_perm_from_ipiv(ctx, w->ipiv, w->perm, (int)N, nq));
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    rc = qb_fail(ctx, QB_E_HIP, "factorisation failed");
    goto fail;
  }
  t2 = now_ms();
  w->st.N = N;
  w->st.nq = nc;
  w->st.precision = precision;
  w->st.t_assemble_ms = t1 - t0;
  w->st.t_factor_ms = t2 - t1;
  w->st.rmin = rmin;
  w->st.rmax = rmax;
  if (stats) *stats = w->st;
  *out = w;
  return QB_OK;
fail:
  // both streams: the first LU panel or the LU's lookahead may still be
  // running on the second one when a later step fails
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamSynchronize(ctx->stream_aux);
  qb_tables_free(ctx, &tab);
  if (rc == QB_OK) rc = QB_E_HIP;
  win_free(w);
  delete w;
  return rc;
}

extern "C" int qb_window_create(qb_ctx* ctx, const qb_geom* g, int32_t op, double k0_re,
                                double k0_im, double R, int32_t nq, const qb_cheb_params* cheb,
                                int32_t precision, qb_window** out, qb_stats* stats) {
  const double k0[2] = {k0_re, k0_im};
  return win_create(ctx, g, op, 1, k0, &R, nq, cheb, precision, out, stats);
}

extern "C" int qb_window_create_multi(qb_ctx* ctx, const qb_geom* g, int32_t op, int32_t nwin,
                                      const double* k0, const double* R, int32_t nq,
                                      const qb_cheb_params* cheb, int32_t precision,
                                      qb_window** out, qb_stats* stats) {
  return win_create(ctx, g, op, nwin, k0, R, nq, cheb, precision, out, stats);
}

extern "C" int32_t qb_window_blocks(const qb_window* w) { return w ? w->nirr * w->nwin : 0; }

extern "C" int qb_window_matrix_cols(qb_window* w, int32_t node, int64_t j0, int64_t ncols,
                                     double* out) {
  if (!w) return QB_E_ARG;
  qb_ctx* ctx = w->ctx;
  if (!out) return qb_fail(ctx, QB_E_ARG, "null argument");
  if (node < 0 || node >= w->nq) return qb_fail(ctx, QB_E_ARG, "node out of range");
  if (j0 < 0 || ncols < 0 || j0 + ncols > w->N)
    return qb_fail(ctx, QB_E_ARG, "column range out of [0, N)");
  if (w->precision != QB_MIXED)
    return qb_fail(ctx, QB_E_STATE,
                   "an FP64 window holds the LU factors, not T(z): only mixed windows keep T");
  if (ncols == 0) return QB_OK;
  (void)hipSetDevice(ctx->device);
  const size_t NN = (size_t)w->N * w->N;
  QB_HIP(ctx, hipMemcpyAsync(out, w->T64 + (size_t)node * NN + (size_t)j0 * w->N,
                             (size_t)ncols * w->N * sizeof(cd), hipMemcpyDeviceToHost,
                             ctx->stream));
  QB_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QB_OK;
}

// FP64 refactorisation of the nodes listed in w->fb_nodes that are not yet
// factored (mixed-precision fallback when refinement does not converge).
static int win_fp64_fallback(qb_window* w, const std::vector<int>& nodes) {
  qb_ctx* ctx = w->ctx;
  const int64_t N = w->N;
  const size_t NN = (size_t)N * N;
  size_t need = w->fb_nodes.size() + nodes.size();
  if (need > w->fb_cap) {
    size_t cap = need;
    cd* lu = nullptr;
    int *ip = nullptr, *pm = nullptr;
    int rc = qb_dev_alloc(ctx, (void**)&lu, cap * NN * sizeof(cd));
    if (!rc) rc = qb_dev_alloc(ctx, (void**)&ip, cap * N * sizeof(int));
    if (!rc) rc = qb_dev_alloc(ctx, (void**)&pm, cap * N * sizeof(int));
    if (rc) {
      qb_dev_free(ctx, lu, cap * NN * sizeof(cd));
      qb_dev_free(ctx, ip, cap * N * sizeof(int));
      return rc;
    }
    if (w->fb_cap) {
      size_t old = w->fb_nodes.size();
      QB_HIP(ctx, hipMemcpyAsync(lu, w->fb_lu, old * NN * sizeof(cd), hipMemcpyDeviceToDevice,
                                 ctx->stream));
      QB_HIP(ctx, hipMemcpyAsync(ip, w->fb_ipiv, old * N * sizeof(int), hipMemcpyDeviceToDevice,
                                 ctx->stream));
      QB_HIP(ctx, hipMemcpyAsync(pm, w->fb_perm, old * N * sizeof(int), hipMemcpyDeviceToDevice,
                                 ctx->stream));
      QB_HIP(ctx, hipStreamSynchronize(ctx->stream));
      qb_dev_free(ctx, w->fb_lu, w->fb_cap * NN * sizeof(cd));
      qb_dev_free(ctx, w->fb_ipiv, w->fb_cap * N * sizeof(int));
      qb_dev_free(ctx, w->fb_perm, w->fb_cap * N * sizeof(int));
    }
    w->fb_lu = lu;
    w->fb_ipiv = ip;
    w->fb_perm = pm;
    w->fb_cap = cap;
  }
  if (nodes.empty()) return QB_OK;
  // copy the nodes' FP64 matrices to consecutive slots and factor them as one batch
  const size_t s0 = w->fb_nodes.size();
  const int nb = (int)nodes.size();
  for (int i = 0; i < nb; ++i)
    QB_HIP(ctx, hipMemcpyAsync(w->fb_lu + (s0 + i) * NN, w->T64 + (size_t)nodes[i] * NN,
                               NN * sizeof(cd), hipMemcpyDeviceToDevice, ctx->stream));
  int rc = qb_getrf_batched<cd>(ctx, w->fb_lu + s0 * NN, (int)N, N, (int64_t)NN,
                                w->fb_ipiv + s0 * N, N, nullptr, nb);
  if (rc) return rc;
  rc = qb_perm_from_ipiv(ctx, w->fb_ipiv + s0 * N, w->fb_perm + s0 * N, (int)N, nb);
  if (rc) return rc;
  for (int node : nodes) w->fb_nodes.push_back(node);
  return QB_OK;
}

// The moment loop.  With dev_out == nullptr, A0 and A1 go to the host
// arrays; otherwise they stay on the device at dev_out (A0) and
// dev_out + N r (A1) and the host arrays are not touched.
static int win_moments(qb_window* w, int32_t r, const double* V, double* A0, double* A1,
                       double refine_tol, qb_stats* stats, cd* dev_out) {
  if (!w) return QB_E_ARG;
  qb_ctx* ctx = w->ctx;
  if (!V || (!dev_out && (!A0 || !A1))) return qb_fail(ctx, QB_E_ARG, "null argument");
  if (r < 1 || r > w->N) return qb_fail(ctx, QB_E_ARG, "r must be in [1, N]");
  (void)hipSetDevice(ctx->device);
  const int64_t N = w->N;
  const int nq = w->nq;
  const size_t NN = (size_t)N * N;
  const int64_t nv = N * (int64_t)r;
  const int nc = w->nc, nwin = w->nwin, nz = nc * nwin;
  const int nblk = w->nirr * nwin;  // moment blocks; nq = nc x nblk matrices
  double t0 = now_ms();
  cd *Vd = nullptr, *X = nullptr, *Rr = nullptr, *A0d = nullptr, *A1d = nullptr;
  cf* D32 = nullptr;
  int* actd = nullptr;
  signed char *ozB = nullptr, *ozC = nullptr;  // int8 residual step buffers
  int* ozF = nullptr;
  size_t oz_bc = 0, oz_f = 0;
  int rc = QB_OK;
  int iters_max = 0;
  double resid_max = 0.0;
  std::vector<double> h3(3 * nq);
  QB_TRY(qb_dev_alloc(ctx, (void**)&Vd, nv * sizeof(cd)));
  QB_TRY(qb_dev_alloc(ctx, (void**)&X, (size_t)nq * nv * sizeof(cd)));
  // moments per block b: device output [A0_b | A1_b] blocks (2 nv each); host
  // outputs A0 and A1 hold the nblk blocks one after the other
  if (!dev_out) {
    QB_TRY(qb_dev_alloc(ctx, (void**)&A0d, (size_t)nblk * nv * sizeof(cd)));
    QB_TRY(qb_dev_alloc(ctx, (void**)&A1d, (size_t)nblk * nv * sizeof(cd)));
  }
  if (hipMemcpyAsync(Vd, V, nv * sizeof(cd), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
    rc = qb_fail(ctx, QB_E_HIP, "V upload failed");
    goto fail;
  }
  if (w->precision == QB_FP64) {
    // X_j = P_j V, then the two triangular solves (ldiv!, beyn.jl:281)
    QB_TRY(qb_gather_rows_cd(ctx, Vd, 0, w->perm, X, (int)N, r, nq));
    QB_TRY(qb_trsm_lu_batched<cd>(ctx, w->T64, (int)N, N, (int64_t)NN, X, r, N, nv, nq));
  } else {
    // complex64 solve + FP64 iterative refinement against the resident T64
    QB_TRY(qb_dev_alloc(ctx, (void**)&Rr, (size_t)nq * nv * sizeof(cd)));
    QB_TRY(qb_dev_alloc(ctx, (void**)&D32, (size_t)nq * nv * sizeof(cf)));
    if (w->ozA) {
      const int ns = std::min(nq, QB_OZ_SLOTS);
      const size_t Np = (size_t)qb_oz_pad(N), rp = (size_t)qb_oz_pad(r);
      oz_bc = 2 * 16 * 2 * rp * Np * ns;
      oz_f = rp * ns * sizeof(int);
      QB_TRY(qb_dev_alloc(ctx, (void**)&ozB, oz_bc / 2));
      QB_TRY(qb_dev_alloc(ctx, (void**)&ozC, oz_bc / 2));
      QB_TRY(qb_dev_alloc(ctx, (void**)&ozF, oz_f));
    }
    if (hipMemsetAsync(X, 0, (size_t)nq * nv * sizeof(cd), ctx->stream) != hipSuccess) {
      rc = qb_fail(ctx, QB_E_HIP, "memset failed");
      goto fail;
    }
    QB_TRY(qb_bcast(ctx, Vd, Rr, nv, nv, nq));
    // host copies of ||T_j||_inf and ||V||_max for the backward error;
    // ||V||_max is formed on the host while the first residual pass runs
    std::vector<double> tn(nq);
    double vmax = -1.0;
    if (hipMemcpyAsync(tn.data(), w->tnorm, nq * sizeof(double), hipMemcpyDeviceToHost,
                       ctx->stream) != hipSuccess) {
      rc = qb_fail(ctx, QB_E_HIP, "norm download failed");
      goto fail;
    }
    std::vector<double> be(nq, 1e300), prev(nq, 1e300);
    std::vector<char> stag(nq, 0);
    std::vector<int> stagnant;
    const int maxit = 8;
    // Only the nodes still above the tolerance take part in an iteration:
    // their indices go to the device (ctx->bix) and every batched launch of
    // the iteration (gather, complex64 TRSM, update, FP64 residual GEMM,
    // norms) covers just them.  Nodes far from the real axis converge in
    // one or two steps; the few close to an eigenvalue need more.
    std::vector<int> act(nq);
    for (int j = 0; j < nq; ++j) act[j] = j;
    QB_TRY(qb_dev_alloc(ctx, (void**)&actd, nq * sizeof(int)));
    const bool trace = getenv("QB_REFINE_TRACE") != nullptr;
    // Residual pass over the nodes of `list` on all r columns: bn[j] = the
    // backward error of node j.  Every node's error at exit is one of these
    // full-column values, so refine_resid_max certifies the whole block.
    // The moments add the nodes' solutions, so their errors must be small in
    // absolute terms, not only relative to each node's own ||X_j||: a node
    // near an eigenvalue has ||X_j|| far above the others, and a normwise
    // backward error tol there leaves absolute errors ||X_j|| / ||X||_typ
    // times those of the other nodes.  The error of node j is therefore
    // measured against min(||X_j||, xref), xref = the median ||X||max over
    // the nodes after the first solve.  A node whose target lies below what
    // an FP64 residual resolves (2 sqrt(N) u64) is refactored in FP64.
    std::vector<double> bn(nq, 0.0);
    double xref = -1.0;
    const double floor64 = 2.0 * sqrt((double)N) * 1.1102230246251565e-16;
    auto residual_pass = [&](const std::vector<int>& list) -> int {
      const int n = (int)list.size();
      if (hipMemcpyAsync(actd, list.data(), n * sizeof(int), hipMemcpyHostToDevice,
                         ctx->stream) != hipSuccess)
        return qb_fail(ctx, QB_E_HIP, "node list upload failed");
      ctx->bix = actd;
      int e = qb_node_maxabs(ctx, X, nv, nv, n, w->scratch_d + nq, nq);
      if (w->ozA) {
        // R = V - T X on the int8 matrix cores, QB_OZ_SLOTS nodes per launch
        for (int c0 = 0; c0 < n && !e; c0 += QB_OZ_SLOTS) {
          ctx->bix = actd + c0;
          e = qb_oz_residual(ctx, w->ozA, w->ozE, N, X, nv, r, Vd, Rr, nv,
                             std::min(QB_OZ_SLOTS, n - c0), ozB, ozF, ozC);
        }
        ctx->bix = actd;
      } else {
        if (!e) e = qb_bcast(ctx, Vd, Rr, nv, nv, n);
        if (!e)
          e = gemm_sub<cd>(ctx, (int)N, r, (int)N, w->T64, N, (int64_t)NN, X, N, nv, Rr, N, nv,
                           n);
      }
      if (!e) e = qb_node_maxabs(ctx, Rr, nv, nv, n, w->scratch_d + 2 * nq, nq);
      ctx->bix = nullptr;
      if (e) return e;
      if (hipMemcpyAsync(h3.data(), w->scratch_d, 3 * nq * sizeof(double),
                         hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        return qb_fail(ctx, QB_E_HIP, "refinement norms failed");
      if (vmax < 0.0) {
        double m2 = 0.0;
        for (int64_t i = 0; i < nv; ++i) m2 = std::max(m2, V[2 * i] * V[2 * i] + V[2 * i + 1] * V[2 * i + 1]);
        vmax = sqrt(m2);
      }
      if (hipStreamSynchronize(ctx->stream) != hipSuccess)
        return qb_fail(ctx, QB_E_HIP, "refinement norms failed");
      if (xref < 0.0) {
        std::vector<double> xs;
        for (int j : list) xs.push_back(h3[nq + j]);
        std::nth_element(xs.begin(), xs.begin() + xs.size() / 2, xs.end());
        xref = xs[xs.size() / 2];
      }
      for (int j : list) bn[j] = h3[2 * nq + j] / (tn[j] * std::min(h3[nq + j], xref) + vmax);
      return QB_OK;
    };
    int it = 0;
    for (; it < maxit && !act.empty(); ++it) {
      const int na = (int)act.size();
      if (hipMemcpyAsync(actd, act.data(), na * sizeof(int), hipMemcpyHostToDevice,
                         ctx->stream) != hipSuccess) {
        rc = qb_fail(ctx, QB_E_HIP, "node list upload failed");
        goto fail;
      }
      ctx->bix = actd;
      // D = LU32^{-1} P R (complex64), X += D on the active nodes
      QB_TRY(qb_gather_rows_cd2cf(ctx, Rr, nv, w->perm, D32, (int)N, r, na));
      QB_TRY(qb_trsm_lu_batched<cf>(ctx, w->LU32, (int)N, N, (int64_t)NN, D32, r, N, nv, na));
      QB_TRY(qb_add_cf_to_cd_nodes(ctx, D32, X, nv, nv, na));
      ctx->bix = nullptr;
      QB_TRY(residual_pass(act));
      for (int j : act) {
        prev[j] = be[j];
        be[j] = bn[j];
      }
      if (trace) {
        double bmax = 0;
        for (int j : act) bmax = std::max(bmax, be[j]);
        fprintf(stderr, "[qb refine] it %d: %d active nodes, max backward error %.3e\n", it, na,
                bmax);
      }
      std::vector<int> next;
      for (int j : act) {
        if (be[j] <= refine_tol) continue;
        if (it >= 1 && !(be[j] < 0.5 * prev[j])) stag[j] = 1;
        // the node's own normwise target refine_tol * xref / ||X_j|| is
        // below the FP64 residual's resolution: refinement cannot reach it
        if (h3[nq + j] > xref && refine_tol * xref / h3[nq + j] < floor64) stag[j] = 1;
        if (!stag[j]) next.push_back(j);
      }
      act.swap(next);
    }
    iters_max = it;
    // nodes that did not reach the tolerance: FP64 refactorisation and solve,
    // then one residual pass that measures their backward error as well
    for (int j = 0; j < nq; ++j)
      if (!(be[j] <= refine_tol)) stagnant.push_back(j);
    if (!stagnant.empty()) {
      std::vector<int> fresh;
      for (int j : stagnant)
        if (std::find(w->fb_nodes.begin(), w->fb_nodes.end(), j) == w->fb_nodes.end())
          fresh.push_back(j);
      QB_TRY(win_fp64_fallback(w, fresh));
      for (int j : stagnant) {
        size_t slot = std::find(w->fb_nodes.begin(), w->fb_nodes.end(), j) - w->fb_nodes.begin();
        cd* Xj = X + (size_t)j * nv;
        QB_TRY(qb_gather_rows_cd(ctx, Vd, 0, w->fb_perm + slot * N, Xj, (int)N, r, 1));
        QB_TRY(qb_trsm_lu_batched<cd>(ctx, w->fb_lu + slot * NN, (int)N, N, (int64_t)NN, Xj, r,
                                      N, nv, 1));
      }
      QB_TRY(residual_pass(stagnant));
      for (int j : stagnant) be[j] = bn[j];
    }
    for (int j = 0; j < nq; ++j) resid_max = std::max(resid_max, be[j]);
    w->st.nodes_fp64_fallback = (int)w->fb_nodes.size();
  }
  for (int b = 0; b < nblk; ++b) {
    const int cw = b % nwin;  // contour of block b (blocks are irrep-major)
    cd* a0 = dev_out ? dev_out + (size_t)2 * b * nv : A0d + (size_t)b * nv;
    cd* a1 = dev_out ? dev_out + (size_t)2 * b * nv + nv : A1d + (size_t)b * nv;
    QB_TRY(qb_moments(ctx, X + (size_t)b * nc * nv, nv, nc, w->w_dev + cw * nc,
                      w->w_dev + nz + cw * nc, a0, a1));
  }
  if (!dev_out &&
      (hipMemcpyAsync(A0, A0d, (size_t)nblk * nv * sizeof(cd), hipMemcpyDeviceToHost,
                      ctx->stream) != hipSuccess ||
       hipMemcpyAsync(A1, A1d, (size_t)nblk * nv * sizeof(cd), hipMemcpyDeviceToHost,
                      ctx->stream) != hipSuccess)) {
    rc = qb_fail(ctx, QB_E_HIP, "moment download failed");
    goto fail;
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    rc = qb_fail(ctx, QB_E_HIP, "moments failed");
    goto fail;
  }
  w->st.r = r;
  w->st.refine_iters_max = iters_max;
  w->st.refine_resid_max = resid_max;
  w->st.t_solve_ms = now_ms() - t0;
  if (stats) *stats = w->st;
fail:
  ctx->bix = nullptr;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamSynchronize(ctx->stream_aux);
  qb_dev_free(ctx, actd, nq * sizeof(int));
  qb_dev_free(ctx, Vd, nv * sizeof(cd));
  qb_dev_free(ctx, X, (size_t)nq * nv * sizeof(cd));
  qb_dev_free(ctx, Rr, (size_t)nq * nv * sizeof(cd));
  qb_dev_free(ctx, D32, (size_t)nq * nv * sizeof(cf));
  qb_dev_free(ctx, ozB, oz_bc / 2);
  qb_dev_free(ctx, ozC, oz_bc / 2);
  qb_dev_free(ctx, ozF, oz_f);
  if (!dev_out) {
    qb_dev_free(ctx, A0d, (size_t)nblk * nv * sizeof(cd));
    qb_dev_free(ctx, A1d, (size_t)nblk * nv * sizeof(cd));
  }
  return rc;
}

extern "C" int qb_window_moments(qb_window* w, int32_t r, const double* V, double* A0, double* A1,
                                 double refine_tol, qb_stats* stats) {
  return win_moments(w, r, V, A0, A1, refine_tol, stats, nullptr);
}

// A0 and A1 stay on the device in w->qr ([A0 | A1], N x 2r); A0 = Q R by
// Householder reflections, Q^H applied to A1; R (r x r upper triangular) and
// C = (Q^H A1)[:r, :] are returned, the reflectors kept for qb_window_basis.
// R (upper triangle of each block's A0 half, zeros below) and C (the first r
// rows of its A1 half) of every moment block, packed as [R_0 .. R_b | C_0 ..
// C_b], r x r each (ld r), so they leave the device in two contiguous copies
// instead of one row transfer per column.  Grid (ceil(r*r/256), blocks, 2).
__global__ void k_pack_rc(const cd* __restrict__ qr, int64_t N, int r, cd* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rr = (int64_t)r * r;
  if (e >= rr) return;
  const int ie = blockIdx.y, half = blockIdx.z;
  const int i = (int)(e % r), c = (int)(e / r);
  const cd* src = qr + ((int64_t)2 * ie + half) * N * r;
  cd v = src[i + (int64_t)c * N];
  if (half == 0 && i > c) v = mk(0.0, 0.0);
  out[((int64_t)half * gridDim.y + ie) * rr + e] = v;
}

extern "C" int qb_window_moments_qr(qb_window* w, int32_t r, const double* V, double* R,
                                    double* Cm, double refine_tol, qb_stats* stats) {
  if (!w) return QB_E_ARG;
  qb_ctx* ctx = w->ctx;
  if (!V || !R || !Cm) return qb_fail(ctx, QB_E_ARG, "null argument");
  if (r < 1 || r > w->N) return qb_fail(ctx, QB_E_ARG, "r must be in [1, N]");
  (void)hipSetDevice(ctx->device);
  const int64_t N = w->N;
  const int nirr = w->nirr * w->nwin;  // moment blocks
  const size_t need = (size_t)2 * N * r * nirr;
  if (w->qr_cap != need) {
    qb_dev_free(ctx, w->qr, w->qr_cap * sizeof(cd));
    qb_dev_free(ctx, w->qr_tau, (size_t)w->qr_rcap * sizeof(cd));
    w->qr = nullptr;
    w->qr_tau = nullptr;
    w->qr_cap = 0;
    w->qr_rcap = 0;
    int rc = qb_dev_alloc(ctx, (void**)&w->qr, need * sizeof(cd));
    if (!rc) rc = qb_dev_alloc(ctx, (void**)&w->qr_tau, (size_t)r * nirr * sizeof(cd));
    if (rc) {
      qb_dev_free(ctx, w->qr, need * sizeof(cd));
      w->qr = nullptr;
      return rc;
    }
    w->qr_cap = need;
    w->qr_rcap = r * nirr;
  }
  w->qr_r = 0;
  int rc = win_moments(w, r, V, nullptr, nullptr, refine_tol, nullptr, w->qr);
  if (rc) return rc;
  // one QR per block on its [A0_b | A1_b], all blocks in the same launches
  rc = qb_house_qr(ctx, w->qr, N, r, 2 * r, w->qr_tau, nirr, (int64_t)2 * N * r, (int64_t)r);
  if (rc) return rc;
  {
    const size_t pbytes = (size_t)2 * nirr * r * r * sizeof(cd);
    cd* pk = nullptr;
    rc = qb_dev_alloc(ctx, (void**)&pk, pbytes);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_rc, dim3((unsigned)(((int64_t)r * r + 255) / 256), nirr, 2),
                       dim3(256), 0, ctx->stream, w->qr, N, r, pk);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(R, pk, pbytes / 2, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(Cm, pk + (size_t)nirr * r * r, pbytes / 2, hipMemcpyDeviceToHost,
                         ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    qb_dev_free(ctx, pk, pbytes);
    if (e != hipSuccess) return qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  }
  w->qr_r = r;
  if (stats) *stats = w->st;
  return QB_OK;
}

// U = Q [Ur; 0] (N x k) with the reflectors of the last qb_window_moments_qr.
extern "C" int qb_window_basis_irrep(qb_window* w, int32_t irrep, const double* Ur, int32_t k,
                                     double* U);
extern "C" int qb_window_basis(qb_window* w, const double* Ur, int32_t k, double* U) {
  return qb_window_basis_irrep(w, 0, Ur, k, U);
}

extern "C" int qb_window_basis_irrep(qb_window* w, int32_t irrep, const double* Ur, int32_t k,
                                     double* U) {
  if (!w) return QB_E_ARG;
  qb_ctx* ctx = w->ctx;
  if (!U || (k > 0 && !Ur) || k < 0) return qb_fail(ctx, QB_E_ARG, "null argument");
  if (w->qr_r == 0) return qb_fail(ctx, QB_E_STATE, "no QR of the moments on this window");
  if (k > w->qr_r) return qb_fail(ctx, QB_E_ARG, "k must be <= r");
  if (irrep < 0 || irrep >= w->nirr * w->nwin)
    return qb_fail(ctx, QB_E_ARG, "block index out of range");
  if (k == 0) return QB_OK;
  (void)hipSetDevice(ctx->device);
  const int64_t N = w->N;
  const int r = w->qr_r;
  const size_t bytes = (size_t)N * k * sizeof(cd);
  const size_t ubytes = (size_t)r * k * sizeof(cd);
  cd *Y = nullptr, *Ud = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&Y, bytes);
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&Ud, ubytes);
  if (rc) {
    qb_dev_free(ctx, Y, bytes);
    return rc;
  }
  // Ur goes up in one contiguous copy and is spread into the first r rows of
  // Y on the device (a host-to-device 2D copy moves one row at a time)
  hipError_t e = hipMemsetAsync(Y, 0, bytes, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(Ud, Ur, ubytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(Y, N * sizeof(cd), Ud, r * sizeof(cd), r * sizeof(cd), k,
                         hipMemcpyDeviceToDevice, ctx->stream);
  if (e != hipSuccess) rc = qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  if (!rc)
    rc = qb_house_apply_q(ctx, w->qr + (size_t)2 * irrep * N * r, N, r,
                          w->qr_tau + (size_t)irrep * r, Y, k);
  if (!rc && (hipMemcpyAsync(U, Y, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
              hipStreamSynchronize(ctx->stream) != hipSuccess))
    rc = qb_fail(ctx, QB_E_HIP, "basis download failed");
  (void)hipStreamSynchronize(ctx->stream);
  qb_dev_free(ctx, Ud, ubytes);
  qb_dev_free(ctx, Y, bytes);
  return rc;
}

// ---------------------------------------------------------------------------
// Bessel evaluation and table tuning
// ---------------------------------------------------------------------------
extern "C" int qb_bessel_eval(qb_ctx* ctx, const double* z, int64_t n, double* h0, double* h1,
                              double* j0, double* j1) {
  if (!ctx) return QB_E_ARG;
  if (n < 0 || (n > 0 && (!z || !h0 || !h1 || !j0 || !j1)))
    return qb_fail(ctx, QB_E_ARG, "null argument");
  if (n == 0) return QB_OK;
  (void)hipSetDevice(ctx->device);
  cd *zd = nullptr, *od = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&zd, n * sizeof(cd));
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&od, 4 * n * sizeof(cd));
  std::vector<cd> ho;
  if (!rc && hipMemcpyAsync(zd, z, n * sizeof(cd), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    rc = qb_fail(ctx, QB_E_HIP, "upload failed");
  if (!rc) rc = qb_bessel_eval_dev(ctx, zd, n, od);
  if (!rc) {
    ho.resize(4 * n);
    if (hipMemcpyAsync(ho.data(), od, 4 * n * sizeof(cd), hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = qb_fail(ctx, QB_E_HIP, "download failed");
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (!rc) {
    for (int64_t i = 0; i < n; ++i) {
      h0[2 * i] = ho[4 * i].re;
      h0[2 * i + 1] = ho[4 * i].im;
      h1[2 * i] = ho[4 * i + 1].re;
      h1[2 * i + 1] = ho[4 * i + 1].im;
      j0[2 * i] = ho[4 * i + 2].re;
      j0[2 * i + 1] = ho[4 * i + 2].im;
      j1[2 * i] = ho[4 * i + 3].re;
      j1[2 * i + 1] = ho[4 * i + 3].im;
    }
  }
  qb_dev_free(ctx, zd, n * sizeof(cd));
  qb_dev_free(ctx, od, 4 * n * sizeof(cd));
  return rc;
}

extern "C" int qb_cheb_tune(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* z,
                            int64_t nz, double tol, int32_t samples, int32_t max_iter,
                            qb_cheb_params* cheb, double* errs) {
  if (!ctx) return QB_E_ARG;
  if (!g || !z || !cheb || nz < 1) return qb_fail(ctx, QB_E_ARG, "null argument");
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS) return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (samples < 2 || max_iter < 1 || !(tol > 0))
    return qb_fail(ctx, QB_E_ARG, "need samples >= 2, max_iter >= 1, tol > 0");
  int rc = check_cheb(ctx, cheb);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  // rmin_cheb = min_j 0.2/|z_j| (optimal_panelization.jl:222)
  double rmin_cheb = 1e300;
  for (int64_t q = 0; q < nz; ++q) rmin_cheb = std::min(rmin_cheb, 0.2 / hypot(z[2 * q], z[2 * q + 1]));
  double rmin, rmax;
  op_range(g, op, &rmin, &rmax);
  double rmin_h = std::max(rmin_cheb, rmin);
  cd* zd = nullptr;
  rc = qb_dev_alloc(ctx, (void**)&zd, nz * sizeof(cd));
  if (rc) return rc;
  if (hipMemcpyAsync(zd, z, nz * sizeof(cd), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
    qb_dev_free(ctx, zd, nz * sizeof(cd));
    return qb_fail(ctx, QB_E_HIP, "upload failed");
  }
  qb_cheb_params cp = *cheb;
  double e4[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  for (int it = 1; it <= max_iter; ++it) {
    if (cp.M_h > 16 || cp.M_j > 16) break;
    DevTables t;
    rc = qb_tables_build(ctx, (const double*)zd, nz, rmin_h, rmax, cp, true, &t);
    if (!rc) rc = qb_cheb_check_dev(ctx, t, zd, nz, samples, e4);
    (void)hipStreamSynchronize(ctx->stream);
    qb_tables_free(ctx, &t);
    if (rc) break;
    *cheb = cp;
    bool okH = e4[0] < tol && e4[1] < tol;
    bool okJ = e4[2] < tol && e4[3] < tol;
    if (okH && okJ) break;
    if (it == max_iter) break;
    if (!okH) {
      if (it % 5 == 0)
        cp.M_h += 2;
      else
        cp.n_panels_h = (int32_t)ceil(1.5 * cp.n_panels_h);
    }
    if (!okJ) {
      if (it % 5 == 0)
        cp.M_j += 2;
      else
        cp.n_panels_j = (int32_t)ceil(1.5 * cp.n_panels_j);
    }
  }
  if (errs)
    for (int c = 0; c < 4; ++c) errs[c] = e4[c];
  qb_dev_free(ctx, zd, nz * sizeof(cd));
  return rc;
}

// ---------------------------------------------------------------------------
// GEMM tile autotuning for the shapes of one window
// ---------------------------------------------------------------------------
extern "C" int qb_ctx_autotune(qb_ctx* ctx, int64_t N, int32_t nq, int32_t r, int32_t precision,
                               char* report, int64_t report_len) {
  if (!ctx) return QB_E_ARG;
  if (N < 64 || N > 65535 || nq < 1 || r < 1 || r > N)
    return qb_fail(ctx, QB_E_ARG, "autotune needs 64 <= N <= 65535, nq >= 1, 1 <= r <= N");
  if (precision != QB_FP64 && precision != QB_MIXED && precision != QB_MIXED_INT8RES)
    return qb_fail(ctx, QB_E_ARG, "precision must be QB_FP64, QB_MIXED or QB_MIXED_INT8RES");
  if (precision == QB_MIXED_INT8RES) precision = QB_MIXED;
  (void)hipSetDevice(ctx->device);
  struct Shape {
    int M, N, K, ty;
  };
  std::vector<Shape> shapes;
  const int n = (int)N;
  const int lu_ty = precision == QB_MIXED ? 0 : 1;
  // LU: trailing updates (K = the panel width), block-row TRSM updates, panel
  // updates over the whole factorisation (the remaining size shrinks from n)
  const int nb = qb_lu_nb(n);
  for (int d : {1, 2, 4, 8, 16})
    shapes.push_back({n / d - (d == 1 ? nb : 0), n / d - (d == 1 ? nb : 0), nb, lu_ty});
  for (int d : {2, 8}) shapes.push_back({128, n / d, 128, lu_ty});
  shapes.push_back({64, n - nb, 64, lu_ty});  // the block-row TRSM's lowest updates
  for (int d : {2, 8, 32})
    for (int w = 128; w >= 32; w /= 2) shapes.push_back({n / d, w, w, lu_ty});
  // triangular solves with r right-hand sides (recursive halves)
  for (int h = n / 2; h >= 32; h /= 2) shapes.push_back({h, r, h, lu_ty});
  // FP64 refinement residual
  if (precision == QB_MIXED) shapes.push_back({n, r, n, 1});
  std::string rep;
  std::vector<char> seen(2 * QB_GEMM_CLASSES, 0);
  for (const Shape& sh : shapes) {
    if (sh.M < 1 || sh.N < 1 || sh.K < 1) continue;
    int cls = qb_gemm_class(sh.M, sh.N, sh.K);
    if (seen[sh.ty * QB_GEMM_CLASSES + cls]) continue;
    seen[sh.ty * QB_GEMM_CLASSES + cls] = 1;
    double tf = 0;
    int v = sh.ty == 0 ? qb_gemm_tune<cf>(ctx, sh.M, sh.N, sh.K, nq, &tf)
                       : qb_gemm_tune<cd>(ctx, sh.M, sh.N, sh.K, nq, &tf);
    if (v < 0) return v;
    char line[200];
    const int vc = ctx->gemm_pick_conc[sh.ty][cls];
    snprintf(line, sizeof(line), "%s %dx%dx%d -> %s (%.1f TF/s)%s%s\n", sh.ty == 0 ? "c64" : "c128",
             sh.M, sh.N, sh.K, qb_gemm_variant_name(sh.ty, v), tf,
             vc != v ? "; beside the lookahead panel: " : "",
             vc != v ? qb_gemm_variant_name(sh.ty, vc) : "");
    rep += line;
  }
  if (report && report_len > 0) {
    size_t m = std::min((size_t)(report_len - 1), rep.size());
    memcpy(report, rep.data(), m);
    report[m] = 0;
  }
  return QB_OK;
}
