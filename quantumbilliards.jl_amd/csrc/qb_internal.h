// Internal declarations of libqbhip (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <functional>
#include <string>
#include <vector>

#include "qb_complex.h"
#include "../../include/qb.h"

#define QB_MAX_COMP 8
#define QB_GEMM_VARIANTS 23
#define QB_GEMM_CLASSES 45

// Boundary geometry resident on the device.  Node arrays are concatenated
// over components; the Kress R matrix and the log term are circulant per
// component, so each is stored as one length-N_c vector indexed by
// (i - j) mod N_c.
struct DevGeom {
  int64_t N;
  int32_t ncomp;
  int64_t off[QB_MAX_COMP + 1];  // component offsets (0-based), off[ncomp] = N
  const double* x;
  const double* y;
  const double* dx;
  const double* dy;
  const double* speed;
  const double* ws;
  const double* kappa;  // curvature / 2pi (boundarypoints.jl:881-882)
  const int32_t* comp;  // component id per node
  const double* rk;     // Kress R circulant columns, concatenated (length N)
  const double* lt;     // log(4 sin^2((t_i - t_j)/2)) circulant columns (length N)
};

// Symmetry orbits of a geometry (symmetry-reduced assembly).  Reduced index
// b stands for the orbit img[b*G + 0 .. b*G + G-1] of full node indices, with
// img[b*G] its representative; entry (a, b) of a reduced matrix is
//   sum_l scale[b*G + l] * A_full(img[a*G], img[b*G + l]).
struct DevSym {
  int64_t m = 0;  // reduced dimension (number of orbits)
  int32_t G = 0;  // orbit size (group order)
  // irreps assembled together (qb_geom_set_orbits_multi): scale holds nirr
  // tables of m x G weights, one per irrep; every assembly then produces
  // nirr matrices per wavenumber, irrep-major (matrix e * nz + q).
  int32_t nirr = 1;
  const int32_t* img = nullptr;
  const cd* scale = nullptr;
  // Pair sharing: when the images come from isometries of the node set
  // (img[b*G + l] = g_l(rep_b) for one g_l per l), |x_{rep_a} - x_{img[b,l]}|
  // = |x_{rep_b} - x_{img[a, inv[l]]}| with g_{inv[l]} = g_l^{-1}, so the
  // specials of one distance serve entry (a, b) and entry (b, a).  pair is set
  // only after that identity was verified for every (a, b, l).
  bool pair = false;
  int8_t inv[64] = {0};
};

struct QbProfRec {
  hipEvent_t a, b;
  int kind;
  double flops, bytes;
};

struct qb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // stream every launch of this context uses
  hipStream_t stream_aux = nullptr;  // second stream for overlapped work (LU lookahead)
  hipEvent_t ev_main = nullptr, ev_aux = nullptr;  // cross-stream ordering
  hipEvent_t ev_first = nullptr;  // end of a first LU panel factored ahead (qb_getrf_first_panel)
  std::string err;
  int64_t bytes = 0;  // device bytes held by objects of this context
  int gemm_pick[2][QB_GEMM_CLASSES];  // autotuned tile variant per shape class (-1: default)
  // The same among the tiles that leave room on each CU for other work: used
  // while gemm_conc is set, i.e. for the LU trailing update that runs beside
  // the lookahead panel factorisation on the second stream.
  int gemm_pick_conc[2][QB_GEMM_CLASSES];
  bool gemm_conc = false;
  // set while a complex128 LU panel runs beside the trailing update: its
  // triangular-solve leaves and GEMMs take the variants that fit on a CU
  // next to two trailing-update workgroups (no LDS, 19 KB of LDS)
  bool panel_lean = false;
  // Node list (device array) for batched launches: when set, batch entry z
  // of a GEMM / TRSM / gather / node-wise helper works on matrix bix[z]
  // (strides unchanged), so a subset of a window's nodes runs in one launch.
  const int* bix = nullptr;
  // hipGraph replay of the batched LU (qb_getrf_graph): instantiated graphs
  // keyed by operand type, buffers and shape; gemm_epoch changes whenever the
  // autotuner changes a tile pick, so stale captures are not replayed.
  bool graphs = false;
  uint64_t gemm_epoch = 0;
  struct GraphRec {
    int ty;
    const void* A;
    int n;
    int64_t lda, sA;
    const int* ipiv;
    int64_t sP;
    const int* info;
    int batch;
    uint64_t epoch;
    hipGraphExec_t exec;
  };
  std::vector<GraphRec> graph_cache;
  // caching allocator: freed blocks are kept (by exact size) for reuse
  std::vector<std::pair<size_t, void*>> cache;
  int64_t cache_bytes = 0;
  int64_t cache_limit = (int64_t)96 << 30;
  bool prof = false;
  std::vector<QbProfRec> prof_recs;      // recorded, not yet reduced
  std::vector<hipEvent_t> prof_pool;     // reusable events
  double prof_ms[QB_K_COUNT] = {0};
  int64_t prof_n[QB_K_COUNT] = {0};
  double prof_flops[QB_K_COUNT] = {0};
  double prof_bytes[QB_K_COUNT] = {0};
  // operations issued to the matrix cores (GEMM classes): 6 M N K for the
  // three-product tiles, 8 M N K otherwise; the int8 products as launched
  double prof_issued[QB_K_COUNT] = {0};
};

// Launch accounting: qb_prof_begin records an event before a launch and
// returns its slot (or -1 when profiling is off); qb_prof_end records the
// closing event.
int qb_prof_begin(qb_ctx* ctx, int kind, double flops, double bytes);
void qb_prof_end(qb_ctx* ctx, int slot);

struct qb_geom {
  qb_ctx* ctx = nullptr;
  DevGeom d{};
  void* buf = nullptr;  // one allocation holding all arrays
  int64_t bytes = 0;
  double rmin_geom = 0, rmax_geom = 0;  // padded (x0.95, x1.05) min / max distance
  double rmin_dlp = 0, rmax_dlp = 0;    // same over the first component only
  DevSym sym{};          // set by qb_geom_set_orbits (m = 0: no reduction)
  void* sym_buf = nullptr;
  int64_t sym_bytes = 0;
};

// Dimension of the matrices assembled on a geometry: the number of orbits
// when symmetry orbits are set, else the number of nodes.
static inline int64_t qb_dim(const qb_geom* g) { return g->sym.m > 0 ? g->sym.m : g->d.N; }
// Matrices per wavenumber of an assembly on g (irreps assembled together).
static inline int qb_nirr(const qb_geom* g) { return g->sym.m > 0 ? g->sym.nirr : 1; }
#define QB_MAX_IRREPS 16

// Chebyshev tables of one or more wavenumbers, resident on the device.
// H: [nz][np_h][M_h+1] of {H0 coeff, H1 coeff}; J: [nz][np_j][M_j+1] of
// {J0 coeff, J1 coeff}.  Panels are uniform on [rmin_h, rmax] (Hankel) and
// [0, rmax] (J), built exactly like plan_h / plan_j.
struct DevTables {
  int64_t nz = 0;
  int32_t np_h = 0, M_h = 0, np_j = 0, M_j = 0;
  double rmin_h = 0, rmax = 0;
  cd* H = nullptr;
  cd* J = nullptr;
};

// error helpers ------------------------------------------------------------
int qb_fail(qb_ctx* ctx, int code, const std::string& msg);
#define QB_HIP(ctx, expr)                                                   \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess)                                                   \
      return qb_fail(ctx, QB_E_HIP, std::string(#expr) + ": " +             \
                                        hipGetErrorString(_e));             \
  } while (0)

int qb_dev_alloc(qb_ctx* ctx, void** p, size_t bytes);
void qb_dev_free(qb_ctx* ctx, void* p, size_t bytes);

// tables (qb_tables.hip) ----------------------------------------------------
int qb_tables_build(qb_ctx* ctx, const double* z_dev_or_host, int64_t nz,
                    double rmin_h, double rmax, const qb_cheb_params& cp,
                    bool want_h0j0, DevTables* t);
void qb_tables_free(qb_ctx* ctx, DevTables* t);
// Bytes of the H and J tables of nz wavenumbers (hbytes: the H part).
size_t qb_tables_bytes(int64_t nz, const qb_cheb_params& cp, size_t* hbytes);
// The same tables built into caller memory (mem_bytes >= qb_tables_bytes):
// H first, J after it.  Allocates nothing; the tables are not freed.
int qb_tables_build_in(qb_ctx* ctx, const double* z_dev, int64_t nz, double rmin_h, double rmax,
                       const qb_cheb_params& cp, void* mem, size_t mem_bytes, DevTables* t);
void qb_lobatto_host(int M, double* t, double* C);  // nodes cospi(j/M), DCT matrix

// assembly (qb_assemble.hip) -------------------------------------------------
// Writes T(z_q) for q in [0, nz) into out64 (ld = N, stride N*N), and, if
// out32 != nullptr, a complex64 copy into out32.
// T(z_q), dT/dk and d2T/dk2 for q in [0, nz) (same layout as qb_assemble_dev).
// ref_sign: the reference's sign of the between-component double-layer
// derivative terms (qb_assemble_deriv, QB_DERIV_REFERENCE_SIGN).
int qb_assemble_deriv_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev, int64_t nz,
                          const DevTables& t, cd* out, cd* dout, cd* ddout, bool ref_sign);
int qb_assemble_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev,
                    int64_t nz, const DevTables& t, cd* out64, cf* out32);
// The same two for a geometry with symmetry orbits: m x m reduced matrices
// (ld = m, stride m*m).  qb_assemble_dev / qb_assemble_deriv_dev call these
// when g->sym.m > 0.
// Both matrices of every conjugate node pair from the Hankel tables alone
// (windows whose contour is symmetric about the real axis); returns 1, having
// done nothing, when the nodes or tables do not allow it.
int qb_assemble_conj_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_host,
                         const cd* z_dev, int64_t nz, const DevTables& t, cd* out64,
                         cf* out32);
// The same matrices from envelope tables of H_n e^{-ikr} held in LDS (no
// uniform tables needed; rmin, rmax: the distance range of the geometry);
// returns 1, having done nothing, when the nodes do not pair up.  With
// `between` set and cols_first a multiple of 64 below N, columns [0,
// cols_first) of every matrix (and rows [0, cols_first)) are assembled by a
// first launch, then between() is called (to enqueue work on those columns
// on another stream) and the rest is assembled; the rest never writes those
// columns.
int qb_assemble_env_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_host,
                        const cd* z_dev, int64_t nz, double rmin, double rmax, cd* out64,
                        cf* out32, int cols_first = 0,
                        const std::function<int()>& between = std::function<int()>());
// Pairs (a, b) of wavenumbers with z_b = conj(z_a) (a real one with itself),
// flattened; false when some wavenumber has no partner.
bool qb_conj_pairs(const cd* z_host, int64_t nz, std::vector<int>* pr);
int qb_assemble_red_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev, int64_t nz,
                        const DevTables& t, cd* out64, cf* out32);
// Checks |x_{rep_a} - x_{img[b,l]}| = |x_{rep_b} - x_{img[a,inv[l]]}| to
// 1e-13 r + floor_abs for every (a, b, l) on the device; ok = 1 if it holds.
int qb_orbit_pairs_ok(qb_ctx* ctx, const DevGeom& g, const DevSym& s, double floor_abs, int* ok);
int qb_assemble_red_deriv_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev,
                              int64_t nz, const DevTables& t, cd* out, cd* dout, cd* ddout,
                              double off_sign);

// bessel (qb_tables.hip)
int qb_bessel_eval_dev(qb_ctx* ctx, const cd* z, int64_t n, cd* out4);
int qb_cheb_check_dev(qb_ctx* ctx, const DevTables& t, const cd* z_dev,
                      int64_t nz, int32_t samples, double* err4_host);
