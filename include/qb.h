/* C ABI of libqbhip: the MI355X (gfx950) boundary-integral hot path of
 * QuantumBilliards.jl — DLP-/CFIE-Kress Chebyshev assembly and the Beyn
 * contour loop.
 *
 * Conventions
 *  - Plain C types only.  Complex numbers are interleaved (re, im) doubles,
 *    layout-identical to Julia's ComplexF64 and C's double _Complex.
 *  - Matrices are column-major with leading dimension N (Julia's layout).
 *  - Every call returns QB_OK (0) or a negative QB_E* code; the message of the
 *    last failing call on a context is returned by qb_last_error(ctx).
 *  - A context owns one HIP stream on one device.  Calls on one context must
 *    not overlap; separate contexts are independent (no global state).
 */
#ifndef QB_H
#define QB_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qb_ctx qb_ctx;
typedef struct qb_geom qb_geom;
typedef struct qb_window qb_window;

enum {
  QB_OK = 0,
  QB_E_ARG = -1,      /* invalid argument (sizes, pointers, parameters)    */
  QB_E_HIP = -2,      /* HIP runtime or kernel launch failure              */
  QB_E_NOMEM = -3,    /* device allocation failed / over the memory budget */
  QB_E_STATE = -4,    /* object used in the wrong state                    */
  QB_E_NUMERIC = -5   /* numerical failure reported by a kernel            */
};

/* operators: the two Kress solver families of the reference */
enum { QB_CFIE_KRESS = 1, QB_DLP_KRESS = 2 };

/* precision of the Beyn LU/solves
 *  QB_FP64           complex128 LU and solves
 *  QB_MIXED          complex64 LU on MFMA, iterative refinement against an
 *                    FP64 residual computed as a complex128 GEMM on the FP64
 *                    matrices kept with the window (24 bytes per matrix entry)
 *  QB_MIXED_INT8RES  the same refinement with the FP64 residual formed
 *                    exactly on the int8 matrix cores from residues of T kept
 *                    with the window (56 bytes per matrix entry; N <= 65535,
 *                    else as QB_MIXED) */
enum { QB_FP64 = 0, QB_MIXED = 1, QB_MIXED_INT8RES = 2 };

/* flags for qb_assemble / qb_assemble_deriv
 *  QB_OUT_DEVICE           out (and dout, ddout) are device pointers
 *  QB_DERIV_REFERENCE_SIGN qb_assemble_deriv: the reference's sign of the
 *                          between-component double-layer derivative terms
 *                          (cheb-CFIE:1031-1032); the default, kept as an
 *                          explicit spelling of it
 *  QB_ASM_CONJ_PAIRS       qb_assemble: take the conjugate-pair kernel that
 *                          Beyn windows with real k0 use: both matrices of a
 *                          pair z, conj(z) together, H_n from envelope tables
 *                          of H_n(zr) e^{-izr} (quarter-octave Chebyshev
 *                          panels of degree 12, ~1e-14 relative; the table
 *                          sizes in cheb are then not used) and
 *                          J_n(zr) = (H_n(zr) + conj H_n(conj(z) r)) / 2;
 *                          QB_E_ARG when the nodes are not closed under
 *                          conjugation or the geometry has symmetry orbits
 *  QB_DERIV_MATRIX_SIGN    qb_assemble_deriv: the wavenumber derivative of the
 *                          matrix between components (the opposite sign of
 *                          the reference's two terms) */
enum {
  QB_OUT_DEVICE = 1,
  QB_DERIV_REFERENCE_SIGN = 2,
  QB_ASM_CONJ_PAIRS = 4,
  QB_DERIV_MATRIX_SIGN = 8
};

/* Chebyshev table sizes: the n_panels_h/M_h/n_panels_j/M_j keywords of
 * construct_boundary_matrices! (chebyshev_euclidian_helmholtz_cfie_kress.jl:1431,
 * chebyshev_euclidian_helmholtz_dlp_kress.jl:1110). */
typedef struct {
  int32_t n_panels_h, M_h, n_panels_j, M_j;
} qb_cheb_params;

/* Per-window statistics returned by qb_window_create / qb_window_moments. */
typedef struct {
  int64_t N;
  int32_t nq;
  int32_t r;
  int32_t precision;
  int32_t refine_iters_max;     /* max refinement steps over nodes (mixed)   */
  int32_t nodes_fp64_fallback;  /* nodes refactored in FP64 (mixed)          */
  double refine_resid_max;      /* max final backward error of refined nodes */
  double t_assemble_ms, t_factor_ms, t_solve_ms;
  double rmin, rmax;            /* padded distance range of the tables       */
} qb_stats;

/* Context (one device, one stream).  Replaces no reference function: the
 * reference runs in one Julia process with no device. */
int qb_ctx_create(int device, qb_ctx** out);
void qb_ctx_destroy(qb_ctx* ctx);
const char* qb_last_error(const qb_ctx* ctx);
/* Bytes currently held on the device by this context (objects plus the
 * allocation cache that keeps freed buffers for windows of the same size). */
int64_t qb_ctx_device_bytes(const qb_ctx* ctx);

/* Replay the batched LU of windows through cached hipGraphs (on != 0): the
 * first window with given buffers and shape captures the launch sequence of
 * both streams, later ones replay it; launch accounting bypasses the graphs.
 * Off by default.  No reference counterpart. */
int qb_ctx_set_graphs(qb_ctx* ctx, int32_t on);

/* Free and total device memory of the context's GPU (hipMemGetInfo); the
 * serving loop sizes its admission budget from it.  No reference
 * counterpart. */
int qb_ctx_mem_info(qb_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);
/* Page-locked host memory (hipHostMalloc) for probe blocks and results: the
 * library's host<->device copies from and to it run as direct DMA instead of
 * being staged through a bounce buffer.  Free with qb_host_free.  No reference
 * counterpart. */
int qb_host_alloc(qb_ctx* ctx, int64_t bytes, void** out);
void qb_host_free(void* p);
/* Set the cache limit in bytes (0 releases the cache and disables it). */
int qb_ctx_set_cache_limit(qb_ctx* ctx, int64_t bytes);

/* Kernel classes for the per-context launch accounting below. */
enum {
  QB_K_TABLES = 0,    /* Chebyshev table construction          */
  QB_K_ASSEMBLE = 1,  /* CFIE/DLP-Kress assembly               */
  QB_K_GEMM_C64 = 2,  /* complex64 C -= A*B (MFMA)             */
  QB_K_GEMM_C128 = 3, /* complex128 C -= A*B (MFMA)            */
  QB_K_PANEL = 4,     /* leaf panel factorisation              */
  QB_K_TRSM = 5,      /* triangular-solve leaves               */
  QB_K_OTHER = 6,     /* swaps, gathers, norms, moments        */
  QB_K_OZAKI = 7,     /* int8 products of the refinement residual (MFMA) */
  QB_K_GEMM_C64_BF16 = 8, /* complex64 C -= A*B on bf16 MFMA (split operands) */
  QB_K_COUNT = 9
};
/* Enable (1) / disable (0) recording a pair of HIP events on the context's
 * stream around every launch, and clear the counters.  No counterpart in the
 * reference (its @benchit timings are wall-clock, macros.jl:52-91). */
int qb_ctx_set_profiling(qb_ctx* ctx, int32_t enable);
/* Totals since profiling was enabled for one kernel class: device time (ms,
 * from the events), launches, and algorithmic flops / bytes of those
 * launches (complex GEMM: 8 M N K flops; assembly: 16 B per entry written;
 * int8 residual products: 2 ops per int8 multiply-add, 64 real products of
 * M N K per complex product, and the int8 operand bytes). */
int qb_ctx_kernel_stats(qb_ctx* ctx, int32_t kind, double* total_ms, int64_t* launches,
                        double* flops, double* bytes);
/* Operations the launches of one class issued to the matrix cores since
 * profiling was enabled: for the complex GEMMs 6 M N K per launch with the
 * three-product (Gauss) tiles and 8 M N K otherwise (the algorithmic count
 * of qb_ctx_kernel_stats is always 8 M N K); for QB_K_OZAKI the int8
 * operations of the padded products as launched; for QB_K_GEMM_C64_BF16 the
 * bf16 operations (48 M N K: 4 real products of 6 split products each).  No
 * reference counterpart. */
int qb_ctx_kernel_issued(qb_ctx* ctx, int32_t kind, double* issued);

/* Tile autotuning: times each MFMA GEMM tile variant on the shapes a window
 * of size N with nq nodes and r probes launches (LU trailing/panel updates,
 * triangular-solve halves, the FP64 refinement residual) and keeps the
 * fastest per shape class for every later launch on this context.  report
 * (may be NULL) receives one line per tuned shape.  No counterpart in the
 * reference (its BLAS picks kernels internally). */
int qb_ctx_autotune(qb_ctx* ctx, int64_t N, int32_t nq, int32_t r, int32_t precision,
                    char* report, int64_t report_len);

/* Boundary discretisation, one or more periodic components, holes already
 * orientation-reversed by the caller.  Replaces the inputs the reference
 * builds in evaluate_points (cfie.jl:822-832, dlp_kress.jl:563-606) and the
 * CFIEGeomCache / Kress R it derives from them (boundarypoints.jl:310-365,
 * :854-884): x, y = xy; dx, dy = tangent; ddx, ddy = tangent_2; ts, ws as in
 * BoundaryPointsCFIE (boundarypoints.jl:487-502), all components concatenated.
 * Each component must carry the uniform Kress grid (ts spaced 2pi/N_c). */
int qb_geom_create(qb_ctx* ctx, int32_t ncomp, const int64_t* n_per_comp,
                   const double* x, const double* y, const double* dx,
                   const double* dy, const double* ddx, const double* ddy,
                   const double* ts, const double* ws, qb_geom** out);
void qb_geom_destroy(qb_geom* g);
int64_t qb_geom_size(const qb_geom* g);

/* Symmetry reduction of a geometry: the reduced workspaces of the reference
 * (build_CFIE_kress_reduced_workspace, cheb-CFIE:475-505, with the orbits of
 * build_symmetry_maps / apply_symmetry_orbit_from_maps, symmetry.jl:746-833;
 * build_dlp_kress_reduced_workspace, dlp_kress.jl:702-708, with the index
 * orbits of periodic_symmetry_index_orbits, symmetry.jl:214-323).  m orbits of
 * G nodes each: img[b*G + l] (0-based node index) is image l of orbit b, the
 * first one its representative (Ifund[b]), and scale[b*G + l] (complex) its
 * weight (fund_to_scale[b][l]; 1 for l = 0).  The orbits must partition the
 * nodes.  After this call every assembly on g (qb_assemble,
 * qb_assemble_deriv, qb_residuals, qb_window_*) produces the m x m reduced
 * matrices of _all_k_symm_CFIE_chebyshev! (cheb-CFIE:698-824, its derivative
 * twin :1087-1243) and of the reduced DLP paths (cheb-DLP:771-880, :955-1050):
 *   A_red(a, b) = sum_l scale[b*G+l] A(img[a*G], img[b*G+l])
 * with A the full operator (identity included).  m = 0 removes the reduction. */
int qb_geom_set_orbits(qb_ctx* ctx, qb_geom* g, int64_t m, int32_t G, const int64_t* img,
                       const double* scale);
/* The same orbits with nirr weight tables (scale: nirr x m x G complex, one
 * table per irrep of the group, e.g. the n characters of a C_n rotation): the
 * image terms of every entry are evaluated once and summed into all nirr
 * reduced matrices.  qb_assemble then writes nirr * nz matrices, matrix
 * e * nz + q for irrep e and wavenumber q.  A window on such a geometry holds
 * all irreps of its contour nodes and factors and solves them as one batch;
 * qb_window_moments then returns nirr blocks of A0 and of A1 (N x r each, one
 * after the other), qb_window_moments_qr nirr blocks of R and of C (r x r
 * each), and qb_window_basis_irrep forms U_k of one irrep; the admission
 * footprint is qb_window_footprint(N, nq * nirr, ...).  No reference
 * counterpart (the reference assembles one irrep per solver); derivative
 * assembly and qb_residuals take one irrep.  1 <= nirr <= 16. */
int qb_geom_set_orbits_multi(qb_ctx* ctx, qb_geom* g, int64_t m, int32_t G, int32_t nirr,
                             const int64_t* img, const double* scales);
/* Irreps assembled together on g (1 without a multi-irrep orbit table). */
int32_t qb_geom_irreps(const qb_geom* g);
/* Matrix dimension of assemblies on g: m after qb_geom_set_orbits, else N
 * (boundary_matrix_size_for_solver, ebim.jl:26-40). */
int64_t qb_geom_dim(const qb_geom* g);

/* construct_boundary_matrices!(Tbufs, solver, pts, zj; n_panels_h, M_h,
 * n_panels_j, M_j) for CFIE_kress (cheb-CFIE:1431-1450) and DLP_kress
 * (cheb-DLP:1110-1124).  Writes nz dense N x N complex128 matrices
 * out[q*N*N + i + j*N] = T(z_q)[i,j] (overwrites, like the reference).  With
 * flags & QB_OUT_DEVICE, out is a device pointer; otherwise host memory.
 * rminmax (may be NULL) receives the padded table range [rmin, rmax]. */
int qb_assemble(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* z,
                int64_t nz, const qb_cheb_params* cheb, double* out,
                int32_t flags, double* rminmax);

/* qb_assemble that allocates no device memory: for callers that cannot
 * allocate (code that runs under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * z_dev (nz complex) and out_dev (the nz * irreps matrices of qb_assemble's
 * layout) are device memory the caller owns; flags must be 0 (the J-table
 * kernel; no conjugate pairing).  The Chebyshev tables, the only memory
 * beyond the inputs and outputs, are built in a fixed device array of
 * QB_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on
 * until the process ends.  They take
 *   nz * (n_panels_h * (M_h + 1) + n_panels_j * (M_j + 1)) * 32 bytes;
 * a call that needs more is rejected with QB_E_ARG (assemble fewer
 * wavenumbers per call).  Every call on one GPU shares that GPU's array and
 * overwrites the tables it reads (nothing needs clearing): calls that overlap
 * on one GPU, from different contexts or threads, race and must be
 * serialised by the caller.  Complete on return, like qb_assemble. */
#define QB_NOALLOC_SCRATCH_BYTES ((int64_t)128 << 20)
int qb_assemble_noalloc(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* z_dev,
                        int64_t nz, const qb_cheb_params* cheb, double* out_dev,
                        int32_t flags, double* rminmax);

/* The first half of construct_B_matrix (beyn.jl:257-277): contour nodes
 * z_j = k0 + R e^{i 2pi j/nq}, assembly of T(z_j), LU of every node.  The
 * factors stay resident in the window so the moment solve can be repeated
 * with more probes (the r -> 2r retry of beyn.jl:291-309).  precision:
 * QB_FP64 (plain complex128 path) or QB_MIXED (complex64 LU on MFMA + FP64
 * iterative refinement against the resident FP64 T).  With k0 real the nodes
 * come in conjugate pairs, and the window assembles each pair together with
 * the kernel of QB_ASM_CONJ_PAIRS (envelope tables in LDS; H_n to ~1e-14
 * relative, J_n = (H_n(kr) + conj H_n(conj(k) r)) / 2), where cheb's table
 * sizes are not used; other contours use the uniform H and J tables of cheb,
 * as the reference does. */
int qb_window_create(qb_ctx* ctx, const qb_geom* g, int32_t op, double k0_re,
                     double k0_im, double R, int32_t nq,
                     const qb_cheb_params* cheb, int32_t precision,
                     qb_window** out, qb_stats* stats);
/* Several disks on one geometry as one window (continuous batching of
 * requests that share N): contour c is k0[2c] + i k0[2c+1] + R[c] e^{i 2pi j/nq},
 * and the nwin * nq matrices (times the irreps of a multi-irrep geometry) are
 * assembled, factored and solved as one batch with one probe block V.  The
 * window then has qb_window_blocks(w) = nirr * nwin moment blocks, block
 * e * nwin + c for irrep e and contour c: qb_window_moments returns that many
 * blocks of A0 and of A1, qb_window_moments_qr that many blocks of R and of
 * C, and qb_window_basis_irrep takes the block index.  The admission
 * footprint is qb_window_footprint(N, nq * nwin * nirr, ...).  Per block the
 * result is that of a window of its own.  No reference counterpart: the
 * reference's driver (compute_spectrum_beyn, beyn.jl:778-921) runs its disks
 * one after another.  nq * nwin * nirr <= 4096. */
int qb_window_create_multi(qb_ctx* ctx, const qb_geom* g, int32_t op, int32_t nwin,
                           const double* k0, const double* R, int32_t nq,
                           const qb_cheb_params* cheb, int32_t precision, qb_window** out,
                           qb_stats* stats);
/* Moment blocks of a window: irreps x contours (1 for a plain window). */
int32_t qb_window_blocks(const qb_window* w);
/* Columns [j0, j0 + ncols) of T(z) of matrix `node` of a window, exactly as
 * the window assembled it (node = (irrep * nwin + contour) * nq + j for
 * contour node j), copied column-major (N x ncols, complex128) into host
 * memory.  Only QB_MIXED windows keep the complex128 matrices resident (for
 * the refinement residual); a QB_FP64 window holds their LU factors instead
 * and returns QB_E_STATE.  No reference counterpart: the reference's Tbufs
 * (beyn.jl:263) are matrices its caller can read. */
int qb_window_matrix_cols(qb_window* w, int32_t node, int64_t j0, int64_t ncols, double* out);
void qb_window_destroy(qb_window* w);
/* Device bytes a window of this size would hold (the admission footprint):
 * matrices, pivots, solve buffers for r probes, and up to 32 MiB of tables
 * per node. */
int64_t qb_window_footprint(int64_t N, int32_t nq, int32_t precision, int32_t r);

/* The moment loop of construct_B_matrix (beyn.jl:278-285): X_j = T(z_j)^{-1} V,
 * A0 = sum_j w_j X_j, A1 = sum_j w_j z_j X_j, w_j = (R/nq) e^{i 2pi j/nq}.
 * V, A0, A1: host N x r complex128.  refine_tol: backward error
 * ||V - T_j X_j||_max / (||T_j||_inf min(||X_j||_max, x_ref) + ||V||_max) at
 * which the mixed-precision refinement of node j stops, x_ref the median of
 * ||X_j||_max over the nodes after the first solve (so that every node's
 * error is small against the typical node's solution, which the moments
 * add); a node whose error stops halving, or whose target lies below the
 * resolution of an FP64 residual, is refactored and solved in FP64 (ignored
 * for QB_FP64). */
int qb_window_moments(qb_window* w, int32_t r, const double* V, double* A0,
                      double* A1, double refine_tol, qb_stats* stats);

/* construct_boundary_matrices_with_derivatives!: T(z_q), dT/dk(z_q) and
 * d2T/dk2(z_q) for the expanded boundary integral method, each in the layout
 * of qb_assemble (CFIE: cheb-CFIE:1507-1530 with the entries of
 * _all_k_nosymm_CFIE_chebyshev_deriv!, :881-1059; DLP: cheb-DLP:1126-1144
 * with _construct_dlp_kress_matrices_derivatives_chebyshev!, :881-952, and
 * its sign flip + identity, :1097-1108).  Between boundary components the
 * two double-layer derivative terms carry the reference's sign
 * (cheb-CFIE:1031-1032) by default, so that the drop-in returns the
 * reference's numbers; with QB_DERIV_MATRIX_SIGN in flags they carry the
 * opposite sign, which makes dT/dk the derivative of T there. */
int qb_assemble_deriv(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* z, int64_t nz,
                      const qb_cheb_params* cheb, double* out, double* dout, double* ddout,
                      int32_t flags, double* rminmax);

/* Residual check of candidate eigenpairs (the spurious-root filter): for
 * q < nl, with T_q = T(lambda_q) assembled on the device exactly as by
 * qb_assemble (same tables, same entries) and phi_q column q of phi,
 *   res2[q] = ||T_q phi_q||_2,  resn1[q] = ||T_q phi_q||_1,  resninf[q] = ||T_q phi_q||_inf,
 *   opn1[q] = ||T_q||_1 (max column sum),  opninf[q] = ||T_q||_inf (max row sum).
 * lam: nl complex; phi: host N x nl complex128, column-major.  Outputs may be
 * NULL.  The matrices are assembled `batch` at a time (0: as many as fit in
 * 8 GiB).  Replaces construct_boundary_matrices! + mul!(y, Tbufs[j], phi) +
 * norm/opnorm in residual_and_norm_select (beyn.jl:557-593) and in
 * imag_k_check_EXPERIMENTAL (beyn.jl:666-687). */
int qb_residuals(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* lam, int64_t nl,
                 const qb_cheb_params* cheb, const double* phi, int32_t batch, double* res2,
                 double* resn1, double* resninf, double* opn1, double* opninf);

/* qb_window_moments followed by a Householder QR of A0 on the device: the
 * N x r blocks stay there, A0 = Q R, and only R (r x r, upper triangular) and
 * C = (Q^H A1)[:r, :] (r x r) are returned, column-major.  svd(R) = U_R S W^H
 * gives the singular values and W of svd(A0) (beyn.jl:286) and
 * B = U_R[:, :k]^H C W[:, :k] S_k^{-1} (beyn.jl:310-321) without moving A0, A1
 * to the host.  The reflectors are kept for qb_window_basis. */
int qb_window_moments_qr(qb_window* w, int32_t r, const double* V, double* R, double* C,
                         double refine_tol, qb_stats* stats);

/* U_k = Q [Ur; 0] (host N x k) from the last qb_window_moments_qr, Ur host
 * r x k: the U_k that construct_B_matrix returns (beyn.jl:311). */
int qb_window_basis(qb_window* w, const double* Ur, int32_t k, double* U);
/* The same for moment block e (irrep e of a window on a multi-irrep geometry,
 * qb_geom_set_orbits_multi; block e * nwin + c of qb_window_create_multi);
 * qb_window_basis is block 0. */
int qb_window_basis_irrep(qb_window* w, int32_t irrep, const double* Ur, int32_t k, double* U);

/* Device evaluation of H0^(1), H1^(1), J0, J1 at n complex points (the values
 * the reference takes from SpecialFunctions.besselh/besselj when sampling its
 * tables, chebyshev_bessels.jl:331-372).  Host arrays of n complex each. */
int qb_bessel_eval(qb_ctx* ctx, const double* z, int64_t n, double* h0,
                   double* h1, double* j0, double* j1);

/* Table-size autotuning of chebyshev_params (chebyshev_optimal_panelization.jl:221-262):
 * grow panels x1.5 (or M += 2 every 5th iteration) until the max abs error of
 * every H0/H1/J0/J1 table over `samples` uniform radii is below tol.
 * cheb holds the initial sizes and receives the chosen ones; errs (may be
 * NULL) receives the 4 worst errors.  Returns QB_OK also when max_iter is
 * reached without meeting tol (best effort, like the reference). */
int qb_cheb_tune(qb_ctx* ctx, const qb_geom* g, int32_t op, const double* z,
                 int64_t nz, double tol, int32_t samples, int32_t max_iter,
                 qb_cheb_params* cheb, double* errs);

#ifdef __cplusplus
}
#endif
#endif
