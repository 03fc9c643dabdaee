/*
 * simpeg_dc.h - C ABI of the MI355X (gfx950) nodal DC-resistivity library (libsimpeg_dc.so).
 *
 * The library replaces, for SimPEG's Simulation3DNodal, the assembled sparse operator and the
 * direct solver behind it (pymatsolver's factorisation: `self.solver(A)` and `Ainv * RHS`) with a
 * matrix-free 7-point stencil, a geometric-multigrid-preconditioned block CG (fp32 inside fp64
 * iterative refinement, or plain fp64), and hand-written kernels for the model derivatives.
 * File:line citations are relative to the simpeg repository root.
 *
 * Conventions
 *  - All array arguments named *_dev are device pointers (HIP global memory); the library never
 *    frees caller memory.  Host arrays are only read during the call.
 *  - A batch of node vectors ("columns", one per source or per adjoint right-hand side) is stored
 *    column-major with column stride sdc_ld(h) doubles/floats: column c, node n lives at
 *    base[c * ld + n], node n = i + (nx+1) * (j + (ny+1) * k) (x fastest, the discretize ordering).
 *  - Column bases (base pointers of a batch) must be 32-byte aligned (sdc_solve checks it);
 *    sdc_ld(h) is a multiple of 64, so every column of an aligned batch is aligned too.
 *  - Every call is enqueued on `stream` (a hipStream_t; NULL = default stream).  Calls that
 *    return host-visible results (sdc_solve statistics) synchronise that stream.
 *  - Every function returns 0 on success or a nonzero SDC_ERR_* code; sdc_last_error() gives the
 *    message for the calling thread.  A handle must not be used from two threads at once, and
 *    calls that overlap on one handle race on its workspace.
 */
#ifndef SIMPEG_DC_H
#define SIMPEG_DC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDC_OK 0
#define SDC_ERR_ARG 1      /* invalid argument (shape, size, null pointer, bad enum) */
#define SDC_ERR_HIP 2      /* HIP runtime error */
#define SDC_ERR_NOCONV 3   /* solver did not reach the requested tolerance */
#define SDC_ERR_STATE 4    /* call order error (e.g. solve before sdc_set_sigma) */
#define SDC_ERR_ALLOC 5    /* device allocation failed */

#define SDC_BC_ROBIN 0     /* Robin/Mixed boundary (reference default) */
#define SDC_BC_NEUMANN 1   /* Neumann: row 0 of A replaced by e_0 */

#define SDC_F32 0
#define SDC_F64 1

#define SDC_MODE_MIXED 0     /* fp32 GMG-PCG with fp64 reliable updates (true residual and
                                solution accumulation in fp64, Krylov space kept); when the fp32
                                phase stalls, fp64 GMG-PCG continues from its result, and later
                                mixed solves on the same operator run fp64 from the start */
#define SDC_MODE_FP64 1      /* plain fp64 GMG-PCG */
#define SDC_MODE_MIXED_IR 2  /* fp32 GMG-PCG restarted inside classic fp64 iterative refinement */

typedef struct sdc_handle_s* sdc_handle;

/* Mesh and boundary description (host arrays).  Replaces the discretize TensorMesh operators
 * used by Simulation3DNodal.getA and setBC (electromagnetics/static/resistivity/simulation.py:
 * 523-555, 573-644): widths of the cells along each axis and, for bc_type Robin, the boundary
 * face weights beta_f = alpha_f * |f| / 4 of the term _AvgBC = AvgN2Fb^T diag(alpha |f|) AvgCC2Fb
 * on each of the six boundary planes (0 on surface faces). */
typedef struct {
  int32_t nx, ny, nz;          /* cells per axis (>= 1) */
  const double* hx;            /* nx widths */
  const double* hy;            /* ny widths */
  const double* hz;            /* nz widths */
  int32_t bc_type;             /* SDC_BC_ROBIN or SDC_BC_NEUMANN */
  const double* beta_xmin;     /* ny*nz, index j + ny*k  (ignored for Neumann) */
  const double* beta_xmax;
  const double* beta_ymin;     /* nx*nz, index i + nx*k */
  const double* beta_ymax;
  const double* beta_zmin;     /* nx*ny, index i + nx*j */
  const double* beta_zmax;
  int32_t coarse_max;          /* multigrid: coarsen until at most this many nodes (0 = 64) */
  int32_t device;              /* HIP device ordinal */
} sdc_mesh_desc;

/* Multigrid / Krylov tuning (defaults from sdc_default_params). */
typedef struct {
  int32_t nu_pre;              /* Chebyshev pre-smoothing steps (>= 1) */
  int32_t nu_post;             /* Chebyshev post-smoothing steps (>= 1) */
  double cheb_lmax;            /* upper bound of the spectrum of D^-1 A (Gershgorin: 2) */
  double cheb_ratio;           /* lower bound = cheb_lmax / cheb_ratio */
  int32_t rhs_per_thread;      /* stencil kernels: 0 = automatic (4), else 1, 2, 4 or 8 columns
                                  per work-item */
  int32_t check_every;         /* PCG iterations recorded per graph replay (or launched) between
                                  two host reads of the device-side stop flag */
  double reliable_delta;       /* mixed mode: replace the fp32 residual by the fp64 true residual
                                  once it dropped by this factor (0 < delta < 1) */
  int32_t xfer_rhs_per_thread; /* grid transfers: 0 = automatic (8), else 1, 2, 4 or 8 */
  int32_t xcd_remap;           /* 1: give each XCD a contiguous range of work-groups, 0: off */
  int32_t vcycle_bf16;         /* 1: the fp32 solver keeps its multigrid vectors, including the
                                  preconditioned residual z, in bfloat16 (arithmetic stays fp32);
                                  2: the same with z in fp32; 3: as 2, and the finest level
                                  reads a bfloat16 copy of the residual; 4: as 1, and the finest
                                  level reads a bfloat16 copy of the residual; 0: fp32 vectors */
  int32_t xfer_separable;      /* restriction: 1 = one pass per coarsened axis, 0 = one 27-point
                                  gather pass */
  int32_t wcycle_level;        /* 0: V-cycle; l >= 1: the coarse correction of level l runs twice
                                  (W-cycle on that level); -1: the first level with at most 1/4
                                  of the finest level's nodes */
  int32_t fused_tail;          /* 1: the fp32 solver runs the small coarse levels that fit into
                                  LDS as one launch (one work-group per column); 0: off */
  int32_t refine_rhs_per_thread; /* fp64 true-residual kernel: 0 = 8, else 1, 2, 4 or 8 columns
                                    per work-item */
  int32_t graph;               /* 1: the solvers record their finest-level V-cycle once per
                                  (workspace, column count) as a HIP graph and replay it every
                                  iteration (not while sdc_profile brackets one of its kernel
                                  classes); 0: launches */
  int32_t interp;              /* multigrid interpolation: 1 = operator-dependent (a node between
                                  two coarse nodes weights them by the conductances of its two
                                  edges on its grid line), 0 = geometric (linear in position) */
  int32_t wcycle_smooth;       /* smoothing steps between the two coarse corrections of the W
                                  step (0 = nu_post; default 1, measured 1.4 % faster than 2 at
                                  the same iteration counts) */
  int32_t nu_coarse;           /* Chebyshev steps before and after the coarse correction on the
                                  levels below the finest (0 = nu_pre / nu_post) */
  int32_t deriv_robin;         /* Robin handles: 1 (default) = the derivative entry points
                                  (sdc_jvec_rhs, sdc_jtvec_contract, sdc_jtj_diag_accumulate)
                                  include the boundary term d(AvgBC sigma)/dsigma; 0 = they leave
                                  it out, as the reference's getADeriv does when the model is a
                                  resistivity (rhoMap; resistivity/simulation.py:566-570) */
  int32_t col_loop;            /* stencil kernels of the fp32 solver: 1 = each work-item takes
                                  rhs_per_thread columns; 8 or 16 = each work-item loops over
                                  that many columns one at a time with its node's coefficients
                                  held (0 = automatic) */
} sdc_params;

typedef struct {
  int32_t mode;                /* SDC_MODE_MIXED or SDC_MODE_FP64 */
  double tol;                  /* per column target |q - A u|_2 / |q|_2 (fp64 residual) */
  double inner_tol;            /* mixed IR: relative residual reduction of each fp32 PCG solve */
  int32_t max_outer;           /* mixed IR: maximum refinement steps */
  int32_t max_iter;            /* maximum PCG iterations per (inner) solve */
} sdc_solve_opts;

typedef struct {
  int32_t outer;               /* fp64 true-residual recomputations (refinement steps) */
  int32_t iters;               /* PCG iterations summed over refinement steps */
  double max_rel_res;          /* max over columns of |q - A u|/|q| after the solve */
  int32_t converged;           /* 1 if every column reached tol */
  int32_t fp64_iters;          /* mixed mode: iterations of the fp64 PCG continuation that takes
                                  over when the fp32 solve stops making progress (0: not used) */
} sdc_solve_stats;

/* Library identification and the last error message of the calling thread. */
const char* sdc_version(void);
const char* sdc_last_error(void);

/* Create / destroy a handle holding the mesh, the multigrid hierarchy and the solver workspace
 * (device memory owned by the handle).  Replaces Simulation3DNodal.__init__ + setBC
 * (resistivity/simulation.py:483-501, 573-644). */
int sdc_create(const sdc_mesh_desc* desc, sdc_handle* out);
int sdc_destroy(sdc_handle h);

/* Column stride (elements) of level-0 node vectors, and node/cell counts. */
int64_t sdc_ld(sdc_handle h);
int64_t sdc_num_nodes(sdc_handle h);
int64_t sdc_num_cells(sdc_handle h);
int32_t sdc_num_levels(sdc_handle h);

int sdc_default_params(sdc_params* p);
int sdc_set_params(sdc_handle h, const sdc_params* p);

/* Set the cell conductivity (device fp64, nC, Fortran cell order) and rebuild the edge
 * conductances of every level and the coarsest dense inverse.  Replaces MeSigma + getA
 * (base/pde_simulation.py:150-162, resistivity/simulation.py:523-555) and the factorisation
 * `self.Ainv = self.solver(A, **self.solver_opts)` (resistivity/simulation.py:108). */
int sdc_set_sigma(sdc_handle h, const double* sigma_dev, void* stream);

/* Several operators on one mesh ("operator sets"): set s is built from the cell conductivity
 * sigma_dev + s * nC (device fp64, nset x nC), with every multigrid level and coarse inverse, in
 * one pass over all sets.  Replaces, for MetaSimulation's sub-simulations that share a mesh
 * (meta/simulation.py:13-330), their separate getA + factorisation calls
 * (resistivity/simulation.py:108, 523-555).  Column solves then pick their set with
 * sdc_solve_sets; sdc_solve, sdc_apply and sdc_vcycle use set 0.  Growing the number of sets
 * reallocates the per-set arrays. */
int sdc_set_sigma_sets(sdc_handle h, int nset, const double* sigma_dev, void* stream);

/* Set the finest level's operator directly (device fp64, nN each): cx/cy/cz = conductance of the
 * +x/+y/+z edge of every node (0 where there is no such edge), bd = the rest of the diagonal
 * (Robin term).  For the solver plugin point, which receives an assembled operator instead of
 * sigma: `self.solver(A, **self.solver_opts)` (resistivity/simulation.py:108; solver contract
 * base/pde_simulation.py:493-511).  Coarse levels and the coarse inverse are rebuilt as in
 * sdc_set_sigma; Neumann handles keep node 0's identity row. */
int sdc_set_coefficients(sdc_handle h, const double* cx_dev, const double* cy_dev,
                         const double* cz_dev, const double* bd_dev, void* stream);

/* y = A x for ncol columns (dtype SDC_F32 or SDC_F64), with A exactly the reference operator
 * getA() (Neumann: identity row 0).  resistivity/simulation.py:523-555. */
int sdc_apply(sdc_handle h, int dtype, const void* x_dev, void* y_dev, int ncol, void* stream);

/* z = M^-1 b: one multigrid V-cycle of the preconditioner (testing / external Krylov use). */
int sdc_vcycle(sdc_handle h, int dtype, const void* b_dev, void* z_dev, int ncol, void* stream);

/* Solve A u = q for ncol columns (fp64 q and u, stride sdc_ld).  Replaces `Ainv * RHS`
 * (resistivity/simulation.py:111, 181, 245).  stats may be NULL. */
int sdc_solve(sdc_handle h, const double* q_dev, double* u_dev, int ncol,
              const sdc_solve_opts* opts, sdc_solve_stats* stats, void* stream);

/* sdc_solve with column c solved against operator set colset[c] (host array of ncol entries,
 * sets built by sdc_set_sigma_sets; each aligned group of 8 columns must use one set, so pad a
 * set's columns to a multiple of 8 with zero right-hand sides).  One batched solve for the
 * sources of many same-mesh problems: `Ainv * RHS` of every sub-simulation of a MetaSimulation
 * (resistivity/simulation.py:111, 181, 245; meta/simulation.py:194-330).  stats cover all
 * columns. */
int sdc_solve_sets(sdc_handle h, const double* q_dev, double* u_dev, int ncol,
                   const int32_t* colset, const sdc_solve_opts* opts, sdc_solve_stats* stats,
                   void* stream);

/* out = -A(dsig) u  (no Neumann row replacement): the Jvec right-hand side -getADeriv(u, v)
 * with dsig = sigmaDeriv @ v (resistivity/simulation.py:557-565, 177-181). */
int sdc_jvec_rhs(sdc_handle h, const double* dsig_dev, const double* u_dev, double* out_dev,
                 int ncol, void* stream);

/* Adjoint contraction of getADeriv(u, lam, adjoint=True) before the map derivative
 * (resistivity/simulation.py:566-570, base/pde_simulation.py:35-36):
 *   g_c = d(u^T A(sigma) lam)/d sigma_c = A_v^T (G u . G lam) + AvgBC^T (u . lam)
 * Column j of lam is paired with column ucol[j] of u (ucol_dev NULL: j).  per_column = 0 sums
 * over columns into out_dev[nC]; per_column = 1 writes out_dev[j*ldo + c] (getJ rows). */
int sdc_jtvec_contract(sdc_handle h, const double* u_dev, const double* lam_dev, int ncol,
                       const int32_t* ucol_dev, int per_column, double* out_dev, int64_t ldo,
                       void* stream);

/* The column-summed contraction of sdc_jtvec_contract (per_column = 0) without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its per-column-group partial sums live in a fixed 32 MiB device array of the library, one per
 * GPU, held until the process ends; groups beyond what it holds are summed in further passes in
 * the same order, so the result equals sdc_jtvec_contract's bit for bit.  Limit: meshes of at
 * most 4,194,304 cells (larger ones return SDC_ERR_ARG).  All calls on one GPU share that array:
 * calls that overlap on one GPU (other streams, other handles, other threads) race. */
int sdc_jtvec_contract_noalloc(sdc_handle h, const double* u_dev, const double* lam_dev, int ncol,
                               const int32_t* ucol_dev, double* out_dev, void* stream);

/* out_dev[c] += sum_j wsq[j] * g_{j,c}^2 with g_j the contraction above for column j (paired
 * with u column ucol[j]): the diagonal of J^T W^T W J before the map derivative, accumulated
 * over chunks of data columns without materialising J (getJtJdiag,
 * resistivity/simulation.py:137-154).  wsq_dev: ncol squared data weights. */
int sdc_jtj_diag_accumulate(sdc_handle h, const double* u_dev, const double* lam_dev, int ncol,
                            const int32_t* ucol_dev, const double* wsq_dev, double* out_dev,
                            void* stream);

/* d[i] = sum_{e in [rowptr[i], rowptr[i+1])} w[e] * u[idx[e]]: the receiver projection
 * P u of BaseRx.eval (receivers.py:156-192), idx = flat column*ld + node. */
int sdc_gather(sdc_handle h, int nrow, const int32_t* rowptr_dev, const int64_t* idx_dev,
               const double* w_dev, const double* u_dev, double* d_dev, void* stream);

/* out[dst[i]] = sum_{e in row i} w[e] * v[src[e]] after zeroing ncol columns of out: the adjoint
 * projection P^T v (receivers.py:241-244) and the source term q (sources.py:64-88). */
int sdc_scatter(sdc_handle h, int nrow, const int64_t* dst_dev, const int32_t* rowptr_dev,
                const int32_t* src_dev, const double* w_dev, const double* v_dev,
                double* out_dev, int ncol, void* stream);

/* Event profiler (no reference counterpart; instrumentation for the roofline report).
 * sdc_profile(h, mask) selects kernel classes (bit 1 << tag per class); each later launch of a
 * selected class is bracketed by HIP events on its stream.  sdc_profile_read(h, tag, ...)
 * returns for one class the number of launches, their summed duration (ms) and summed
 * algorithmic bytes (the minimum HBM traffic of each launch: every array element it must
 * read or write counted once), then clears that record.  mask 0 = off.  Finest level only. */
#define SDC_PROF_SMOOTH 1     /* last post-smoothing step (writes z and the r.z partials) */
#define SDC_PROF_APPLY 2      /* operator application inside PCG (q = A p, p.q partials) */
#define SDC_PROF_RESIDUAL 3   /* V-cycle residual */
#define SDC_PROF_REFINE 4     /* fp64 residual q - A u of a reliable update or refinement step */
#define SDC_PROF_JTVEC 5      /* adjoint edge-gradient contraction */
#define SDC_PROF_RESTRICT 6   /* one per-axis restriction pass (or the 27-point gather) */
#define SDC_PROF_PROLONG 7    /* interpolation from level 1, added into the smoother iterate */
#define SDC_PROF_UPDATE 8     /* PCG x += alpha p, r -= alpha q and |r|^2 partials */
#define SDC_PROF_PAIR 9       /* fused first two pre-smoothing steps */
#define SDC_PROF_SMOOTH0 10   /* first post-smoothing step */
#define SDC_PROF_UPDATE_P 11  /* PCG p = z + beta p */
#define SDC_PROF_FOLD 12      /* reliable update u += x, x = 0 */
#define SDC_PROF_NTAGS 13
int sdc_profile(sdc_handle h, uint32_t mask);
int sdc_profile_read(sdc_handle h, int tag, int64_t* count, double* total_ms, double* total_bytes);

#ifdef __cplusplus
}
#endif

#endif /* SIMPEG_DC_H */
