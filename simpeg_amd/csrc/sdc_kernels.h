// HIP kernels for the matrix-free 3D nodal DC operator A(sigma) = G^T M_e(sigma) G + diag(AvgBC sigma),
// its geometric-multigrid preconditioner, block-PCG vector updates, fp64 iterative
// refinement, and the adjoint (Jtvec) edge-gradient contraction.
//
// Reference arithmetic restated here (paths relative to the simpeg repository):
//   A(sigma):            electromagnetics/static/resistivity/simulation.py:523-555
//   Robin term AvgBC:    electromagnetics/static/resistivity/simulation.py:579-644
//   dA/dm (adjoint):     electromagnetics/static/resistivity/simulation.py:557-571,
//                        base/pde_simulation.py:15-83, 296-323
#pragma once
#include "sdc_device.h"

// ------------------------------------------------------------------------------------------
// generic launch geometry: blockIdx.x -> 256 nodes (XCD-remapped), blockIdx.y -> R columns

// The linear work-group id (x fastest, as the dispatcher numbers them) is remapped so that
// every XCD works through one contiguous range of (column group, node block) pairs.
// Work-group -> (node block bx, column group by) of a (node blocks, column groups) grid.
// remap 0: blockIdx order; 1: each XCD takes one contiguous run of the linear order (node
// blocks fastest), so neighbouring planes of a sweep share that XCD's L2; B > 1: in addition
// the column groups come in bands of B whose work-groups for one node block are consecutive,
// so that block's coefficients are fetched once per band and then read from L2.
__device__ __forceinline__ void block_of(int remap, uint32_t& bx, uint32_t& by) {
  const uint32_t nbx = gridDim.x, ng = gridDim.y;
  const uint32_t lin0 = blockIdx.x + nbx * blockIdx.y;
  const uint32_t lin = remap ? xcd_remap(lin0, nbx * ng) : lin0;
  if (remap <= 1) {
    by = lin / nbx;
    bx = lin - by * nbx;
    return;
  }
  const uint32_t B = (uint32_t)remap;
  const uint32_t band = lin / (nbx * B);
  const uint32_t w = lin - band * nbx * B;
  const uint32_t cb = min(B, ng - band * B);
  bx = w / cb;
  by = band * B + (w - bx * cb);
}

#define SDC_NODE_PROLOGUE_C(NN, REMAP, CPB)                                     \
  const uint32_t nbx = gridDim.x;                                               \
  uint32_t bx, by_;                                                             \
  block_of((REMAP), bx, by_);                                                   \
  const uint32_t n = bx * SDC_BLOCK + threadIdx.x;                              \
  const int c0 = (int)by_ * (CPB);                                              \
  const bool live = n < (NN);

#define SDC_NODE_PROLOGUE_R(NN, REMAP) SDC_NODE_PROLOGUE_C(NN, REMAP, R)

// Column loops: the stencil kernels take G groups of R columns per work-item, one group after
// the other, with the node's coefficients loaded once (G = 1: one group).  With operator sets
// the columns of one work-item may cross a set boundary (sets are aligned to 8 columns): the
// operator is reselected where the set changes.
template <typename T>
__device__ __forceinline__ bool set_changes(const Sten<T>& sa, int c, int cprev) {
  return sa.cset && sa.cset[c] != sa.cset[cprev];
}

#define SDC_NODE_PROLOGUE(NN) SDC_NODE_PROLOGUE_R(NN, 1)

// Control block of the mixed PCG (device ints), read and written by the kernels so that the
// iteration runs without a host decision: STOP ends the solve (every kernel of an iteration
// returns at once), UPD asks for a reliable update in the current iteration, PEND counts the
// search directions in the ring not yet folded into u, FOLD asks for a fold without a
// residual replacement (the ring is full), UFRESH marks u as not yet written, IT counts
// iterations, LAST is the iteration of the last update, REASON says why the solve stopped
// (1 converged, 2 iteration limit, 3 stalled) and OUTER counts reliable updates; MAXIT and
// WINDOW are the iteration limit and the stall window of the solve.  The solve's parameters
// live here, not in kernel arguments, so that a recorded iteration serves solves with any
// tolerance or limit.
#define SDC_CTL_STOP 0
#define SDC_CTL_UPD 1
#define SDC_CTL_PEND 2
#define SDC_CTL_UFRESH 3
#define SDC_CTL_IT 4
#define SDC_CTL_LAST 5
#define SDC_CTL_REASON 6
#define SDC_CTL_OUTER 7
#define SDC_CTL_MAXIT 8
#define SDC_CTL_WINDOW 9
#define SDC_CTL_FOLD 10
#define SDC_CTL_N 16
// doubles after the ints: the largest true relative residual of the previous update and of the
// last one, tol^2 and reliable_delta^2
#define SDC_DCTL_MPREV 0
#define SDC_DCTL_MREL 1
#define SDC_DCTL_TOL2 2
#define SDC_DCTL_DELTA2 3
#define SDC_DCTL_N 4

__device__ __forceinline__ double* ctl_dbl(int* ctl) { return (double*)(ctl + SDC_CTL_N); }
__device__ __forceinline__ const double* ctl_dbl(const int* ctl) {
  return (const double*)(ctl + SDC_CTL_N);
}

// a reliable update runs in this iteration
__device__ __forceinline__ bool ctl_update(const int* ctl) {
  return ld_flag(ctl + SDC_CTL_STOP) == 0 && ld_flag(ctl + SDC_CTL_UPD) != 0;
}

// the pending search directions are folded into u in this iteration (a reliable update, or
// a full ring)
__device__ __forceinline__ bool ctl_fold(const int* ctl) {
  return ld_flag(ctl + SDC_CTL_STOP) == 0 &&
         (ld_flag(ctl + SDC_CTL_UPD) != 0 || ld_flag(ctl + SDC_CTL_FOLD) != 0);
}

// ------------------------------------------------------------------------------------------
// operator application: y = A x, optional per-column partial sums of x.y

template <typename T, int R, bool DOT, int G = 1>
__global__ __launch_bounds__(SDC_BLOCK) void k_apply(Sten<T> sa, const T* __restrict__ x,
                                                     T* __restrict__ y, size_t ld, int ncol,
                                                     double* __restrict__ partial,
                                                     const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE_C(sa.nN, sa.remap, R * G)
  double acc[R * G];
#pragma unroll
  for (int r = 0; r < R * G; ++r) acc[r] = 0.0;
  if (live) {
    Sten<T> s = sten_col(sa, c0);
    NodeC<T> c = node_coef(s, n);
#pragma unroll 2
    for (int g = 0; g < G; ++g) {
      const int cg = c0 + g * R;
      if (cg >= ncol) break;
      if (G > 1 && g > 0 && set_changes(sa, cg, cg - R)) {
        s = sten_col(sa, cg);
        c = node_coef(s, n);
      }
      T xc[R], v[R];
      stencil_cols<T, R>(c, x, n, ld, cg, ncol, xc, v);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (cg + r < ncol) {
          st_once(y + (size_t)(cg + r) * ld + n, v[r]);
          if (DOT) acc[g * R + r] += (double)xc[r] * (double)v[r];
        }
      }
    }
  }
  if (DOT) block_partials<R * G>(acc, partial, c0, ncol, nbx, bx);
}

// r = b - A x.  b in type B, x and r in the V-cycle storage type X, arithmetic in T.
template <typename T, typename B, typename X, int R, int G = 1>
__global__ __launch_bounds__(SDC_BLOCK) void k_residual(Sten<T> sa, const B* __restrict__ b,
                                                        const X* __restrict__ x,
                                                        X* __restrict__ res, size_t ld, int ncol,
                                                        const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE_C(sa.nN, sa.remap, R * G)
  if (!live) return;
  Sten<T> s = sten_col(sa, c0);
  NodeC<T> c = node_coef(s, n);
#pragma unroll 2
  for (int g = 0; g < G; ++g) {
    const int cg = c0 + g * R;
    if (cg >= ncol) break;
    if (G > 1 && g > 0 && set_changes(sa, cg, cg - R)) {
      s = sten_col(sa, cg);
      c = node_coef(s, n);
    }
    T xc[R], v[R], bv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) bv[r] = (T)ld_once(col_ptr(b, ld, cg, r, ncol) + n);
    stencil_cols<T, R>(c, x, n, ld, cg, ncol, xc, v);
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (cg + r < ncol) st_once(res + (size_t)(cg + r) * ld + n, (X)(bv[r] - v[r]));
  }
}

// First Chebyshev step from a zero guess: x = c0 * D^-1 b
template <typename T, typename B, typename X, int R>
__global__ __launch_bounds__(SDC_BLOCK) void k_smooth0(Sten<T> sa, const B* __restrict__ b,
                                                       X* __restrict__ x, size_t ld, int ncol,
                                                       T coef, const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE_R(sa.nN, sa.remap)
  const Sten<T> s = sten_col(sa, c0);
  if (!live) return;
  const T w = coef * s.idg[n];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (c0 + r < ncol) {
      const size_t o = (size_t)(c0 + r) * ld;
      st_once(x + o + n, (X)(w * (T)ld_once(b + o + n)));
    }
  }
}

// Chebyshev step in three-term form:  xout = x + c1 (x - xprev) + c2 D^-1 (b - A x).
//   PREV = 0: first step from a non-zero guess (no c1 term);
//   PREV = 1: xprev read pointwise; xout may be the same buffer (each work-item reads its
//             xprev entries before it writes xout, and nobody else reads them);
//   PREV = 2: xprev = cz D^-1 b, the first iterate from a zero guess, recomputed here.
// DOT adds per-column partial sums of b.xout (the r.z product of PCG).
// b in type B, x and xprev in the storage type X, xout in type O (O = T on the last step of
// the finest level, which hands z back to PCG).
template <typename T, typename B, typename X, typename O, int R, int PREV, bool DOT, int G = 1>
__global__ __launch_bounds__(SDC_BLOCK) void k_cheb(Sten<T> sa, const B* __restrict__ b,
                                                    const X* __restrict__ x, const X* xprev,
                                                    O* xout, size_t ld, int ncol, T c1, T c2,
                                                    T cz, double* __restrict__ partial,
                                                    const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE_C(sa.nN, sa.remap, R * G)
  double acc[R * G];
#pragma unroll
  for (int r = 0; r < R * G; ++r) acc[r] = 0.0;
  if (live) {
    Sten<T> s = sten_col(sa, c0);
    NodeC<T> c = node_coef(s, n);
    T g = s.idg[n];
#pragma unroll 2
    for (int q = 0; q < G; ++q) {
      const int cg = c0 + q * R;
      if (cg >= ncol) break;
      if (G > 1 && q > 0 && set_changes(sa, cg, cg - R)) {
        s = sten_col(sa, cg);
        c = node_coef(s, n);
        g = s.idg[n];
      }
      const T w = c2 * g;
      T xc[R], v[R], bv[R], pv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        bv[r] = (T)ld_once(col_ptr(b, ld, cg, r, ncol) + n);
        pv[r] = (PREV == 1) ? (T)col_ptr(xprev, ld, cg, r, ncol)[n] : T(0);
      }
      stencil_cols<T, R>(c, x, n, ld, cg, ncol, xc, v);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (cg + r < ncol) {
          const size_t o = (size_t)(cg + r) * ld + n;
          T xn = xc[r] + w * (bv[r] - v[r]);
          if (PREV == 1) xn += c1 * (xc[r] - pv[r]);
          if (PREV == 2) xn += c1 * (xc[r] - cz * g * bv[r]);
          const O xs = (O)xn;
          st_once(xout + o, xs);
          // the product uses the value stored (z may be rounded to bf16), as the next step does
          if (DOT) acc[q * R + r] += (double)bv[r] * (double)(T)xs;
        }
      }
    }
  }
  if (DOT) block_partials<R * G>(acc, partial, c0, ncol, nbx, bx);
}

// The first two Chebyshev steps from a zero guess in one pass:
//   x1 = c0 D^-1 b (recomputed at the 7 stencil points, never stored), r1 = b - A x1,
//   x2 = x1 + c1 x1 + c2 D^-1 r1.  Writes x2 only (a following step recomputes x1 from b).
template <typename T, typename B, typename X, int R, int G = 1>
__global__ __launch_bounds__(SDC_BLOCK) void k_smooth_pair(Sten<T> sa, const B* __restrict__ b,
                                                           X* __restrict__ xout, size_t ld,
                                                           int ncol, T cz, T c1, T c2,
                                                           const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE_C(sa.nN, sa.remap, R * G)
  if (!live) return;
  // A x1 = c0 * A' b with the neighbour weights scaled by their 1/diagonal
  auto scaled = [&](const Sten<T>& s, T& g) {
    NodeC<T> c = node_coef(s, n);
    g = s.idg[n];
    c.d *= g;
    c.wxm *= s.idg[c.ixm];
    c.wxp *= s.idg[c.ixp];
    c.wym *= s.idg[c.iym];
    c.wyp *= s.idg[c.iyp];
    c.wzm *= s.idg[c.izm];
    c.wzp *= s.idg[c.izp];
    return c;
  };
  T g;
  NodeC<T> c = scaled(sten_col(sa, c0), g);
#pragma unroll 2
  for (int q = 0; q < G; ++q) {
    const int cg = c0 + q * R;
    if (cg >= ncol) break;
    if (G > 1 && q > 0 && set_changes(sa, cg, cg - R)) c = scaled(sten_col(sa, cg), g);
    T bc[R], ab[R];
    stencil_cols<T, R>(c, b, n, ld, cg, ncol, bc, ab);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (cg + r < ncol) {
        const size_t o = (size_t)(cg + r) * ld + n;
        const T x1 = cz * g * bc[r];
        const T r1 = bc[r] - cz * ab[r];
        st_once(xout + o, (X)(x1 + c1 * x1 + c2 * g * r1));
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// inter-grid transfer (vertex-based coarsening of each tensor axis, linear interpolation
// in physical coordinates, restriction = transpose of interpolation)

template <typename T>
struct Xfer {
  // restriction tables indexed by coarse index: fine index of the coarse node and the
  // weights of its fine neighbours (0 when that neighbour is itself a coarse node)
  const int* rix;
  const int* riy;
  const int* riz;
  const T* rlx;
  const T* rly;
  const T* rlz;
  const T* rhx;
  const T* rhy;
  const T* rhz;
  // interpolation tables indexed by fine index: left coarse node and weight of the right one
  const int* pix;
  const int* piy;
  const int* piz;
  const T* pwx;
  const T* pwy;
  const T* pwz;
  uint32_t fn1x, fn1y, fplane, fN;
  uint32_t cn1x, cn1y, cplane, cN;
  FastDiv ffx, ffy, cfx, cfy;
  int remap;
  // operator-dependent interpolation (null: the geometric tables above).  P = Px Py Pz, each
  // factor a 1D interpolation along its axis whose weight of the right coarse node is stored
  // per point of the factor's grid: wx over the fine grid (fn1x, fn1y, fn1z), wy over
  // (cn1x, fn1y, fn1z), wz over (cn1x, cn1y, fn1z)
  const T* wx;
  const T* wy;
  const T* wz;
  // operator sets: column c uses wx + cset[c] * wxs (and likewise wy, wz); cset null: one set
  const int32_t* cset;
  uint32_t wxs, wys, wzs;
};

// The transfer of column c (operator-dependent weights of its operator set).
template <typename T>
__device__ __forceinline__ Xfer<T> xfer_col(Xfer<T> t, int c) {
  if (t.cset && t.wx) {
    const size_t k = (size_t)t.cset[c];
    t.wx += k * t.wxs;
    t.wy += k * t.wys;
    t.wz += k * t.wzs;
  }
  return t;
}

// Weights of the 8 coarse nodes around fine node n = (i, j, k); corner a + 2b + 4c is
// (I0 + a, J0 + b, K0 + c).  P = Px Py Pz makes the corner weight the product
// wx_a(i, j, k) * wy_b(I0 + a, j, k) * wz_c(I0 + a, J0 + b, k), each factor read on its own grid;
// with the geometric tables the three factors depend on i, j and k only (trilinear).
template <typename T>
__device__ __forceinline__ void prolong_weights(const Xfer<T>& t, uint32_t n, uint32_t i,
                                                uint32_t j, uint32_t k, uint32_t I0, uint32_t J0,
                                                T w[8]) {
  T fx, fy[2], fz[2][2];
  if (t.wx) {
    fx = t.wx[n];
    const uint32_t py = I0 + t.cn1x * (j + t.fn1y * k);
    fy[0] = t.wy[py];
    fy[1] = t.wy[py + 1];
    const uint32_t pz = I0 + t.cn1x * J0 + t.cplane * k;
    fz[0][0] = t.wz[pz];
    fz[1][0] = t.wz[pz + 1];
    fz[0][1] = t.wz[pz + t.cn1x];
    fz[1][1] = t.wz[pz + t.cn1x + 1];
  } else {
    fx = t.pwx[i];
    fy[0] = fy[1] = t.pwy[j];
    fz[0][0] = fz[1][0] = fz[0][1] = fz[1][1] = t.pwz[k];
  }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a)
        w[a + 2 * b + 4 * c] = (a ? fx : T(1) - fx) * (b ? fy[a] : T(1) - fy[a]) *
                               (c ? fz[a][b] : T(1) - fz[a][b]);
}

// Weights of P^T for coarse node (I, J, K) at fine (fi, fj, fk) over the 3x3x3 fine nodes
// (xi[a], yj[b], zk[c]) (index a + 3b + 9c), the transpose of prolong_weights: a fine neighbour
// halfway to the next coarse node on the left contributes its weight of the right coarse node,
// one on the right 1 minus that; 0 where the neighbour is missing or is itself a coarse node
// (the geometric tables' zeros), the index then being clamped into the grid.
template <typename T>
__device__ __forceinline__ void restrict_weights(const Xfer<T>& t, uint32_t I, uint32_t J,
                                                 uint32_t K, const int xi[3], const int yj[3],
                                                 const int zk[3], T w[27]) {
  const T gx[3] = {t.rlx[I], T(1), t.rhx[I]};
  const T gy[3] = {t.rly[J], T(1), t.rhy[J]};
  const T gz[3] = {t.rlz[K], T(1), t.rhz[K]};
  if (!t.wx) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) w[a + 3 * b + 9 * c] = gx[a] * gy[b] * gz[c];
    return;
  }
  const uint32_t pz = I + t.cn1x * J;
  const T z[3] = {gz[0] != T(0) ? t.wz[pz + t.cplane * (uint32_t)zk[0]] : T(0), T(1),
                  gz[2] != T(0) ? T(1) - t.wz[pz + t.cplane * (uint32_t)zk[2]] : T(0)};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const uint32_t py = I + t.cn1x * t.fn1y * (uint32_t)zk[c];
    const T y[3] = {gy[0] != T(0) ? t.wy[py + t.cn1x * (uint32_t)yj[0]] : T(0), T(1),
                    gy[2] != T(0) ? T(1) - t.wy[py + t.cn1x * (uint32_t)yj[2]] : T(0)};
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const uint32_t row = (uint32_t)yj[b] * t.fn1x + (uint32_t)zk[c] * t.fplane;
      const T yz = y[b] * z[c];
      w[3 * b + 9 * c] = gx[0] != T(0) ? t.wx[row + (uint32_t)xi[0]] * yz : T(0);
      w[1 + 3 * b + 9 * c] = yz;
      w[2 + 3 * b + 9 * c] = gx[2] != T(0) ? (T(1) - t.wx[row + (uint32_t)xi[2]]) * yz : T(0);
    }
  }
}

// Operator-dependent interpolation weights of the factor of P = Px Py Pz along axis AX, over
// the factor's grid (g0, g1, g2) (see Xfer::wx).  A fine node halfway between two coarse nodes
// along AX gets the weight cR / (cL + cR) for the right one, cL and cR being the conductances of
// its two edges along AX on the fine line through it: the exact interpolation of the 1D
// operator, equal to the geometric weight in a homogeneous medium.  Coincident nodes keep their
// 0 / 1 and nodes of longer gaps the geometric weight.  t: the geometric tables (fp64);
// n: the point of the factor's grid; c: the fine level's conductances along AX (of one operator
// set, as w64 and w32).  Run by k_coarsen_coef.
template <int AX>
__device__ __forceinline__ void interp_weight(const Xfer<double>& t, size_t n, uint32_t g0,
                                              uint32_t g1, const double* __restrict__ c,
                                              double* __restrict__ w64, float* __restrict__ w32) {
  const uint32_t p0 = (uint32_t)(n % g0);
  const size_t q = n / g0;
  const uint32_t p1 = (uint32_t)(q % g1);
  const uint32_t p2 = (uint32_t)(q / g1);
  // fine indices of the grid point: coarse x on the Py and Pz grids, coarse y on the Pz grid
  const uint32_t fi = AX == 0 ? p0 : (uint32_t)t.rix[p0];
  const uint32_t fj = AX == 2 ? (uint32_t)t.riy[p1] : p1;
  const uint32_t fk = p2;
  const uint32_t f = AX == 0 ? fi : AX == 1 ? fj : fk;
  const int* pi = AX == 0 ? t.pix : AX == 1 ? t.piy : t.piz;
  const int* ri = AX == 0 ? t.rix : AX == 1 ? t.riy : t.riz;
  const double* pw = AX == 0 ? t.pwx : AX == 1 ? t.pwy : t.pwz;
  const size_t stride = AX == 0 ? 1 : AX == 1 ? (size_t)t.fn1x : (size_t)t.fplane;
  const int I0 = pi[f];
  double w = pw[f];
  if (ri[I0] != (int)f && ri[I0 + 1] != (int)f && ri[I0 + 1] - ri[I0] == 2) {
    const size_t node = fi + (size_t)fj * t.fn1x + (size_t)fk * t.fplane;
    const double cl = c[node - stride], cr = c[node];
    if (cl + cr > 0.0) w = cr / (cl + cr);
  }
  w64[n] = w;
  w32[n] = (float)w;
}

// rc = P^T rf: every coarse node gathers the 3x3x3 fine nodes around its fine position
// (restrict_weights); a missing neighbour (domain edge or coincident coarse node) has weight 0
// and its index is clamped into the grid.
template <typename T, typename X, int R>
__global__ __launch_bounds__(SDC_BLOCK) void k_restrict(Xfer<T> ta, const X* __restrict__ rf,
                                                        X* __restrict__ rc, size_t ldf, size_t ldc,
                                                        int ncol, const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE_R(ta.cN, ta.remap)
  const Xfer<T> t = xfer_col(ta, c0);
  if (!live) return;
  const uint32_t rr = fdiv(n, t.cfx);
  const uint32_t I = n - rr * t.cn1x;
  const uint32_t K = fdiv(rr, t.cfy);
  const uint32_t J = rr - K * t.cn1y;
  const int fi = t.rix[I], fj = t.riy[J], fk = t.riz[K];
  const int fn1z = (int)(t.fN / t.fplane);
  const int xi[3] = {fi > 0 ? fi - 1 : 0, fi, fi + 1 < (int)t.fn1x ? fi + 1 : fi};
  const int yj[3] = {fj > 0 ? fj - 1 : 0, fj, fj + 1 < (int)t.fn1y ? fj + 1 : fj};
  const int zk[3] = {fk > 0 ? fk - 1 : 0, fk, fk + 1 < fn1z ? fk + 1 : fk};
  T w[27];
  restrict_weights(t, I, J, K, xi, yj, zk, w);
  T acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = T(0);
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const T* wb = w + 3 * b + 9 * c;
      const uint32_t row = (uint32_t)yj[b] * t.fn1x + (uint32_t)zk[c] * t.fplane;
      const uint32_t i0 = row + xi[0], i1 = row + xi[1], i2 = row + xi[2];
      T v[R][3];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const X* f = col_ptr(rf, ldf, c0, r, ncol);
        v[r][0] = (T)f[i0];
        v[r][1] = (T)f[i1];
        v[r][2] = (T)f[i2];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] += wb[0] * v[r][0] + wb[1] * v[r][1] + wb[2] * v[r][2];
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (c0 + r < ncol) rc[(size_t)(c0 + r) * ldc + n] = (X)acc[r];
}

// One factor of the separable restriction P^T = Pz^T Py^T Px^T: along axis AX the output has
// the coarse count, every output entry gathers the three input entries around its fine index
// (weights lo, 1, hi, or from wop for the operator-dependent P; a weight is 0 where that
// neighbour is missing or is itself a coarse node, and the index is clamped).  Input dims (a0, a1, a2), output dims (b0, b1, b2), column
// strides ldi / ldo.  Work-items run over output entries with b0 fastest, so every load and
// store of a wavefront covers consecutive addresses (stride 2 at most along x).
struct AxisXfer {
  uint32_t a0, a1, a2, b0, b1, b2, nout;
  FastDiv f0, f1;  // division by b0 and b1
  int remap;
  const int32_t* cset;  // operator sets: wop + cset[c] * wstride (null: one set)
  uint32_t wstride;
};

// r = b - A x and its restriction along x in one pass: out = Px^T r on the grid (coarse x, fine
// y, fine z), the first pass of the separable P^T (k_restrict_axis<..., 0, ...>), without
// storing r.  A work-group computes r at SDC_BLOCK consecutive fine nodes starting one node
// before its tile of SDC_BLOCK - 2 nodes and keeps it in LDS, rounded to X as a stored residual
// would be; every node of the tile that is a coarse x position then writes
// wl r(n-1) + r(n) + wh r(n+1) with the x-restriction weights of k_restrict_axis (OPD: from
// wop over the fine grid, masked by the geometric ones).  A node's x neighbours n-1, n+1 lie in
// the same grid row wherever the weight is non-zero.  Results equal the two-pass form's (up to
// the sign of zero terms).  ax: the x pass's AxisXfer; fi / pix: the x tables (coarse -> fine
// index, fine -> left coarse node).
template <typename T, typename B, typename X, int R, bool OPD>
__global__ __launch_bounds__(SDC_BLOCK) void k_residual_rx(Sten<T> sa, const B* __restrict__ b,
                                                           const X* __restrict__ x,
                                                           X* __restrict__ out, size_t ld,
                                                           size_t ldo, int ncol, AxisXfer ax,
                                                           const int* __restrict__ fi,
                                                           const int* __restrict__ pix,
                                                           const T* __restrict__ wlo,
                                                           const T* __restrict__ whi,
                                                           const T* __restrict__ wop,
                                                           const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  constexpr uint32_t TS = SDC_BLOCK - 2;
  __shared__ T rl[R][SDC_BLOCK];
  uint32_t bx, by;
  block_of(sa.remap, bx, by);
  const int c0 = (int)by * R;
  const int64_t nn = (int64_t)bx * TS - 1 + threadIdx.x;
  const bool live = nn >= 0 && nn < (int64_t)sa.nN;
  const uint32_t n = live ? (uint32_t)nn : 0u;
  T v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = T(0);
  if (live) {
    const NodeC<T> c = node_coef(sten_col(sa, c0), n);
    T xc[R], av[R], bv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) bv[r] = (T)ld_once(col_ptr(b, ld, c0, r, ncol) + n);
    stencil_cols<T, R>(c, x, n, ld, c0, ncol, xc, av);
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = (T)(X)(bv[r] - av[r]);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) rl[r][threadIdx.x] = v[r];
  __syncthreads();
  if (!live || threadIdx.x == 0 || threadIdx.x == SDC_BLOCK - 1) return;
  // fine x index i of n and the rest of its index (j + n1y k); n restricts to coarse x node I
  // when it is one's fine position
  const uint32_t rest = fdiv(n, sa.fx);  // division by the fine row length (= ax.a0)
  const uint32_t i = n - rest * sa.n1x;
  const int I0 = pix[i];
  const int I = fi[I0] == (int)i ? I0 : (I0 + 1 < (int)ax.b0 && fi[I0 + 1] == (int)i) ? I0 + 1 : -1;
  if (I < 0) return;
  T wl = wlo[I], wh = whi[I];
  if (OPD) {
    const T* w = ax.cset ? wop + (size_t)ax.cset[c0] * ax.wstride : wop;
    const uint32_t il = i > 0 ? n - 1 : n, ih = i + 1 < ax.a0 ? n + 1 : n;
    wl = w[il] * (T)(wl != T(0));
    wh = (T(1) - w[ih]) * (T)(wh != T(0));
  }
  const size_t o = (size_t)I + (size_t)ax.b0 * rest;
  const int t = threadIdx.x;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (c0 + r < ncol)
      st_once(out + (size_t)(c0 + r) * ldo + o,
              (X)(wl * rl[r][t - 1] + rl[r][t] + wh * rl[r][t + 1]));
}

// OPD: operator-dependent weights wop (else the geometric wlo / whi alone).  The weight loads
// are unconditional (clamped indices) and the missing neighbours are masked by multiplying, so
// they go out with the value loads in the second round trip after the fi lookup.
template <typename T, typename X, int AX, int R, bool OPD>
__global__ __launch_bounds__(SDC_BLOCK) void k_restrict_axis(AxisXfer t, const int* __restrict__ fi,
                                                             const T* __restrict__ wlo,
                                                             const T* __restrict__ whi,
                                                             const T* __restrict__ wop,
                                                             const X* __restrict__ in,
                                                             X* __restrict__ out, size_t ldi,
                                                             size_t ldo, int ncol,
                                                             const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE_R(t.nout, t.remap)
  if (!live) return;
  const uint32_t q = fdiv(n, t.f0);
  const uint32_t o0 = n - q * t.b0;
  const uint32_t o2 = fdiv(q, t.f1);
  const uint32_t o1 = q - o2 * t.b1;
  const uint32_t C = AX == 0 ? o0 : AX == 1 ? o1 : o2;
  const uint32_t na = AX == 0 ? t.a0 : AX == 1 ? t.a1 : t.a2;
  const uint32_t stride = AX == 0 ? 1u : AX == 1 ? t.a0 : t.a0 * t.a1;
  const uint32_t f = (uint32_t)fi[C];
  const uint32_t fl = f > 0 ? f - 1 : 0, fh = f + 1 < na ? f + 1 : f;
  const uint32_t c0i = AX == 0 ? 0u : o0;
  const uint32_t c1i = AX == 1 ? 0u : o1;
  const uint32_t c2i = AX == 2 ? 0u : o2;
  const uint32_t base = c0i + t.a0 * (c1i + t.a1 * c2i);
  const uint32_t il = base + fl * stride, im = base + f * stride, ih = base + fh * stride;
  // operator-dependent weights (wop, over this pass's input grid): the left neighbour gives
  // its weight of the right coarse node, the right neighbour 1 minus its own
  T wl = wlo[C], wh = whi[C];
  if (OPD && t.cset) wop += (size_t)t.cset[c0] * t.wstride;
  if (OPD) {
    wl = wop[il] * (T)(wl != T(0));
    wh = (T(1) - wop[ih]) * (T)(wh != T(0));
  }
  T v[R][3];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const X* col = col_ptr(in, ldi, c0, r, ncol);
    v[r][0] = (T)col[il];
    v[r][1] = (T)col[im];
    v[r][2] = (T)col[ih];
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (c0 + r < ncol)
      st_once(out + (size_t)(c0 + r) * ldo + n, (X)(wl * v[r][0] + v[r][1] + wh * v[r][2]));
}

// xf += P xc, one work-item per coarse node (I, J, K): it updates the fine nodes it owns, the
// fine nodes from its own fine position up to (not including) the next coarse node's along
// every axis (1 or 2 per axis, up to 8), all of which interpolate from the 2x2x2 coarse nodes
// (I..I+1, J..J+1, K..K+1): one coarse load per owned fine node and column, and the weight
// loads are shared by the owned nodes and the R columns.  A fine node's value is
//   xf + sum over the parent-cell corners m = a + 2b + 4c (in that order) of w_m xc_m,
//   w_m = (a ? fx : 1 - fx) (b ? fy_a : 1 - fy_a) (c ? fz_ab : 1 - fz_ab)
// with fx the weight of the right x parent at the fine node, fy_a the weight of the upper y
// parent on the Py grid at x parent a, fz_ab that of the upper z parent on the Pz grid
// (operator-dependent: t.wx, t.wy, t.wz; else the geometric tables pwx, pwy, pwz).  An owned
// node coincident with the coarse grid along an axis has weight 0 on the next coarse node
// there, whose index is clamped into the grid.
template <typename T, typename X, int R, bool OPD>
__global__ __launch_bounds__(SDC_BLOCK) void k_prolong_own(Xfer<T> ta, const X* __restrict__ xc,
                                                           X* __restrict__ xf, size_t ldf,
                                                           size_t ldc, int ncol,
                                                           const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE_R(ta.cN, ta.remap)
  const Xfer<T> t = xfer_col(ta, c0);
  if (!live) return;
  const uint32_t rr = fdiv(n, t.cfx);
  const uint32_t I = n - rr * t.cn1x;
  const uint32_t K = fdiv(rr, t.cfy);
  const uint32_t J = rr - K * t.cn1y;
  const bool hx = I + 1 < t.cn1x, hy = J + 1 < t.cn1y, hz = (K + 1) * t.cplane < t.cN;
  const uint32_t Ip = hx ? I + 1 : I, Jp = hy ? J + 1 : J, Kp = hz ? K + 1 : K;
  const uint32_t i0 = (uint32_t)t.rix[I], j0 = (uint32_t)t.riy[J], k0 = (uint32_t)t.riz[K];
  // owned nodes per axis: 2 where the coarse cell spans 2 fine cells
  const bool ox = hx && (uint32_t)t.rix[Ip] == i0 + 2, oy = hy && (uint32_t)t.riy[Jp] == j0 + 2,
             oz = hz && (uint32_t)t.riz[Kp] == k0 + 2;
  const uint32_t c000 = I + J * t.cn1x + K * t.cplane;
  const uint32_t dx = Ip - I, dy = (Jp - J) * t.cn1x, dz = (Kp - K) * t.cplane;
  T cv[R][8];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const X* c = col_ptr(xc, ldc, c0, r, ncol);
    cv[r][0] = (T)c[c000];
    cv[r][1] = (T)c[c000 + dx];
    cv[r][2] = (T)c[c000 + dy];
    cv[r][3] = (T)c[c000 + dy + dx];
    cv[r][4] = (T)c[c000 + dz];
    cv[r][5] = (T)c[c000 + dz + dx];
    cv[r][6] = (T)c[c000 + dz + dy];
    cv[r][7] = (T)c[c000 + dz + dy + dx];
  }
#pragma unroll
  for (int cz = 0; cz < 2; ++cz) {
    if (cz && !oz) break;
    const uint32_t k = k0 + cz;
#pragma unroll
    for (int by = 0; by < 2; ++by) {
      if (by && !oy) continue;
      const uint32_t j = j0 + by;
      // y weights of the two x parents at fine (j, k); z weights of the four (x, y) parents
      T fy[2] = {T(0), T(0)}, fz[2][2] = {{T(0), T(0)}, {T(0), T(0)}};
      if (by) {
        if (OPD) {
          const uint32_t py = I + t.cn1x * (j + t.fn1y * k);
          fy[0] = t.wy[py];
          fy[1] = t.wy[py + dx];
        } else {
          fy[0] = fy[1] = t.pwy[j];
        }
      }
      if (cz) {
        if (OPD) {
          const uint32_t pz = I + t.cn1x * J + t.cplane * k;
          fz[0][0] = t.wz[pz];
          fz[1][0] = t.wz[pz + dx];
          fz[0][1] = t.wz[pz + dy];
          fz[1][1] = t.wz[pz + dy + dx];
        } else {
          fz[0][0] = fz[1][0] = fz[0][1] = fz[1][1] = t.pwz[k];
        }
      }
#pragma unroll
      for (int ax = 0; ax < 2; ++ax) {
        if (ax && !ox) continue;
        const uint32_t f = i0 + ax + t.fn1x * (j + t.fn1y * k);
        const T fx = ax ? (OPD ? t.wx[f] : t.pwx[i0 + 1]) : T(0);
        T fv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) fv[r] = (T)col_ptr((const X*)xf, ldf, c0, r, ncol)[f];
        T w[8];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a)
              w[a + 2 * b + 4 * c] = (a ? fx : T(1) - fx) * (b ? fy[a] : T(1) - fy[a]) *
                                     (c ? fz[a][b] : T(1) - fz[a][b]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (c0 + r < ncol) {
            T acc = fv[r];
#pragma unroll
            for (int q = 0; q < 8; ++q) acc += w[q] * cv[r][q];
            xf[(size_t)(c0 + r) * ldf + f] = (X)acc;
          }
        }
      }
    }
  }
}

// coarsest level: x = Ainv b (dense inverse, n x n; symmetric, so column m is read as row m
// and consecutive work-items read consecutive addresses)
// (operator sets: column col uses ainv + cset[col] * astride)
template <typename T, typename B, typename O>
__global__ __launch_bounds__(SDC_BLOCK) void k_dense_solve(const T* __restrict__ ainv, int nn,
                                                           const B* __restrict__ b,
                                                           O* __restrict__ x, size_t ld, int ncol,
                                                           const int* __restrict__ done,
                                                           const int32_t* __restrict__ cset,
                                                           size_t astride) {
  if (done && ld_flag(done)) return;
  const int row = blockIdx.x * SDC_BLOCK + threadIdx.x;
  const int col = blockIdx.y;
  if (row >= nn || col >= ncol) return;
  if (cset) ainv += (size_t)cset[col] * astride;
  const B* bc = b + (size_t)col * ld;
  T acc = T(0);
  for (int m = 0; m < nn; ++m) acc += ainv[(size_t)m * nn + row] * (T)bc[m];
  x[(size_t)col * ld + row] = (O)acc;
}

// ------------------------------------------------------------------------------------------
// PCG vector updates and reductions

template <typename T, int R>
__global__ __launch_bounds__(SDC_BLOCK) void k_dot(const T* __restrict__ a, const T* __restrict__ b,
                                                   size_t ld, uint32_t nN, int ncol,
                                                   double* __restrict__ partial,
                                                   const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  SDC_NODE_PROLOGUE(nN)
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  if (live) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (c0 + r < ncol) acc[r] = (double)col_ptr(a, ld, c0, r, ncol)[n] * (double)col_ptr(b, ld, c0, r, ncol)[n];
  }
  block_partials<R>(acc, partial, c0, ncol, nbx, bx);
}

// ------------------------------------------------------------------------------------------
// streaming PCG updates, 4 consecutive nodes per work-item (one 16-byte access per array for
// fp32), one column per blockIdx.y; the last, ragged group of nodes takes a scalar path.
#define SDC_NODES4 (SDC_BLOCK * 4)

template <typename T>
using vec4_t = T __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double block_reduce1(double v) {
  __shared__ double red1[SDC_BLOCK / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red1[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < SDC_BLOCK / 64; ++w) t += red1[w];
  return t;
}

// x += alpha p ; r -= alpha q ; partial |r|^2.  x is read from the device pointer table xp[0],
// so that a recorded iteration serves any solution buffer.
template <typename T>
__global__ __launch_bounds__(SDC_BLOCK) void k_update_xr4(T* const* xp, const T* __restrict__ p,
                                                          T* __restrict__ r, const T* __restrict__ q,
                                                          const double* __restrict__ alpha,
                                                          size_t ld, uint32_t nN,
                                                          double* __restrict__ partial,
                                                          const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  T* __restrict__ x = xp[0];
  const int c = blockIdx.y;
  const size_t o = (size_t)c * ld;
  const uint32_t n0 = (blockIdx.x * SDC_BLOCK + threadIdx.x) * 4;
  const T al = (T)alpha[c];
  double acc = 0.0;
  if (n0 + 3 < nN) {
    vec4_t<T> xv = *(const vec4_t<T>*)(x + o + n0);
    const vec4_t<T> pv = *(const vec4_t<T>*)(p + o + n0);
    vec4_t<T> rv = *(const vec4_t<T>*)(r + o + n0);
    const vec4_t<T> qv = *(const vec4_t<T>*)(q + o + n0);
    xv += al * pv;
    rv -= al * qv;
    *(vec4_t<T>*)(x + o + n0) = xv;
    *(vec4_t<T>*)(r + o + n0) = rv;
    acc = (double)rv.x * rv.x + (double)rv.y * rv.y + (double)rv.z * rv.z + (double)rv.w * rv.w;
  } else {
    for (uint32_t n = n0; n < nN && n < n0 + 4; ++n) {
      x[o + n] += al * p[o + n];
      const T rn = r[o + n] - al * q[o + n];
      r[o + n] = rn;
      acc += (double)rn * rn;
    }
  }
  const double t = block_reduce1(acc);
  if (threadIdx.x == 0) partial[(size_t)c * gridDim.x + blockIdx.x] = t;
}

// The mixed-precision PCG step that replaces "p = z + beta p; q = A p": by linearity
// q = A z + beta q, so the stencil reads z (the preconditioner output, possibly bf16) instead of
// p, and p, q are updated pointwise; partial sums of p.q per column.  The search directions
// live in a ring of slots: the new p goes to p_out, the previous one is read from p_prev, and
// the solution update x += alpha p is not carried along: the pending directions are folded
// into u (fp64) by k_ir_fold_ring, so no fp32 iterate is read and written every step.
// FIRST: p = z, q = A z.  check_stop: return at once when ctl[STOP] is set (only needed when
// several iterations run between host reads of the stop flag).
template <typename Z, int R, bool FIRST, int G = 1>
__global__ __launch_bounds__(SDC_BLOCK) void k_apply_pq(Sten<float> sa, const Z* __restrict__ z,
                                                        const float* __restrict__ p_prev,
                                                        float* __restrict__ p_out,
                                                        float* __restrict__ q,
                                                        const double* __restrict__ beta,
                                                        size_t ld, int ncol,
                                                        double* __restrict__ partial,
                                                        const int* __restrict__ ctl,
                                                        int check_stop) {
  if (check_stop && ld_flag(ctl + SDC_CTL_STOP)) return;
  SDC_NODE_PROLOGUE_C(sa.nN, sa.remap, R * G)
  double acc[R * G];
#pragma unroll
  for (int r = 0; r < R * G; ++r) acc[r] = 0.0;
  if (live) {
    NodeC<float> c = node_coef(sten_col(sa, c0), n);
#pragma unroll 2
    for (int gq = 0; gq < G; ++gq) {
      const int cg = c0 + gq * R;
      if (cg >= ncol) break;
      if (G > 1 && gq > 0 && set_changes(sa, cg, cg - R)) c = node_coef(sten_col(sa, cg), n);
      float zc[R], az[R], po[R], qo[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (!FIRST) {
          const size_t o = (size_t)(cg + r < ncol ? cg + r : ncol - 1) * ld + n;
          po[r] = ld_once(p_prev + o);
          qo[r] = ld_once(q + o);
        }
      }
      stencil_cols<float, R>(c, z, n, ld, cg, ncol, zc, az);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (cg + r < ncol) {
          const size_t o = (size_t)(cg + r) * ld + n;
          float pn = zc[r], qn = az[r];
          if (!FIRST) {
            const float be = (float)beta[cg + r];
            pn += be * po[r];
            qn += be * qo[r];
          }
          st_once(p_out + o, pn);
          st_once(q + o, qn);
          acc[gq * R + r] += (double)pn * (double)qn;
        }
      }
    }
  }
  block_partials<R * G>(acc, partial, c0, ncol, nbx, bx);
}

// r -= alpha q and per-column |r|^2 partials: a work-group takes CH chunks of SDC_NODES4 nodes
// (4 per work-item and chunk, all loads issued before the arithmetic), one partial per
// work-group, so that the single-work-group decision sums few partials per column
// (R16: also write a bfloat16 copy of r, the multigrid input)
template <bool R16, int CH>
__global__ __launch_bounds__(SDC_BLOCK) void k_update_r4(float* __restrict__ r,
                                                         const float* __restrict__ q,
                                                         const double* __restrict__ alpha,
                                                         size_t ld, uint32_t nN,
                                                         double* __restrict__ partial,
                                                         __bf16* __restrict__ r16,
                                                         const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  const int c = blockIdx.y;
  const size_t o = (size_t)c * ld;
  const float al = (float)alpha[c];
  const uint32_t base = (blockIdx.x * CH * SDC_BLOCK + threadIdx.x) * 4;
  vec4_t<float> rv[CH], qv[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const uint32_t n0 = base + k * SDC_NODES4;
    if (n0 + 3 < nN) {
      rv[k] = *(const vec4_t<float>*)(r + o + n0);
      qv[k] = *(const vec4_t<float>*)(q + o + n0);
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const uint32_t n0 = base + k * SDC_NODES4;
    if (n0 + 3 < nN) {
      const vec4_t<float> rn = rv[k] - al * qv[k];
      *(vec4_t<float>*)(r + o + n0) = rn;
      if (R16) {
        typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
        bf4 h;
        h.x = (__bf16)rn.x;
        h.y = (__bf16)rn.y;
        h.z = (__bf16)rn.z;
        h.w = (__bf16)rn.w;
        *(bf4*)(r16 + o + n0) = h;
      }
      acc += (double)rn.x * rn.x + (double)rn.y * rn.y + (double)rn.z * rn.z +
             (double)rn.w * rn.w;
    } else {
      for (uint32_t n = n0; n < nN && n < n0 + 4; ++n) {
        const float rn = r[o + n] - al * q[o + n];
        r[o + n] = rn;
        if (R16) r16[o + n] = (__bf16)rn;
        acc += (double)rn * rn;
      }
    }
  }
  const double t = block_reduce1(acc);
  if (threadIdx.x == 0) partial[(size_t)c * gridDim.x + blockIdx.x] = t;
}

#define SDC_FOLD_CHUNKS 4

// u += sum of alpha_j p_j over the pending search directions of the ring (oldest first, in
// fp64), ending at slot phi: ctl[PEND] slots phi - PEND + 1 .. phi (mod nslot), slot k at
// ptab[k] with its per-column alpha at ahist[k * ncol + c].  Runs only when the control block
// asks for a fold (reliable update or full ring).  UFRESH: u holds nothing yet (first fold of a
// solve): u = the sum without reading u, and the column padding [nN, ld) is zeroed, so the
// solve needs no separate clearing pass.  u comes from the pointer table (u = qu[1]); a
// work-group covers SDC_FOLD_CHUNKS chunks of SDC_NODES4 nodes (most launches only read the
// flags and return).
__global__ __launch_bounds__(SDC_BLOCK) void k_ir_fold_ring(const double* const* qu,
                                                            const float* const* ptab,
                                                            const double* __restrict__ ahist,
                                                            int nslot, int phi, int ncol,
                                                            size_t ld, uint32_t nN,
                                                            const int* __restrict__ ctl) {
  if (!ctl_fold(ctl)) return;
  double* __restrict__ u = const_cast<double*>(qu[1]);
  const bool ufresh = ld_flag(ctl + SDC_CTL_UFRESH) != 0;
  const int pend = ld_flag(ctl + SDC_CTL_PEND);
  const int c = blockIdx.y;
  const size_t o = (size_t)c * ld;
  const uint32_t end = ufresh ? (uint32_t)ld : nN;
#pragma unroll
  for (int ch = 0; ch < SDC_FOLD_CHUNKS; ++ch) {
    const uint32_t n0 =
        ((blockIdx.x * SDC_FOLD_CHUNKS + ch) * SDC_BLOCK + threadIdx.x) * 4;
    if (n0 + 3 < nN) {
      vec4_t<double> uv = ufresh ? vec4_t<double>(0.0) : *(const vec4_t<double>*)(u + o + n0);
      for (int j = pend - 1; j >= 0; --j) {
        const int k = (phi - j + nslot) % nslot;
        const double al = ahist[(size_t)k * ncol + c];
        const vec4_t<float> pv = *(const vec4_t<float>*)(ptab[k] + o + n0);
        uv.x += al * (double)pv.x;
        uv.y += al * (double)pv.y;
        uv.z += al * (double)pv.z;
        uv.w += al * (double)pv.w;
      }
      *(vec4_t<double>*)(u + o + n0) = uv;
    } else {
      for (uint32_t n = n0; n < end && n < n0 + 4; ++n) {
        double v = 0.0;
        if (n < nN) {
          v = ufresh ? 0.0 : u[o + n];
          for (int j = pend - 1; j >= 0; --j) {
            const int k = (phi - j + nslot) % nslot;
            v += ahist[(size_t)k * ncol + c] * (double)ptab[k][o + n];
          }
        }
        u[o + n] = v;
      }
    }
  }
}

// Start of a plain PCG: counters cleared, the iteration limit in counters[5] and tol^2 in the
// control block's double area (read by k_fin_rr), x in the pointer table.
template <typename T>
__global__ void k_pcg_init(int* __restrict__ counters, int* __restrict__ ctl,
                           T** __restrict__ xp, T* x, double tol2, int maxit) {
  if (threadIdx.x == 0) {
    for (int i = 0; i < 8; ++i) counters[i] = 0;
    counters[5] = maxit;
    ctl_dbl(ctl)[SDC_DCTL_TOL2] = tol2;
    xp[0] = x;
  }
}

// p = z + beta p (FIRST: p = z)
template <typename T, bool FIRST>
__global__ __launch_bounds__(SDC_BLOCK) void k_update_p4(const T* __restrict__ z, T* __restrict__ p,
                                                         const double* __restrict__ beta,
                                                         size_t ld, uint32_t nN,
                                                         const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  const int c = blockIdx.y;
  const size_t o = (size_t)c * ld;
  const uint32_t n0 = (blockIdx.x * SDC_BLOCK + threadIdx.x) * 4;
  const T be = FIRST ? T(0) : (T)beta[c];
  if (n0 + 3 < nN) {
    const vec4_t<T> zv = *(const vec4_t<T>*)(z + o + n0);
    if (FIRST) {
      *(vec4_t<T>*)(p + o + n0) = zv;
    } else {
      const vec4_t<T> pv = *(const vec4_t<T>*)(p + o + n0);
      *(vec4_t<T>*)(p + o + n0) = zv + be * pv;
    }
  } else {
    for (uint32_t n = n0; n < nN && n < n0 + 4; ++n)
      p[o + n] = FIRST ? z[o + n] : z[o + n] + be * p[o + n];
  }
}

template <typename T>
__global__ __launch_bounds__(SDC_BLOCK) void k_zero(T* __restrict__ x, size_t len) {
  const size_t i = (size_t)blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (i < len) x[i] = T(0);
}

// ------------------------------------------------------------------------------------------
// per-column scalar finalisation (one work-group per column, sums the block partials)

__device__ __forceinline__ double block_sum(const double* __restrict__ part, uint32_t nbx) {
  __shared__ double red[SDC_BLOCK / 64];
  double v = 0.0;
  for (uint32_t b = threadIdx.x; b < nbx; b += SDC_BLOCK) v += part[b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < SDC_BLOCK / 64; ++w) t += red[w];
  return t;
}

// out[c] = sum of partials; mode 1 additionally initialises the PCG state
__global__ __launch_bounds__(SDC_BLOCK) void k_fin_norm(const double* __restrict__ partial,
                                                        uint32_t nbx, double* __restrict__ out,
                                                        int* __restrict__ active, int init,
                                                        const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  const int c = blockIdx.x;
  const double v = block_sum(partial + (size_t)c * nbx, nbx);
  if (threadIdx.x == 0) {
    out[c] = v;
    if (init) active[c] = (v > 0.0) ? 1 : 0;
  }
}

// The mixed PCG's per-iteration finalisers run as one work-group of SDC_FIN_THREADS: a wave
// sums the block partials of SDC_FIN_COLS columns at a time (lane-strided, then a butterfly; a
// fixed order), so the partials come from kernels that write few per column.
#define SDC_FIN_THREADS 1024
#define SDC_FIN_COLS 8

// The sums of the n partials of columns c + k step, k < C (those below ncol; column j's
// partials at p + j * n) by one wave: each lane issues the loads of all C columns at 4 strides
// together (guarded), then one butterfly per column.  Every lane returns all sums.
template <int C>
__device__ __forceinline__ void wave_sums(const double* __restrict__ p, uint32_t n, int c,
                                          int step, int ncol, double (&v)[C]) {
  const uint32_t lane = threadIdx.x & 63;
  const double* q[C];
#pragma unroll
  for (int k = 0; k < C; ++k) {
    v[k] = 0.0;
    const int ck = c + k * step < ncol ? c + k * step : c;
    q[k] = p + (size_t)ck * n;
  }
  for (uint32_t b = lane; b < n; b += 256) {
    double a[C][4];
#pragma unroll
    for (int k = 0; k < C; ++k)
#pragma unroll
      for (int u = 0; u < 4; ++u) a[k][u] = b + 64 * u < n ? q[k][b + 64 * u] : 0.0;
#pragma unroll
    for (int k = 0; k < C; ++k) v[k] += (a[k][0] + a[k][1]) + (a[k][2] + a[k][3]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < C; ++k) v[k] += __shfl_xor(v[k], off, 64);
}

// |r|^2 per column (rr) from the update's partials, then the reliable-update decision: counts
// the iteration and its search direction (now pending in the ring) and asks for an update
// (replace the fp32 residual by the fp64 true residual, fold the pending directions into u)
// when every active column's |r|^2 passed tol^2 |q|^2, some active column's dropped by
// delta^2 since its last replacement (rmax), the iteration limit is reached, or `window`
// iterations passed without an update.  Otherwise, when all nslot slots of the ring are
// pending, it asks for a fold alone (the next iteration writes the oldest slot).
__global__ __launch_bounds__(SDC_FIN_THREADS) void k_norm_decide(const double* __restrict__ partial,
                                                                 uint32_t nbx,
                                                                 double* __restrict__ rr,
                                                                 int* __restrict__ ctl,
                                                                 const double* __restrict__ qn2,
                                                                 const double* __restrict__ rmax,
                                                                 const int* __restrict__ active,
                                                                 int ncol, int nslot) {
  if (ld_flag(ctl + SDC_CTL_STOP)) return;
  const double tol2 = ld_flag(ctl_dbl(ctl) + SDC_DCTL_TOL2);
  const double delta2 = ld_flag(ctl_dbl(ctl) + SDC_DCTL_DELTA2);
  const int maxit = ld_flag(ctl + SDC_CTL_MAXIT), window = ld_flag(ctl + SDC_CTL_WINDOW);
  int cand = 1, drop = 0;
  constexpr int NW = SDC_FIN_THREADS / 64;
  for (int c0 = (int)threadIdx.x >> 6; c0 < ncol; c0 += SDC_FIN_COLS * NW) {
    double vs[SDC_FIN_COLS];
    wave_sums<SDC_FIN_COLS>(partial, nbx, c0, NW, ncol, vs);
#pragma unroll
    for (int k = 0; k < SDC_FIN_COLS; ++k) {
      const int c = c0 + k * NW;
      if (c >= ncol) break;
      const double v = vs[k];
      if ((threadIdx.x & 63) == 0) rr[c] = v;
      if (!active[c]) continue;
      if (v > tol2 * qn2[c]) cand = 0;
      if (v <= delta2 * rmax[c]) drop = 1;
    }
  }
  cand = __syncthreads_and(cand);
  drop = __syncthreads_or(drop);
  if (threadIdx.x == 0) {
    const int it = ld_flag(ctl + SDC_CTL_IT) + 1;
    ctl[SDC_CTL_IT] = it;
    const int pend = ld_flag(ctl + SDC_CTL_PEND) + 1;
    ctl[SDC_CTL_PEND] = pend;
    const int upd =
        (cand || drop || it >= maxit || it - ld_flag(ctl + SDC_CTL_LAST) >= window) ? 1 : 0;
    ctl[SDC_CTL_UPD] = upd;
    ctl[SDC_CTL_FOLD] = (!upd && pend >= nslot) ? 1 : 0;
  }
}

// After a fold: the ring is empty and u is written.  After a reliable update also: the true
// residual per column from the fp64 residual's partials (orr = rmax = |q - A u|^2, active =
// orr > tol^2 |q|^2), then the stop rules: the solve stops when no column is open (converged),
// at the iteration limit, or when the largest true relative residual failed to halve since the
// previous update (the fp32 recursion stalled; the caller continues in fp64).  At the stop the
// control block is also written to hctl (pinned host memory, may be null) for the host to read.
__global__ __launch_bounds__(SDC_FIN_THREADS) void k_true_after(const double* __restrict__ partial,
                                                                uint32_t nbx,
                                                                const double* __restrict__ qn2,
                                                                double* __restrict__ orr,
                                                                double* __restrict__ rmax,
                                                                int* __restrict__ active,
                                                                int* __restrict__ ctl, int ncol,
                                                                int* __restrict__ hctl) {
  if (!ctl_fold(ctl)) return;
  if (ld_flag(ctl + SDC_CTL_UPD) == 0) {  // a fold alone (full ring)
    if (threadIdx.x == 0) {
      ctl[SDC_CTL_PEND] = 0;
      ctl[SDC_CTL_FOLD] = 0;
      ctl[SDC_CTL_UFRESH] = 0;
    }
    return;
  }
  const double tol2 = ld_flag(ctl_dbl(ctl) + SDC_DCTL_TOL2);
  const int maxit = ld_flag(ctl + SDC_CTL_MAXIT);
  constexpr int NW = SDC_FIN_THREADS / 64;
  __shared__ double sm[NW];
  __shared__ int sn[NW];
  const int lane = threadIdx.x & 63;
  double m = 0.0;
  int nact = 0;
  // 4 columns per wave at a time: this kernel holds more state (runs at reliable updates only)
  for (int c0 = (int)threadIdx.x >> 6; c0 < ncol; c0 += 4 * NW) {
    double vs[4];
    wave_sums<4>(partial, nbx, c0, NW, ncol, vs);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int c = c0 + k * NW;
      if (c >= ncol) break;
      const double v = vs[k];
      const double q2 = qn2[c];
      const int a = (q2 > 0.0 && v > tol2 * q2) ? 1 : 0;
      if (lane == 0) {
        orr[c] = v;
        rmax[c] = v;
        active[c] = a;
        nact += a;
      }
      if (q2 > 0.0) m = fmax(m, sqrt(v / q2));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    m = fmax(m, __shfl_xor(m, off, 64));
    nact += __shfl_xor(nact, off, 64);
  }
  if (lane == 0) {
    sm[threadIdx.x >> 6] = m;
    sn[threadIdx.x >> 6] = nact;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NW; ++w) {
      m = fmax(m, sm[w]);
      nact += sn[w];
    }
    double* dctl = ctl_dbl(ctl);
    const int it = ld_flag(ctl + SDC_CTL_IT);
    ctl[SDC_CTL_OUTER] = ld_flag(ctl + SDC_CTL_OUTER) + 1;
    ctl[SDC_CTL_UPD] = 0;
    ctl[SDC_CTL_PEND] = 0;
    ctl[SDC_CTL_FOLD] = 0;
    ctl[SDC_CTL_UFRESH] = 0;
    dctl[SDC_DCTL_MREL] = m;
    int reason = 0;
    if (nact == 0)
      reason = 1;
    else if (it >= maxit)
      reason = 2;
    else if (m > 0.5 * ld_flag(dctl + SDC_DCTL_MPREV))
      reason = 3;
    if (reason) {
      ctl[SDC_CTL_REASON] = reason;
      ctl[SDC_CTL_STOP] = 1;
      if (hctl) {
        // the stop flag goes last, after a system-scope fence
        for (int i = 0; i < SDC_CTL_N; ++i)
          if (i != SDC_CTL_STOP) hctl[i] = ld_flag(ctl + i);
        double* hd = (double*)(hctl + SDC_CTL_N);
        for (int i = 0; i < SDC_DCTL_N; ++i) hd[i] = ld_flag(dctl + i);
        __threadfence_system();
        __hip_atomic_store(hctl + SDC_CTL_STOP, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    } else {
      dctl[SDC_DCTL_MPREV] = m;
      ctl[SDC_CTL_LAST] = it;
    }
  }
}

// Start of the mixed PCG: the node mask of non-zero right-hand sides cleared (work-groups over
// the nodes), and in work-group 0 the control block with the solve's parameters and q, u in the
// pointer table.
__global__ __launch_bounds__(SDC_BLOCK) void k_mixed_init(unsigned* __restrict__ qmask,
                                                          uint32_t nN, int* __restrict__ ctl,
                                                          const double** __restrict__ qu,
                                                          const double* q, double* u,
                                                          double tol2, double delta2, int maxit,
                                                          int window) {
  const uint32_t n = blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (n < nN) qmask[n] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int i = 0; i < SDC_CTL_N; ++i) ctl[i] = 0;
    ctl[SDC_CTL_UFRESH] = 1;
    ctl[SDC_CTL_MAXIT] = maxit;
    ctl[SDC_CTL_WINDOW] = window;
    double* d = ctl_dbl(ctl);
    d[SDC_DCTL_MPREV] = 1.0;
    d[SDC_DCTL_MREL] = 0.0;
    d[SDC_DCTL_TOL2] = tol2;
    d[SDC_DCTL_DELTA2] = delta2;
    qu[0] = q;
    qu[1] = u;
  }
}

// |q|^2 per column into qn2 and rmax (the residual at the last replacement: q, u = 0) and
// active = |q| > 0, from the first residual's partials (one work-group per column)
__global__ __launch_bounds__(SDC_BLOCK) void k_fin_qnorm(const double* __restrict__ partial,
                                                         uint32_t nbx, double* __restrict__ qn2,
                                                         double* __restrict__ rmax,
                                                         int* __restrict__ active) {
  const int c = blockIdx.x;
  const double v = block_sum(partial + (size_t)c * nbx, nbx);
  if (threadIdx.x == 0) {
    qn2[c] = v;
    rmax[c] = v;
    active[c] = (v > 0.0) ? 1 : 0;
  }
}

// alpha = rz / (p.Ap) on active columns, 0 elsewhere
__global__ __launch_bounds__(SDC_BLOCK) void k_fin_alpha(const double* __restrict__ partial,
                                                         uint32_t nbx,
                                                         const double* __restrict__ rz,
                                                         double* __restrict__ alpha,
                                                         const int* __restrict__ active,
                                                         const int* __restrict__ done,
                                                         double* __restrict__ ahist) {
  if (done && ld_flag(done)) return;
  const int c = blockIdx.x;
  const double pq = block_sum(partial + (size_t)c * nbx, nbx);
  if (threadIdx.x == 0) {
    const double a = (active[c] && pq > 0.0) ? rz[c] / pq : 0.0;
    alpha[c] = a;
    if (ahist) ahist[c] = a;  // the ring slot's alpha (mixed PCG)
  }
}

// convergence test |r| <= tol |b| per column (tol^2 from the control block ctl); the last
// work-group counts the iteration (counters[4]) and publishes the number of still-active columns
// and the stop flag: all converged, or counters[5] iterations done (device copy and host-mapped
// copy)
__global__ __launch_bounds__(SDC_BLOCK) void k_fin_rr(const double* __restrict__ partial,
                                                      uint32_t nbx,
                                                      const double* __restrict__ bnorm2,
                                                      const int* __restrict__ ctl,
                                                      double* __restrict__ rr,
                                                      int* __restrict__ active,
                                                      int* __restrict__ counters, int* done,
                                                      int* host_flag) {
  if (ld_flag(done)) return;
  const double tol2 = ld_flag(ctl_dbl(ctl) + SDC_DCTL_TOL2);
  const int c = blockIdx.x;
  const double v = block_sum(partial + (size_t)c * nbx, nbx);
  if (threadIdx.x == 0) {
    rr[c] = v;
    int a = active[c];
    if (a && v <= tol2 * bnorm2[c]) a = 0;
    active[c] = a;
    if (a) atomicAdd(&counters[0], 1);
    __threadfence();
    const int ticket = atomicAdd(&counters[1], 1);
    if (ticket == (int)gridDim.x - 1) {
      const int nact = atomicAdd(&counters[0], 0);
      const int it = ld_flag(counters + 4) + 1;
      counters[4] = it;
      const int fin = (nact == 0 || it >= ld_flag(counters + 5)) ? 1 : 0;
      *done = fin;
      counters[2] = nact;
      counters[0] = 0;
      counters[1] = 0;
      if (host_flag) {
        host_flag[0] = fin;
        host_flag[1] = nact;
        __threadfence_system();
      }
    }
  }
}

// beta = rz_new / rz_old on active columns
__global__ __launch_bounds__(SDC_BLOCK) void k_fin_beta(const double* __restrict__ partial,
                                                        uint32_t nbx, double* __restrict__ rz,
                                                        double* __restrict__ beta,
                                                        const int* __restrict__ active, int first,
                                                        const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  const int c = blockIdx.x;
  const double v = block_sum(partial + (size_t)c * nbx, nbx);
  if (threadIdx.x == 0) {
    const double old = rz[c];
    beta[c] = (!first && active[c] && old != 0.0) ? v / old : 0.0;
    rz[c] = v;
  }
}

// active[c] &= mask[c]
__global__ __launch_bounds__(SDC_BLOCK) void k_and_active(int* __restrict__ active,
                                                          const int* __restrict__ mask, int ncol) {
  const int c = blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (c < ncol && !mask[c]) active[c] = 0;
}

// refinement convergence: mask[c] = |r_c|^2 > tol^2 |q_c|^2; counters[0] += number still open
__global__ __launch_bounds__(SDC_BLOCK) void k_outer_check(const double* __restrict__ rr,
                                                           const double* __restrict__ qn2,
                                                           double tol2, int* __restrict__ mask,
                                                           int* __restrict__ nopen, int ncol) {
  const int c = blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (c >= ncol) return;
  const int open = rr[c] > tol2 * qn2[c] ? 1 : 0;
  mask[c] = open;
  if (open) atomicAdd(nopen, 1);
}

// ------------------------------------------------------------------------------------------
// fp64 iterative refinement

struct PinFix {
  int on;      // Neumann gauge: rows next to node 0 get the column-0 term moved to the rhs
  uint32_t n1x, plane;
  const double* cx;  // level-0 conductances; the node-0 edges are cx[0], cy[0], cz[0]
  const double* cy;
  const double* cz;
};

// r = q - A' u (A' = reference operator: identity row 0 for Neumann), written in type TO;
// with U_ZERO the u loads are skipped (first refinement step).  The rows next to node 0 are
// shifted into the symmetric (identity row and column 0) form the inner PCG solves.
// Partial sums of |r|^2 (before the shift) per column.
// qmask (per node, never null): with U_ZERO it is set to 1 on every node where some column of
// q is non-zero (the caller clears it first); otherwise q is read only on flagged nodes.  A
// right-hand side of point sources touches a few nodes, so the fp64 residual then streams u.
// ctl non-null: runs only when the control block asks for a reliable update, and q, u are read
// from qu[0], qu[1] (device pointer table) instead of the arguments, so that a recorded
// iteration serves every solve.  CH: a work-group covers CH consecutive blocks of SDC_BLOCK
// nodes (one partial per work-group and column), so that a launch that returns at once
// dispatches fewer work-groups.
template <typename TO, int R, bool U_ZERO, int G = 1, int CH = 1>
__global__ __launch_bounds__(SDC_BLOCK) void k_ir_residual(Sten<double> sa, const double* q,
                                                           const double* u,
                                                           TO* __restrict__ rout, size_t ld,
                                                           int ncol, PinFix pf0,
                                                           double* __restrict__ partial,
                                                           unsigned* __restrict__ qmask,
                                                           __bf16* __restrict__ r16,
                                                           const int* __restrict__ ctl,
                                                           const double* const* qu) {
  if (ctl) {
    if (!ctl_update(ctl)) return;
    q = qu[0];
    u = qu[1];
  }
  const uint32_t nbx = gridDim.x;
  const uint32_t lin0_ = blockIdx.x + nbx * blockIdx.y;
  const uint32_t lin_ = sa.remap ? xcd_remap(lin0_, nbx * gridDim.y) : lin0_;
  const uint32_t by_ = lin_ / nbx;
  const uint32_t bx = lin_ - by_ * nbx;
  const int c0 = (int)by_ * (R * G);
  // the operator (and the node-0 edges of the Neumann shift) of column cg's set
  auto select = [&](int cg, PinFix& pf) {
    pf = pf0;
    if (sa.cset) {
      const size_t o = (size_t)sa.cset[cg] * sa.sstride;
      pf.cx += o;
      pf.cy += o;
      pf.cz += o;
    }
    return sten_col(sa, cg);
  };
  double acc[R * G];
#pragma unroll
  for (int r = 0; r < R * G; ++r) acc[r] = 0.0;
  for (int ch = 0; ch < CH; ++ch) {
    const uint32_t n = (bx * CH + ch) * SDC_BLOCK + threadIdx.x;
    if (n >= sa.nN) break;
    PinFix pf;
    Sten<double> s = select(c0, pf);
    NodeC<double> c;
    if (!U_ZERO) c = node_coef(s, n);
    // Load order: the mask word, then the 7R stencil loads of u, then the q loads that depend
    // on the mask.  Issued the other way round, the branch on the mask held the stencil loads
    // back by one memory round trip per wavefront.
    const unsigned qm = U_ZERO ? 1u : qmask[n];
    bool anyq = false;
#pragma unroll 2
    for (int gq = 0; gq < G; ++gq) {
      const int cg = c0 + gq * R;
      if (cg >= ncol) break;
      if (G > 1 && gq > 0 && set_changes(sa, cg, cg - R)) {
        s = select(cg, pf);
        if (!U_ZERO) c = node_coef(s, n);
      }
      double wfix = 0.0;
      if (pf.on) {
        if (n == 1) wfix = pf.cx[0];
        else if (n == pf.n1x) wfix = pf.cy[0];
        else if (n == pf.plane) wfix = pf.cz[0];
      }
      double qv[R], uc[R], au[R];
      if (!U_ZERO) stencil_cols<double, R>(c, u, n, ld, cg, ncol, uc, au);
      const bool readq = qm != 0u;
#pragma unroll
      for (int r = 0; r < R; ++r) qv[r] = readq ? ld_once(col_ptr(q, ld, cg, r, ncol) + n) : 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (cg + r < ncol) {
          const size_t o = (size_t)(cg + r) * ld;
          double rv = qv[r];
          anyq |= rv != 0.0;
          if (!U_ZERO) rv -= au[r];
          acc[gq * R + r] += rv * rv;
          if (wfix != 0.0) {
            const double r0 = U_ZERO ? q[o] : q[o] - u[o];
            rv += wfix * r0;
          }
          st_once(rout + o + n, (TO)rv);
          if (r16) r16[o + n] = (__bf16)rv;
        }
    }
    // every writer stores the same value; the mask is cleared before this pass
    if (U_ZERO && anyq) qmask[n] = 1u;
  }
  block_partials<R * G>(acc, partial, c0, ncol, nbx, bx);
}

// u += e
template <typename TE, int R>
__global__ __launch_bounds__(SDC_BLOCK) void k_ir_update(double* __restrict__ u,
                                                         const TE* __restrict__ e, size_t ld,
                                                         uint32_t nN, int ncol) {
  SDC_NODE_PROLOGUE(nN)
  if (!live) return;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (c0 + r < ncol) {
      const size_t o = (size_t)(c0 + r) * ld + n;
      u[o] += (double)e[o];
    }
}

// ------------------------------------------------------------------------------------------
// coefficient construction

struct Geo {
  const double* hx;
  const double* hy;
  const double* hz;
  // Robin face weights beta_f = alpha_f |f| / 4 on the six boundary planes
  const double* bxm;  // x = xmin faces, index j + ny*k
  const double* bxp;
  const double* bym;  // index i + nx*k
  const double* byp;
  const double* bzm;  // index i + nx*j
  const double* bzp;
  int robin;
  uint32_t nx, ny, nz, n1x, n1y, nN, plane;
  FastDiv fx, fy;     // by n1x, n1y (nodes)
  FastDiv cfx, cfy;   // by nx, ny (cells)
};

// Edge conductances of the +x/+y/+z edges of every node and the Robin diagonal, from a cell
// property sig (sigma, or a perturbation of it; A is linear in sigma).  blockIdx.z selects an
// operator set: sig + z * sig_stride in, arrays + z * ostride out.
//   x edge (i,j,k)-(i+1,j,k): sum over the <=4 cells (i, j', k') touching it of
//   sig * hy_j' * hz_k' / (4 hx_i), i.e. M_e = diag(A_v sig) with A_v[e,c] = V_c/4 divided by hx_i^2
//   Robin: b_n = sum over boundary faces f containing n of beta_f * sig(cell of f)
__global__ __launch_bounds__(SDC_BLOCK) void k_coef_from_sigma(Geo g, const double* __restrict__ sig,
                                                               double* __restrict__ cx,
                                                               double* __restrict__ cy,
                                                               double* __restrict__ cz,
                                                               double* __restrict__ bd,
                                                               double scale, size_t sig_stride,
                                                               size_t ostride) {
  const uint32_t n = blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (n >= g.nN) return;
  // operator set blockIdx.z: its cell property and its coefficient arrays
  const size_t zi = (size_t)blockIdx.z * sig_stride, zo = (size_t)blockIdx.z * ostride;
  sig += zi;
  cx += zo;
  cy += zo;
  cz += zo;
  bd += zo;
  const uint32_t r = fdiv(n, g.fx);
  const int i = (int)(n - r * g.n1x);
  const int k = (int)fdiv(r, g.fy);
  const int j = (int)(r - (uint32_t)k * g.n1y);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  // s[a][b][c] = sig of cell (i-1+a, j-1+b, k-1+c), 0 outside the mesh
  double s[2][2][2];
  double hxa[2], hyb[2], hzc[2];
  for (int a = 0; a < 2; ++a) {
    const int ii = i - 1 + a;
    hxa[a] = (ii >= 0 && ii < nx) ? g.hx[ii] : 0.0;
  }
  for (int b = 0; b < 2; ++b) {
    const int jj = j - 1 + b;
    hyb[b] = (jj >= 0 && jj < ny) ? g.hy[jj] : 0.0;
  }
  for (int c = 0; c < 2; ++c) {
    const int kk = k - 1 + c;
    hzc[c] = (kk >= 0 && kk < nz) ? g.hz[kk] : 0.0;
  }
  for (int c = 0; c < 2; ++c)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) {
        const int ii = i - 1 + a, jj = j - 1 + b, kk = k - 1 + c;
        const bool ok = ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz;
        s[a][b][c] = ok ? sig[(size_t)ii + (size_t)nx * ((size_t)jj + (size_t)ny * kk)] : 0.0;
      }
  double wx = 0.0, wy = 0.0, wz = 0.0;
  if (i < nx) {
    double t = 0.0;
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) t += s[1][b][c] * hyb[b] * hzc[c];
    wx = t / (4.0 * hxa[1]);
  }
  if (j < ny) {
    double t = 0.0;
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c) t += s[a][1][c] * hxa[a] * hzc[c];
    wy = t / (4.0 * hyb[1]);
  }
  if (k < nz) {
    double t = 0.0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) t += s[a][b][1] * hxa[a] * hyb[b];
    wz = t / (4.0 * hzc[1]);
  }
  double rb = 0.0;
  if (g.robin) {
    if (i == 0 || i == nx) {
      const double* bf = (i == 0) ? g.bxm : g.bxp;
      const int a = (i == 0) ? 1 : 0;
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c) {
          const int jj = j - 1 + b, kk = k - 1 + c;
          if (jj >= 0 && jj < ny && kk >= 0 && kk < nz) rb += bf[jj + ny * kk] * s[a][b][c];
        }
    }
    if (j == 0 || j == ny) {
      const double* bf = (j == 0) ? g.bym : g.byp;
      const int b = (j == 0) ? 1 : 0;
      for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c) {
          const int ii = i - 1 + a, kk = k - 1 + c;
          if (ii >= 0 && ii < nx && kk >= 0 && kk < nz) rb += bf[ii + nx * kk] * s[a][b][c];
        }
    }
    if (k == 0 || k == nz) {
      const double* bf = (k == 0) ? g.bzm : g.bzp;
      const int c = (k == 0) ? 1 : 0;
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          const int ii = i - 1 + a, jj = j - 1 + b;
          if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) rb += bf[ii + nx * jj] * s[a][b][c];
        }
    }
  }
  cx[n] = scale * wx;
  cy[n] = scale * wy;
  cz[n] = scale * wz;
  bd[n] = scale * rb;
}

// fp32 copies of a level's coefficients (null: none)
struct Coef32 {
  float* cx;
  float* cy;
  float* cz;
  float* dg;
  float* idg;
};

// One level of k_coef_diag: its shape, coefficient arrays (one operator per set, stride
// entries apart), its first work-group in the launch, and pin = 1 to make node 0 an identity row.
struct DiagLevel {
  uint32_t n1x, n1y, nN, blk0;
  FastDiv fx, fy;
  const double *cx, *cy, *cz, *bd;
  double *dg, *idg;
  size_t stride;
  Coef32 c32;
  int pin;
};
#define SDC_DIAG_LEVELS 12
struct DiagLevels {
  DiagLevel lv[SDC_DIAG_LEVELS];
  int n;
};

// Full diagonal and its inverse of up to SDC_DIAG_LEVELS levels in one launch (work-groups
// [lv[l].blk0, lv[l+1].blk0) take level l, blockIdx.z the operator set); with c32.cx set, also
// the fp32 copies of cx, cy, cz, dg, idg (the fp32 solver's operator) in the same pass.
__global__ __launch_bounds__(SDC_BLOCK) void k_coef_diag(DiagLevels tab) {
  int l = 0;
  while (l + 1 < tab.n && blockIdx.x >= tab.lv[l + 1].blk0) ++l;
  const DiagLevel& v = tab.lv[l];
  const uint32_t n = (blockIdx.x - v.blk0) * SDC_BLOCK + threadIdx.x;
  if (n >= v.nN) return;
  const size_t zo = (size_t)blockIdx.z * v.stride;  // operator set
  const double* cx = v.cx + zo;
  const double* cy = v.cy + zo;
  const double* cz = v.cz + zo;
  const uint32_t r = fdiv(n, v.fx);
  const uint32_t i = n - r * v.n1x;
  const uint32_t k = fdiv(r, v.fy);
  const uint32_t j = r - k * v.n1y;
  const uint32_t plane = v.n1x * v.n1y;
  double d = cx[n] + cy[n] + cz[n] + v.bd[zo + n];
  if (i > 0) d += cx[n - 1];
  if (j > 0) d += cy[n - v.n1x];
  if (k > 0) d += cz[n - plane];
  if (v.pin && n == 0) d = 1.0;
  v.dg[zo + n] = d;
  v.idg[zo + n] = 1.0 / d;
  if (v.c32.cx) {
    v.c32.cx[zo + n] = (float)cx[n];
    v.c32.cy[zo + n] = (float)cy[n];
    v.c32.cz[zo + n] = (float)cz[n];
    v.c32.dg[zo + n] = (float)d;
    v.c32.idg[zo + n] = (float)(1.0 / d);
  }
}

// Grids and outputs of the operator-dependent interpolation weights of a transfer's three
// factors (axis a: a g0[a] x g1[a] x g2[a] grid, weights of operator set z at w64[a] + z *
// wstride[a] and their fp32 copies at w32[a] + z * wstride[a]).
struct InterpGrids {
  double* w64[3];
  float* w32[3];
  size_t wstride[3];
  uint32_t g0[3], g1[3], g2[3];
};

// Coarse-level conductances from fine ones (blockIdx.y = 0): along the edge direction the (one
// or two) fine edges between the two coarse nodes act in series; across it the fine edge lines
// are summed with the interpolation weights of the transverse axes.  The Robin diagonal is
// restricted like a vector (lumped Galerkin P^T b), so A_c 1 = P^T A_f 1.  In the same launch
// (blockIdx.y = 1 + axis) the operator-dependent interpolation weights of the transfer's
// factors, which read the same fine conductances.
__global__ __launch_bounds__(SDC_BLOCK) void k_coarsen_coef(Xfer<double> t,
                                                            const double* __restrict__ fcx,
                                                            const double* __restrict__ fcy,
                                                            const double* __restrict__ fcz,
                                                            const double* __restrict__ fbd,
                                                            double* __restrict__ ccx,
                                                            double* __restrict__ ccy,
                                                            double* __restrict__ ccz,
                                                            double* __restrict__ cbd,
                                                            uint32_t cn1z, size_t fstride,
                                                            size_t cstride, InterpGrids iw) {
  if (blockIdx.y > 0) {
    const int ax = (int)blockIdx.y - 1;
    const size_t m = (size_t)blockIdx.x * SDC_BLOCK + threadIdx.x;
    if (m >= (size_t)iw.g0[ax] * iw.g1[ax] * iw.g2[ax]) return;
    const size_t zf = (size_t)blockIdx.z * fstride, zw = (size_t)blockIdx.z * iw.wstride[ax];
    if (ax == 0)
      interp_weight<0>(t, m, iw.g0[0], iw.g1[0], fcx + zf, iw.w64[0] + zw, iw.w32[0] + zw);
    else if (ax == 1)
      interp_weight<1>(t, m, iw.g0[1], iw.g1[1], fcy + zf, iw.w64[1] + zw, iw.w32[1] + zw);
    else
      interp_weight<2>(t, m, iw.g0[2], iw.g1[2], fcz + zf, iw.w64[2] + zw, iw.w32[2] + zw);
    return;
  }
  const uint32_t n = blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (n >= t.cN) return;
  // operator set blockIdx.z: fine arrays + z * fstride, coarse arrays + z * cstride
  const size_t zf = (size_t)blockIdx.z * fstride, zc = (size_t)blockIdx.z * cstride;
  fcx += zf;
  fcy += zf;
  fcz += zf;
  fbd += zf;
  ccx += zc;
  ccy += zc;
  ccz += zc;
  cbd += zc;
  const uint32_t rr = fdiv(n, t.cfx);
  const uint32_t I = n - rr * t.cn1x;
  const uint32_t K = fdiv(rr, t.cfy);
  const uint32_t J = rr - K * t.cn1y;
  const int fi = t.rix[I], fj = t.riy[J], fk = t.riz[K];
  const double wx[3] = {t.rlx[I], 1.0, t.rhx[I]};
  const double wy[3] = {t.rly[J], 1.0, t.rhy[J]};
  const double wz[3] = {t.rlz[K], 1.0, t.rhz[K]};
  auto fidx = [&](int a, int b, int c) -> size_t {
    return (size_t)a + (size_t)b * t.fn1x + (size_t)c * t.fplane;
  };
  auto series = [](double c1, double c2) -> double {
    const double s = c1 + c2;
    return s > 0.0 ? c1 * c2 / s : 0.0;
  };
  // x edges
  double sx = 0.0;
  if (I + 1 < t.cn1x) {
    const int span = t.rix[I + 1] - fi;
    for (int c = 0; c < 3; ++c) {
      if (wz[c] == 0.0) continue;
      for (int b = 0; b < 3; ++b) {
        if (wy[b] == 0.0) continue;
        const size_t f = fidx(fi, fj + b - 1, fk + c - 1);
        const double e = (span == 2) ? series(fcx[f], fcx[f + 1]) : fcx[f];
        sx += wy[b] * wz[c] * e;
      }
    }
  }
  double sy = 0.0;
  if (J + 1 < t.cn1y) {
    const int span = t.riy[J + 1] - fj;
    for (int c = 0; c < 3; ++c) {
      if (wz[c] == 0.0) continue;
      for (int a = 0; a < 3; ++a) {
        if (wx[a] == 0.0) continue;
        const size_t f = fidx(fi + a - 1, fj, fk + c - 1);
        const double e = (span == 2) ? series(fcy[f], fcy[f + t.fn1x]) : fcy[f];
        sy += wx[a] * wz[c] * e;
      }
    }
  }
  double sz = 0.0;
  if (K + 1 < cn1z) {
    const int span = t.riz[K + 1] - fk;
    for (int b = 0; b < 3; ++b) {
      if (wy[b] == 0.0) continue;
      for (int a = 0; a < 3; ++a) {
        if (wx[a] == 0.0) continue;
        const size_t f = fidx(fi + a - 1, fj + b - 1, fk);
        const double e = (span == 2) ? series(fcz[f], fcz[f + t.fplane]) : fcz[f];
        sz += wx[a] * wy[b] * e;
      }
    }
  }
  double sb = 0.0;
  for (int c = 0; c < 3; ++c) {
    if (wz[c] == 0.0) continue;
    for (int b = 0; b < 3; ++b) {
      if (wy[b] == 0.0) continue;
      for (int a = 0; a < 3; ++a) {
        if (wx[a] == 0.0) continue;
        sb += wx[a] * wy[b] * wz[c] * fbd[fidx(fi + a - 1, fj + b - 1, fk + c - 1)];
      }
    }
  }
  ccx[n] = sx;
  ccy[n] = sy;
  ccz[n] = sz;
  cbd[n] = sb;
}

// Inverse of the coarsest level's operator, one work-group of 1024 per operator set (blockIdx.x:
// coefficients + set * s.sstride, matrices + set * astride, scratch + set * nn):
//   1. the dense operator a (row-major nn x nn; pin = 1: identity row / column 0), one row per
//      work-item;
//   2. in-place Gauss-Jordan elimination without pivoting (the operator is SPD), every pivot
//      step updating all nn*nn entries of a and inv in parallel (f: the elimination factors);
//   3. inv <- (inv + inv^T) / 2 into inv64 and its fp32 copy inv32.
// SH: a, inv and f live in LDS ((2 nn^2 + nn) doubles of dynamic shared memory), so the pivot
// steps do not go through L2; otherwise in the global scratch ga, gf and in inv64.
template <bool SH>
__global__ __launch_bounds__(1024) void k_coarse_inverse(Sten<double> s, double* __restrict__ ga,
                                                         double* __restrict__ inv64,
                                                         float* __restrict__ inv32,
                                                         double* __restrict__ gf,
                                                         size_t astride) {
  extern __shared__ __align__(16) double cinv_smem[];
  const size_t zo = (size_t)blockIdx.x * s.sstride;
  s.cx += zo;
  s.cy += zo;
  s.cz += zo;
  s.dg += zo;
  s.idg += zo;
  inv64 += (size_t)blockIdx.x * astride;
  inv32 += (size_t)blockIdx.x * astride;
  const int N = (int)s.nN;
  const int NN = N * N;
  double* a = SH ? cinv_smem : ga + (size_t)blockIdx.x * astride;
  double* inv = SH ? cinv_smem + NN : inv64;
  double* f = SH ? cinv_smem + 2 * NN : gf + (size_t)blockIdx.x * N;
  const int tid = threadIdx.x;
  for (int n = tid; n < N; n += 1024) {
    double* row = a + (size_t)n * N;
    for (int m = 0; m < N; ++m) row[m] = 0.0;
    const NodeC<double> c = node_coef(s, (uint32_t)n);
    row[n] = c.d;
    if (c.ixm != (uint32_t)n) row[c.ixm] = -c.wxm;
    if (c.ixp != (uint32_t)n) row[c.ixp] = -c.wxp;
    if (c.iym != (uint32_t)n) row[c.iym] = -c.wym;
    if (c.iyp != (uint32_t)n) row[c.iyp] = -c.wyp;
    if (c.izm != (uint32_t)n) row[c.izm] = -c.wzm;
    if (c.izp != (uint32_t)n) row[c.izp] = -c.wzp;
  }
  for (int e = tid; e < NN; e += 1024) inv[e] = ((e / N) == (e % N)) ? 1.0 : 0.0;
  __syncthreads();
  for (int p = 0; p < N; ++p) {
    const double ip = 1.0 / a[p * N + p];
    __syncthreads();
    for (int m = tid; m < N; m += 1024) {
      a[p * N + m] *= ip;
      inv[p * N + m] *= ip;
    }
    __syncthreads();
    for (int r = tid; r < N; r += 1024) f[r] = (r == p) ? 0.0 : a[r * N + p];
    __syncthreads();
    for (int e = tid; e < NN; e += 1024) {
      const int r = e / N;
      const int m = e - r * N;
      const double fr = f[r];
      if (fr != 0.0) {
        a[e] -= fr * a[p * N + m];
        inv[e] -= fr * inv[p * N + m];
      }
    }
    __syncthreads();
  }
  // each unordered pair (r, c) is read and written by the one work-item with c >= r
  for (int e = tid; e < NN; e += 1024) {
    const int r = e / N, c = e % N;
    if (c < r) continue;
    const double v = 0.5 * (inv[r * N + c] + inv[c * N + r]);
    inv64[r * N + c] = v;
    inv64[c * N + r] = v;
    inv32[r * N + c] = (float)v;
    inv32[c * N + r] = (float)v;
  }
}

// ------------------------------------------------------------------------------------------
// adjoint contraction (Jtvec):
//   g_c = sum_j  (hy hz / 4hx) sum_{4 x-edges} du dl + (hx hz / 4hy) sum_{4 y-edges} du dl
//              + (hx hy / 4hz) sum_{4 z-edges} du dl
//              + sum_{boundary faces f of c} beta_f sum_{4 nodes of f} u l
// (the derivative of u^T A(sigma) lambda with respect to the cell value sigma_c).
// (per-column output for getJ, weighted squares for getJtJdiag, or the sum over columns)

// columns per work-item of the adjoint contraction
#ifndef SDC_JT_COLS
#define SDC_JT_COLS 8
#endif

// per-cell factors of the contraction: edge weights hy hz / 4hx ... and the Robin face weights
struct JtCell {
  double fxw, fyw, fzw, bx0, bx1, by0, by1, bz0, bz1;
};

// g_c of one column from the 8 corner values of u and lambda (corner m = a + 2b + 4c)
__device__ __forceinline__ double jt_cell_value(const JtCell& w, bool robin, const double (&uv)[8],
                                                const double (&lv)[8]) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    // x edges: pairs (0,1),(2,3),(4,5),(6,7)
    sx += (uv[2 * q + 1] - uv[2 * q]) * (lv[2 * q + 1] - lv[2 * q]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    // y edges: (0,2),(1,3),(4,6),(5,7)
    const int m0 = (q & 1) + 4 * (q >> 1);
    sy += (uv[m0 + 2] - uv[m0]) * (lv[m0 + 2] - lv[m0]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    // z edges: (m, m+4)
    sz += (uv[q + 4] - uv[q]) * (lv[q + 4] - lv[q]);
  }
  double gval = w.fxw * sx + w.fyw * sy + w.fzw * sz;
  if (robin) {
    double ul[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) ul[m] = uv[m] * lv[m];
    gval += w.bx0 * (ul[0] + ul[2] + ul[4] + ul[6]) + w.bx1 * (ul[1] + ul[3] + ul[5] + ul[7]);
    gval += w.by0 * (ul[0] + ul[1] + ul[4] + ul[5]) + w.by1 * (ul[2] + ul[3] + ul[6] + ul[7]);
    gval += w.bz0 * (ul[0] + ul[1] + ul[2] + ul[3]) + w.bz1 * (ul[4] + ul[5] + ul[6] + ul[7]);
  }
  return gval;
}

// MODE 0: out[c] = sum_j g_j;  MODE 1: out[j*ldo + c] = g_j;
// MODE 2: out[c] += sum_j wsq[j] g_j^2 (diagonal of J^T W^2 J before the map derivative).
// Grid (cell blocks, column groups of SDC_JT_COLS columns): a work-item sums its group's
// columns.  The column groups keep the work-items that share a node column in flight
// together, so the 8 cells around a node read it from L2 (one long loop over every column let
// the cell planes drift apart and fetched the shared node planes from HBM ~2.7 times).  With several groups,
// MODE 0 / 2 write one partial per group to part[g*nC + c], summed in group order by
// k_jtvec_reduce (deterministic); with one group they write out directly.
template <int MODE>
__global__ __launch_bounds__(SDC_BLOCK) void k_jtvec(Geo g, const double* __restrict__ u,
                                                     const double* __restrict__ lam, size_t ld,
                                                     int ncol, const int* __restrict__ ucol,
                                                     double* __restrict__ out, size_t ldo,
                                                     const double* __restrict__ wsq,
                                                     double* __restrict__ part) {
  const uint32_t nC = g.nx * g.ny * g.nz;
  const uint32_t nbx = gridDim.x;
  const uint32_t c = xcd_remap(blockIdx.x, nbx) * SDC_BLOCK + threadIdx.x;
  if (c >= nC) return;
  constexpr int G = SDC_JT_COLS;
  const int cbeg = (int)blockIdx.y * G;
  const int nc = min(ncol - cbeg, G);
  const uint32_t r = fdiv(c, g.cfx);
  const uint32_t i = c - r * g.nx;
  const uint32_t k = fdiv(r, g.cfy);
  const uint32_t j = r - k * g.ny;
  const double hx = g.hx[i], hy = g.hy[j], hz = g.hz[k];
  const double fxw = hy * hz / (4.0 * hx), fyw = hx * hz / (4.0 * hy), fzw = hx * hy / (4.0 * hz);
  double bx0 = 0.0, bx1 = 0.0, by0 = 0.0, by1 = 0.0, bz0 = 0.0, bz1 = 0.0;
  if (g.robin) {
    // unconditional loads (the face index is in range for every cell) masked by multiplying,
    // so that they go out with the first column's loads
    bx0 = g.bxm[j + g.ny * k] * (double)(i == 0);
    bx1 = g.bxp[j + g.ny * k] * (double)(i == g.nx - 1);
    by0 = g.bym[i + g.nx * k] * (double)(j == 0);
    by1 = g.byp[i + g.nx * k] * (double)(j == g.ny - 1);
    bz0 = g.bzm[i + g.nx * j] * (double)(k == 0);
    bz1 = g.bzp[i + g.nx * j] * (double)(k == g.nz - 1);
  }
  const size_t n000 = (size_t)i + (size_t)j * g.n1x + (size_t)k * g.plane;
  size_t off[8];
  for (int cc = 0; cc < 2; ++cc)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) off[a + 2 * b + 4 * cc] = n000 + a + b * g.n1x + cc * g.plane;
  // the group's column indices and weights first, so that no load in the loop waits on
  // another; then two columns per step (32 loads in flight per work-item)
  int scol[G];
  double wq[G];
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int col = cbeg + (q < nc ? q : nc - 1);
    scol[q] = ucol ? ucol[col] : col;
    wq[q] = MODE == 2 ? wsq[col] : 0.0;
  }
  const JtCell jc{fxw, fyw, fzw, bx0, bx1, by0, by1, bz0, bz1};
  double total = 0.0;
#pragma unroll
  for (int q = 0; q < G; q += 2) {
    if (q >= nc) break;
    const int qb = q + 1 < nc ? q + 1 : q;
    const double* ua = u + (size_t)scol[q] * ld;
    const double* la = lam + (size_t)(cbeg + q) * ld;
    const double* ub = u + (size_t)scol[qb] * ld;
    const double* lb = lam + (size_t)(cbeg + qb) * ld;
    double uva[8], lva[8], uvb[8], lvb[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      uva[m] = ua[off[m]];
      lva[m] = la[off[m]];
      uvb[m] = ub[off[m]];
      lvb[m] = lb[off[m]];
    }
    const double gv[2] = {jt_cell_value(jc, g.robin, uva, lva), jt_cell_value(jc, g.robin, uvb, lvb)};
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      if (q + e >= nc) break;
      if (MODE == 1)
        out[(size_t)(cbeg + q + e) * ldo + c] = gv[e];
      else if (MODE == 2)
        total += wq[q + e] * gv[e] * gv[e];
      else
        total += gv[e];
    }
  }
  if (MODE != 1) {
    if (part)
      part[(size_t)blockIdx.y * nC + c] = total;
    else if (MODE == 0)
      out[c] = total;
    else
      out[c] += total;
  }
}

// Sum of the ngroups partials of cell c, in group order.
//   MODE 0: out[c] = sum;  MODE 1: out[c] += sum;
//   MODE 2: out[c] = (...((out[c] + p_0) + p_1) ...) -- the running sum of earlier groups
//           continued, so that groups reduced in several passes add up in the same order (and
//           to the same bits) as one pass over all of them.
template <int MODE>
__global__ __launch_bounds__(SDC_BLOCK) void k_jtvec_reduce(const double* __restrict__ part,
                                                            int ngroups, uint32_t nC,
                                                            double* __restrict__ out) {
  const uint32_t c = blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (c >= nC) return;
  double t = MODE == 2 ? out[c] : 0.0;
  for (int g = 0; g < ngroups; ++g) t += ld_once(part + (size_t)g * nC + c);
  out[c] = MODE == 1 ? out[c] + t : t;
}

// ------------------------------------------------------------------------------------------
// sparse projection to the receivers and its transpose (fp64)

// d[i] = sum_{e in row i} w[e] * u[idx[e]]   (idx = flat index col*ld + node)
__global__ __launch_bounds__(SDC_BLOCK) void k_gather(int nrow, const int* __restrict__ rowptr,
                                                      const int64_t* __restrict__ idx,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ u,
                                                      double* __restrict__ d) {
  const int i = blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (i >= nrow) return;
  double acc = 0.0;
  for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) acc += w[e] * u[idx[e]];
  d[i] = acc;
}

// out[dst[i]] = sum_{e in row i} w[e] * v[src[e]]   (rows are distinct flat destinations)
__global__ __launch_bounds__(SDC_BLOCK) void k_scatter(int nrow, const int64_t* __restrict__ dst,
                                                       const int* __restrict__ rowptr,
                                                       const int* __restrict__ src,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ out) {
  const int i = blockIdx.x * SDC_BLOCK + threadIdx.x;
  if (i >= nrow) return;
  double acc = 0.0;
  for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) acc += w[e] * v[src[e]];
  out[dst[i]] = acc;
}
