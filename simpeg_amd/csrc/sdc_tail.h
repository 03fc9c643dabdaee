// The coarse end of the V-cycle in one launch.
//
// Levels l_t .. L-1 of the hierarchy are small (7,406 + 1,352 + 245 + 48 nodes below a
// 932,137-node mesh), so the separate launches for their smoothing steps, residuals and
// transfers are bound by launch latency, not by memory.  Here one work-group owns one
// right-hand-side column and runs the whole V-cycle below level l_t - 1 out of LDS (gfx950: up
// to 160 KB per work-group): the column's rhs of level l_t is read once from HBM, the solution
// written once.  The arithmetic is the multi-launch V-cycle's: Chebyshev steps on the same
// interval (three-term form), P^T as the 27-point gather, the same P (geometric or
// operator-dependent weights), and the dense inverse on the coarsest level.  Coefficients and
// transfer tables are read from global memory (they are shared by all columns and stay in L2).
#pragma once
#include "sdc_kernels.h"

#define SDC_TAIL_MAX 6

struct TailLevel {
  Sten<float> s;    // operator of this level (pin0 as in the multi-launch cycle)
  Xfer<float> t;    // transfer to the next coarser level (unused on the coarsest)
  uint32_t offB;    // LDS offsets (floats) of this level's rhs and solution
  uint32_t offX;
};

struct TailDesc {
  TailLevel lv[SDC_TAIL_MAX];
  int nlev;          // levels in the tail; the last one is the coarsest (dense inverse)
  uint32_t offY;     // LDS offset of the shared scratch vector (size of the largest level)
  const float* ainv; // coarsest dense inverse, nn x nn, symmetric
  int nn;
  const int32_t* cset;  // operator sets: column c uses set cset[c] (null: one set); the levels'
                        // Sten / Xfer carry their own strides, the dense inverse astride
  uint32_t astride;
  int nu_pre, nu_post;
  float c1[16], c2[16];
};

namespace tail {

// (A x)_n on an LDS vector
__device__ __forceinline__ float apply_at(const Sten<float>& s, const float* x, uint32_t n,
                                          float& idg) {
  const NodeC<float> c = node_coef(s, n);
  idg = s.idg[n];
  return c.d * x[n] - c.wxm * x[c.ixm] - c.wxp * x[c.ixp] - c.wym * x[c.iym] -
         c.wyp * x[c.iyp] - c.wzm * x[c.izm] - c.wzp * x[c.izp];
}

// Chebyshev smoothing on one level.  b, x: this level's rhs and iterate; y: scratch.  On entry
// x holds the initial guess unless zero_guess; on exit the result is in x.
__device__ void smooth(const TailDesc& d, const Sten<float>& s, const float* b, float* x, float* y,
                       int nsteps, bool zero_guess) {
  const uint32_t nN = s.nN;
  float* cur = x;
  float* prev = y;
  int k0 = 0;
  if (zero_guess) {
    for (uint32_t n = threadIdx.x; n < nN; n += blockDim.x) cur[n] = d.c2[0] * s.idg[n] * b[n];
    __syncthreads();
    k0 = 1;
  }
  for (int k = k0; k < nsteps; ++k) {
    // new iterate into prev: x+ = x + c1 (x - x-) + c2 D^-1 (b - A x); x- = 0 before step 1 of
    // a zero-guess sweep, absent (c1 = 0) on the first step from a non-zero guess
    const bool has_prev = zero_guess ? (k >= 2) : (k >= 1);
    const float c1 = (k == 0) ? 0.0f : d.c1[k];
    for (uint32_t n = threadIdx.x; n < nN; n += blockDim.x) {
      float g;
      const float ax = apply_at(s, cur, n, g);
      const float xc = cur[n];
      const float xm = has_prev ? prev[n] : 0.0f;
      prev[n] = xc + c1 * (xc - xm) + d.c2[k] * g * (b[n] - ax);
    }
    __syncthreads();
    float* t = cur;
    cur = prev;
    prev = t;
  }
  if (cur != x) {
    for (uint32_t n = threadIdx.x; n < nN; n += blockDim.x) x[n] = cur[n];
    __syncthreads();
  }
}

// y = b - A x
__device__ void residual(const Sten<float>& s, const float* b, const float* x, float* y) {
  for (uint32_t n = threadIdx.x; n < s.nN; n += blockDim.x) {
    float g;
    y[n] = b[n] - apply_at(s, x, n, g);
  }
  __syncthreads();
}

// bc = P^T r (27-point gather, clamped indices, zero weights where a neighbour is missing)
__device__ void restrict_to(const Xfer<float>& t, const float* r, float* bc) {
  const int fn1z = (int)(t.fN / t.fplane);
  for (uint32_t n = threadIdx.x; n < t.cN; n += blockDim.x) {
    const uint32_t rr = fdiv(n, t.cfx);
    const uint32_t I = n - rr * t.cn1x;
    const uint32_t K = fdiv(rr, t.cfy);
    const uint32_t J = rr - K * t.cn1y;
    const int fi = t.rix[I], fj = t.riy[J], fk = t.riz[K];
    const int xi[3] = {fi > 0 ? fi - 1 : 0, fi, fi + 1 < (int)t.fn1x ? fi + 1 : fi};
    const int yj[3] = {fj > 0 ? fj - 1 : 0, fj, fj + 1 < (int)t.fn1y ? fj + 1 : fj};
    const int zk[3] = {fk > 0 ? fk - 1 : 0, fk, fk + 1 < fn1z ? fk + 1 : fk};
    float w[27];
    restrict_weights(t, I, J, K, xi, yj, zk, w);
    float acc = 0.0f;
    for (int c = 0; c < 3; ++c)
      for (int bb = 0; bb < 3; ++bb) {
        const uint32_t row = (uint32_t)yj[bb] * t.fn1x + (uint32_t)zk[c] * t.fplane;
        const float* wb = w + 3 * bb + 9 * c;
        acc += wb[0] * r[row + xi[0]] + wb[1] * r[row + xi[1]] + wb[2] * r[row + xi[2]];
      }
    bc[n] = acc;
  }
  __syncthreads();
}

// xf += P xc (prolong_weights)
__device__ void prolong_add(const Xfer<float>& t, const float* xc, float* xf) {
  for (uint32_t n = threadIdx.x; n < t.fN; n += blockDim.x) {
    const uint32_t rr = fdiv(n, t.ffx);
    const uint32_t i = n - rr * t.fn1x;
    const uint32_t k = fdiv(rr, t.ffy);
    const uint32_t j = rr - k * t.fn1y;
    const uint32_t I0 = (uint32_t)t.pix[i], J0 = (uint32_t)t.piy[j];
    float w[8];
    prolong_weights(t, n, i, j, k, I0, J0, w);
    const uint32_t c000 = I0 + J0 * t.cn1x + (uint32_t)t.piz[k] * t.cplane;
    const uint32_t c010 = c000 + t.cn1x, c001 = c000 + t.cplane, c011 = c001 + t.cn1x;
    xf[n] += w[0] * xc[c000] + w[1] * xc[c000 + 1] + w[2] * xc[c010] + w[3] * xc[c010 + 1] +
             w[4] * xc[c001] + w[5] * xc[c001 + 1] + w[6] * xc[c011] + w[7] * xc[c011 + 1];
  }
  __syncthreads();
}

}  // namespace tail

// One V-cycle of the tail levels per column: x = M_tail^-1 b, b and x are column c = blockIdx.x of
// the level-l_t blocks (storage type S, column stride ld).
template <typename S>
__global__ __launch_bounds__(1024) void k_tail_vcycle(TailDesc d, const S* __restrict__ b,
                                                      S* __restrict__ x, size_t ld,
                                                      const int* __restrict__ done) {
  if (done && ld_flag(done)) return;
  extern __shared__ float lds[];
  const size_t col = blockIdx.x;
  float* y = lds + d.offY;
  // the operator set of this column: each level's coefficients and weights are selected into
  // local copies where they are used (the descriptor itself is not modified: a written kernel
  // argument of this size is copied to scratch memory), and the dense inverse here
  const float* ainv = d.ainv + (d.cset ? (size_t)d.cset[col] * d.astride : (size_t)0);
  {
    const TailLevel& L0 = d.lv[0];
    float* B = lds + L0.offB;
    const S* src = b + col * ld;
    for (uint32_t n = threadIdx.x; n < L0.s.nN; n += blockDim.x) B[n] = (float)src[n];
    __syncthreads();
  }
  const int last = d.nlev - 1;
  // down: pre-smoothing, residual, restriction
  for (int l = 0; l < last; ++l) {
    const TailLevel& L = d.lv[l];
    const Sten<float> s = sten_col(L.s, (int)col);
    float* B = lds + L.offB;
    float* X = lds + L.offX;
    tail::smooth(d, s, B, X, y, d.nu_pre, true);
    tail::residual(s, B, X, y);
    tail::restrict_to(xfer_col(L.t, (int)col), y, lds + d.lv[l + 1].offB);
  }
  // coarsest: dense inverse (symmetric: column m read as row m, consecutive threads consecutive)
  {
    const TailLevel& L = d.lv[last];
    const float* B = lds + L.offB;
    float* X = lds + L.offX;
    for (int r = threadIdx.x; r < d.nn; r += blockDim.x) {
      float acc = 0.0f;
      for (int m = 0; m < d.nn; ++m) acc += ainv[(size_t)m * d.nn + r] * B[m];
      X[r] = acc;
    }
    __syncthreads();
  }
  // up: interpolation, post-smoothing
  for (int l = last - 1; l >= 0; --l) {
    const TailLevel& L = d.lv[l];
    float* B = lds + L.offB;
    float* X = lds + L.offX;
    tail::prolong_add(xfer_col(L.t, (int)col), lds + d.lv[l + 1].offX, X);
    tail::smooth(d, sten_col(L.s, (int)col), B, X, y, d.nu_post, false);
  }
  {
    const TailLevel& L0 = d.lv[0];
    const float* X = lds + L0.offX;
    S* dst = x + col * ld;
    for (uint32_t n = threadIdx.x; n < L0.s.nN; n += blockDim.x) dst[n] = (S)X[n];
  }
}
