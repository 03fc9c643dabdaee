// Device-side building blocks shared by the nodal DC kernels (gfx950 / CDNA4).
//
// Layout: a batch of node vectors is column-major, column c of a level with
// column stride `ld` starts at base + c*ld; node n = i + n1x*(j + n1y*k)
// (x fastest, the discretize node ordering).  One work-item owns one node
// and R columns, so every load of a 64-wide wavefront is 64 consecutive nodes
// of one column (256 B for fp32, 512 B for fp64) and the per-node stencil
// coefficients are loaded once and reused for R columns.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SDC_BLOCK 256

// Unsigned division by a run-time constant: q = (umulhi(n, m) + n) >> s,
// valid for n < 2^31 (Granlund & Montgomery, "Division by invariant integers
// using multiplication", PLDI 1994, round-up variant).
struct FastDiv {
  uint32_t d, m, s;
};

static inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  f.s = l;
  f.m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
  return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
  return (__umulhi(n, f.m) + n) >> f.s;
}

// Work-groups are handed to the 8 XCDs round-robin by linear id.  Remap so
// that XCD x receives one contiguous run of logical blocks: neighbouring
// x/y/z rows of a sweep then share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t per = nb >> 3, rem = nb & 7u;
  const uint32_t x = b & 7u, l = b >> 3;
  return x * per + (x < rem ? x : rem) + l;
}

// Coefficients of the 7-point nodal operator on one grid level.
//   cx[n]: conductance of the +x edge of node n (0 on the last x node), same for cy, cz
//   dg[n]: full diagonal sum(edge conductances) + Robin term
//   idg[n]: 1/dg[n]
// pin0 = 1: node 0 is an identity row and column (Neumann gauge, symmetric form)
// pin0 = 2: node 0 is an identity row only (the nonsymmetric reference operator)
template <typename T>
struct Sten {
  const T* cx;
  const T* cy;
  const T* cz;
  const T* dg;
  const T* idg;
  uint32_t n1x, n1y, n1z, nN, plane;
  FastDiv fx, fy;
  int pin0;
};

template <typename T>
struct NodeC {
  T wxm, wxp, wym, wyp, wzm, wzp, d;
  bool hxm, hxp, hym, hyp, hzm, hzp;
};

template <typename T>
__device__ __forceinline__ NodeC<T> node_coef(const Sten<T>& s, uint32_t n) {
  NodeC<T> c;
  const uint32_t r = fdiv(n, s.fx);
  const uint32_t i = n - r * s.n1x;
  const uint32_t k = fdiv(r, s.fy);
  const uint32_t j = r - k * s.n1y;
  c.hxm = i > 0;
  c.hxp = i + 1 < s.n1x;
  c.hym = j > 0;
  c.hyp = j + 1 < s.n1y;
  c.hzm = k > 0;
  c.hzp = k + 1 < s.n1z;
  c.wxp = s.cx[n];
  c.wyp = s.cy[n];
  c.wzp = s.cz[n];
  c.wxm = c.hxm ? s.cx[n - 1] : T(0);
  c.wym = c.hym ? s.cy[n - s.n1x] : T(0);
  c.wzm = c.hzm ? s.cz[n - s.plane] : T(0);
  c.d = s.dg[n];
  if (s.pin0) {
    if (n == 0) {
      c.wxp = c.wyp = c.wzp = T(0);
    } else if (s.pin0 == 1) {
      if (n == 1) c.wxm = T(0);
      if (n == s.n1x) c.wym = T(0);
      if (n == s.plane) c.wzm = T(0);
    }
  }
  return c;
}

// Loads/stores of operands that a kernel touches exactly once (right-hand sides, search
// directions, outputs) carry the non-temporal hint so that L2 keeps the stencil operand's
// neighbouring planes.
template <typename T>
__device__ __forceinline__ T ld_once(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void st_once(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// (A x)_n for one column (x points at the column base).
template <typename T>
__device__ __forceinline__ T stencil(const NodeC<T>& c, const T* __restrict__ x, uint32_t n,
                                     uint32_t n1x, uint32_t plane) {
  T v = c.d * x[n];
  if (c.hxm) v -= c.wxm * x[n - 1];
  if (c.hxp) v -= c.wxp * x[n + 1];
  if (c.hym) v -= c.wym * x[n - n1x];
  if (c.hyp) v -= c.wyp * x[n + n1x];
  if (c.hzm) v -= c.wzm * x[n - plane];
  if (c.hzp) v -= c.wzp * x[n + plane];
  return v;
}

// (A x)_n for R columns starting at column c0 (columns >= ncol are skipped).  All loads are
// issued before the arithmetic so that a wavefront keeps 7R requests in flight.
template <typename T, int R>
__device__ __forceinline__ void stencil_cols(const NodeC<T>& c, const T* __restrict__ x, uint32_t n,
                                             size_t ld, int c0, int ncol, uint32_t n1x,
                                             uint32_t plane, T (&xc)[R], T (&v)[R]) {
  T xw[R], xe[R], xs[R], xn[R], xb[R], xt[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool ok = c0 + r < ncol;
    const T* col = x + (size_t)(c0 + r) * ld;
    xc[r] = ok ? col[n] : T(0);
    xw[r] = (ok && c.hxm) ? col[n - 1] : T(0);
    xe[r] = (ok && c.hxp) ? col[n + 1] : T(0);
    xs[r] = (ok && c.hym) ? col[n - n1x] : T(0);
    xn[r] = (ok && c.hyp) ? col[n + n1x] : T(0);
    xb[r] = (ok && c.hzm) ? col[n - plane] : T(0);
    xt[r] = (ok && c.hzp) ? col[n + plane] : T(0);
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    v[r] = c.d * xc[r] - c.wxm * xw[r] - c.wxp * xe[r] - c.wym * xs[r] - c.wyp * xn[r] -
           c.wzm * xb[r] - c.wzp * xt[r];
}

// Sum R per-thread doubles over the work-group and write one partial per
// column: partial[(c0 + r) * nbx + bx].
template <int R>
__device__ __forceinline__ void block_partials(double (&acc)[R], double* __restrict__ partial,
                                               int c0, int ncol, uint32_t nbx, uint32_t bx) {
  __shared__ double red[R][SDC_BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double v = acc[r];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[r][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < R) {
    const int r = threadIdx.x;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < SDC_BLOCK / 64; ++w) v += red[r][w];
    if (c0 + r < ncol) partial[(size_t)(c0 + r) * nbx + bx] = v;
  }
}
