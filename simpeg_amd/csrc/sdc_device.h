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
  int remap;  // work-group order (block_of): 0 plain, 1 XCD-contiguous, B > 1 also column bands
  // Operator sets (several operators on one mesh, one per column group): column c uses the
  // arrays at offset cset[c] * sstride.  cset null: one operator.
  const int32_t* cset;
  uint32_t sstride;
};

// The operator of column c (every column of a work-item's group shares one set).
template <typename T>
__device__ __forceinline__ Sten<T> sten_col(Sten<T> s, int c) {
  if (s.cset) {
    const size_t o = (size_t)s.cset[c] * s.sstride;
    s.cx += o;
    s.cy += o;
    s.cz += o;
    s.dg += o;
    s.idg += o;
  }
  return s;
}

// Per-node stencil data.  The neighbour indices of a missing neighbour (domain edge) are the
// node itself and its weight is 0, so stencil loads need no branches and stay in bounds.
template <typename T>
struct NodeC {
  T wxm, wxp, wym, wyp, wzm, wzp, d;
  uint32_t ixm, ixp, iym, iyp, izm, izp;
};

template <typename T>
__device__ __forceinline__ NodeC<T> node_coef(const Sten<T>& s, uint32_t n) {
  NodeC<T> c;
  const uint32_t r = fdiv(n, s.fx);
  const uint32_t i = n - r * s.n1x;
  const uint32_t k = fdiv(r, s.fy);
  const uint32_t j = r - k * s.n1y;
  const bool hxm = i > 0, hxp = i + 1 < s.n1x;
  const bool hym = j > 0, hyp = j + 1 < s.n1y;
  const bool hzm = k > 0, hzp = k + 1 < s.n1z;
  c.ixm = hxm ? n - 1 : n;
  c.ixp = hxp ? n + 1 : n;
  c.iym = hym ? n - s.n1x : n;
  c.iyp = hyp ? n + s.n1x : n;
  c.izm = hzm ? n - s.plane : n;
  c.izp = hzp ? n + s.plane : n;
  const T ax = s.cx[n], ay = s.cy[n], az = s.cz[n];
  const T bx = s.cx[c.ixm], by = s.cy[c.iym], bz = s.cz[c.izm];
  // Missing and pinned edges are removed by multiplying with 0/1 masks (the clamped loads read
  // finite coefficients), not by a branch or a select: with a branch, or a select the compiler
  // turned into a conditional load, the coefficient loads were waited for before the operand
  // loads were issued, one extra memory round trip per wavefront.
  const bool z0 = s.pin0 != 0 && n == 0;
  const bool p1 = s.pin0 == 1;
  c.wxp = ax * (T)(hxp && !z0);
  c.wyp = ay * (T)(hyp && !z0);
  c.wzp = az * (T)(hzp && !z0);
  c.wxm = bx * (T)(hxm && !(p1 && n == 1));
  c.wym = by * (T)(hym && !(p1 && n == s.n1x));
  c.wzm = bz * (T)(hzm && !(p1 && n == s.plane));
  c.d = s.dg[n];
  return c;
}

// Flags and counters written by an earlier kernel of the same stream (convergence flags, the
// PCG control block) are read from L2 with device-scope loads, never through a compute unit's
// scalar cache, which may still hold the line from before that kernel wrote it.
__device__ __forceinline__ int ld_flag(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_flag(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
__device__ __forceinline__ T stencil(const NodeC<T>& c, const T* __restrict__ x, uint32_t n) {
  const T xc = x[n], xw = x[c.ixm], xe = x[c.ixp], xs = x[c.iym], xn = x[c.iyp], xb = x[c.izm],
          xt = x[c.izp];
  return c.d * xc - c.wxm * xw - c.wxp * xe - c.wym * xs - c.wyp * xn - c.wzm * xb - c.wzp * xt;
}

// Column c0 + r of a block, clamped to the last column: loads of columns past ncol read valid
// memory and their results are discarded, so no load needs a branch.
template <typename T>
__device__ __forceinline__ const T* col_ptr(const T* x, size_t ld, int c0, int r, int ncol) {
  const int c = c0 + r < ncol ? c0 + r : ncol - 1;
  return x + (size_t)c * ld;
}

// (A x)_n for R columns starting at column c0.  All 7R loads are issued before the arithmetic
// so that a wavefront keeps them in flight; the per-lane indices are shared by every column.
// The block may be stored in a narrower type X (bf16) than the arithmetic type T.
template <typename T, int R, typename X>
__device__ __forceinline__ void stencil_cols(const NodeC<T>& c, const X* __restrict__ x, uint32_t n,
                                             size_t ld, int c0, int ncol, T (&xc)[R], T (&v)[R]) {
  T xw[R], xe[R], xs[R], xn[R], xb[R], xt[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const X* col = col_ptr(x, ld, c0, r, ncol);
    xc[r] = (T)col[n];
    xw[r] = (T)col[c.ixm];
    xe[r] = (T)col[c.ixp];
    xs[r] = (T)col[c.iym];
    xn[r] = (T)col[c.iyp];
    xb[r] = (T)col[c.izm];
    xt[r] = (T)col[c.izp];
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    v[r] = c.d * xc[r] - c.wxm * xw[r] - c.wxp * xe[r] - c.wym * xs[r] - c.wyp * xn[r] -
           c.wzm * xb[r] - c.wzp * xt[r];
}

// Sum R per-thread doubles over the work-group and write one partial per
// column: partial[(c0 + r) * nbx + bx].
// Inside a wavefront the R sums share one butterfly: at each of the first log2(R) xor steps a
// lane keeps half of its values and trades the other half with its partner, so R columns cost
// R - 1 + 6 - log2(R) exchanges instead of 6R.  Afterwards every lane holds the wave sum of
// column (lane >> (6 - log2 R)).
template <int R>
__device__ __forceinline__ void block_partials(double (&acc)[R], double* __restrict__ partial,
                                               int c0, int ncol, uint32_t nbx, uint32_t bx) {
  static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 16 || R == 32, "R: power of 2 <= 32");
  constexpr int LOG_R = R == 32 ? 5 : R == 16 ? 4 : R == 8 ? 3 : R == 4 ? 2 : R == 2 ? 1 : 0;
  __shared__ double red[R][SDC_BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double a[R];
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] = acc[r];
  int mask = 32;
#pragma unroll
  for (int w = R; w > 1; w >>= 1) {
    const int half = w >> 1;
    const bool hi = (lane & mask) != 0;
#pragma unroll
    for (int k = 0; k < half; ++k) {
      const double mine = hi ? a[half + k] : a[k];
      const double other = hi ? a[k] : a[half + k];
      a[k] = mine + __shfl_xor(other, mask, 64);
    }
    mask >>= 1;
  }
#pragma unroll
  for (int m = 32 >> LOG_R; m > 0; m >>= 1) a[0] += __shfl_xor(a[0], m, 64);
  constexpr int GROUP = 64 / R;
  if ((lane & (GROUP - 1)) == 0) red[lane / GROUP][wid] = a[0];
  __syncthreads();
  if (threadIdx.x < R) {
    const int r = threadIdx.x;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < SDC_BLOCK / 64; ++w) v += red[r][w];
    if (c0 + r < ncol) partial[(size_t)(c0 + r) * nbx + bx] = v;
  }
}
