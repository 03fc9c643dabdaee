// Host side of libsimpeg_dc.so: handle, multigrid hierarchy, PCG / V-cycle / iterative
// refinement drivers and the C ABI declared in include/simpeg_dc.h.
#include "sdc_kernels.h"
#include "sdc_tail.h"

// 1: the residual and the x pass of the separable restriction run as one kernel
#ifndef SDC_FUSE_RX
#define SDC_FUSE_RX 1
#endif

// k_apply_pq: column groups per band whose work-groups for one node block run back to back
// (its coefficients fetched once per band; measured on the 932k-node / 128-column solve: 461 ->
// 438 us per launch with 4, while the smoothing kernels, which read b at 7 points, gained
// nothing and keep the plain XCD order)
constexpr int kApplyBand = 4;

// search directions held by the mixed PCG before they are folded into the solution (a full
// ring forces a fold)
constexpr int kRingSlots = 6;
#include "simpeg_dc.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

// Device scratch of the allocation-free adjoint contraction (sdc_jtvec_contract_noalloc): the
// per-column-group partial sums, SDC_SCRATCH_DOUBLES doubles (32 MiB).  Every GPU the library
// runs on holds its own copy until the process ends.  Host code never passes this name to a
// kernel (on the host it is only a stand-in for the symbol): it uses the per-device address
// that hipGetSymbolAddress returned at sdc_create.
#define SDC_SCRATCH_DOUBLES (4u << 20)
static __device__ __align__(256) double scratch[SDC_SCRATCH_DOUBLES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCK(x)                                                                     \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      return fail(SDC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define CK(x)                 \
  do {                        \
    int rc_ = (x);            \
    if (rc_ != SDC_OK) return rc_; \
  } while (0)

inline uint32_t nblk(size_t n) { return (uint32_t)((n + SDC_BLOCK - 1) / SDC_BLOCK); }

template <typename T>
int dmalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(SDC_ERR_ALLOC, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  return SDC_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// one tensor axis of a level: nodes 0..n (n cells)
struct Axis {
  int n = 0;
  std::vector<double> X;  // node coordinates (cumulative widths)
};

// Coarsen an axis toward isotropic cells: walking from the first node, two neighbouring
// cells merge into one coarse cell when neither is wider than `target`; otherwise the cell is
// kept.  On a stretched (padded) tensor mesh the small core cells coarsen first and the wide
// padding cells only once the target width has grown to theirs, so coarse cells stay close to
// isotropic and the point smoother keeps working.  Every coarse cell spans 1 or 2 fine cells.
void coarsen_axis(const Axis& f, Axis& c, std::vector<int>& fidx, double target) {
  fidx.assign(1, 0);
  int i = 0;
  while (i < f.n) {
    const double w0 = f.X[i + 1] - f.X[i];
    const double w1 = (i + 1 < f.n) ? f.X[i + 2] - f.X[i + 1] : 0.0;
    if (i + 1 < f.n && w0 <= target && w1 <= target)
      i += 2;
    else
      i += 1;
    fidx.push_back(i);
  }
  c.n = (int)fidx.size() - 1;
  c.X.resize(fidx.size());
  for (size_t I = 0; I < fidx.size(); ++I) c.X[I] = f.X[fidx[I]];
}

template <typename T>
struct XferDev {
  int *rix = nullptr, *riy = nullptr, *riz = nullptr;
  T *rlx = nullptr, *rly = nullptr, *rlz = nullptr, *rhx = nullptr, *rhy = nullptr, *rhz = nullptr;
  int *pix = nullptr, *piy = nullptr, *piz = nullptr;
  T *pwx = nullptr, *pwy = nullptr, *pwz = nullptr;
  // operator-dependent interpolation weights (Xfer::wx), recomputed with the conductances;
  // one array per operator set, wxs / wys / wzs entries apart
  T *wx = nullptr, *wy = nullptr, *wz = nullptr;
  size_t wxs = 0, wys = 0, wzs = 0;
  void release() {
    dfree(rix); dfree(riy); dfree(riz);
    dfree(rlx); dfree(rly); dfree(rlz); dfree(rhx); dfree(rhy); dfree(rhz);
    dfree(pix); dfree(piy); dfree(piz);
    dfree(pwx); dfree(pwy); dfree(pwz);
    dfree(wx); dfree(wy); dfree(wz);
  }
};

struct Level {
  Axis ax[3];
  uint32_t n1x = 0, n1y = 0, n1z = 0, nN = 0, plane = 0;
  size_t ld = 0;
  // coefficient arrays hold one operator per set, sstride entries apart (round_ld(nN))
  size_t sstride = 0;
  double *cx64 = nullptr, *cy64 = nullptr, *cz64 = nullptr, *bd64 = nullptr, *dg64 = nullptr,
         *idg64 = nullptr;
  float *cx32 = nullptr, *cy32 = nullptr, *cz32 = nullptr, *dg32 = nullptr, *idg32 = nullptr;
  // transfer from this level to the next coarser one
  bool has_coarse = false;
  XferDev<double> x64;
  XferDev<float> x32;
  void release() {
    dfree(cx64); dfree(cy64); dfree(cz64); dfree(bd64); dfree(dg64); dfree(idg64);
    dfree(cx32); dfree(cy32); dfree(cz32); dfree(dg32); dfree(idg32);
    x64.release();
    x32.release();
  }
};

// doubles taken by the int part of a control block
constexpr int kCtlInts = SDC_CTL_N * (int)sizeof(int) / (int)sizeof(double);

// dynamic LDS of the coarsest inverse when it runs in LDS (160 KB: all a gfx950 work-group
// may use)
constexpr size_t kCoarseInvLds = 160 * 1024;

template <typename T>
struct Work {
  int ncap = 0;
  // level 0: PCG vectors and smoother scratch
  T *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  // per level V-cycle vectors in the V-cycle storage type (vbytes per entry: T, or bf16 for
  // the fp32 solver): rhs b and solution xs (levels >= 1; level 0 reads r and writes z),
  // smoother ping-pong xa (and xb0 on level 0), residual res
  std::vector<void*> b, xs, xa, res;
  void* xb0 = nullptr;
  int vbytes = 0;
  double* partial = nullptr;
  double *alpha = nullptr, *beta = nullptr, *rz = nullptr, *rr = nullptr, *bn2 = nullptr,
         *aux = nullptr;
  int* active = nullptr;
  int* counters = nullptr;  // [0] active count, [1] ticket, [2] last active count, [3] done,
                            // [4] iterations
  // refinement (outer) state: |q|^2 in aux, |q - A u|^2 in orr, open columns in omask
  double* orr = nullptr;
  int* omask = nullptr;
  int* ocount = nullptr;
  // level 0 nodes where some right-hand-side column is non-zero (fp64 residual skips q elsewhere)
  unsigned* qmask = nullptr;
  // bfloat16 copy of r, the finest-level multigrid input (vcycle_bf16 = 3 or 4)
  __bf16* r16 = nullptr;
  // device-driven mixed PCG: control block (SDC_CTL_N ints, then SDC_DCTL_N doubles in the same
  // allocation), per-column true residual at the last update, and the (q, u) pointer table
  double* ctlmem = nullptr;
  int* ctl = nullptr;
  double* dctl = nullptr;
  double* rmax = nullptr;
  const double** qu = nullptr;
  // plain PCG: the solution buffer, read by k_update_xr4 through this device pointer
  T** xtab = nullptr;
  // mixed PCG: ring of nslot search-direction blocks (slot k at pr + k * ld0 * ncap, device
  // pointer table ptab) and their per-column alphas (ahist, nslot x ncap)
  T* pr = nullptr;
  T** ptab = nullptr;
  double* ahist = nullptr;
  int nslot = 0;
  void release() {
    dfree(x); dfree(r); dfree(z); dfree(p); dfree(q); dfree(r16);
    dfree(ctlmem); dfree(rmax); dfree(qu); dfree(xtab);
    dfree(pr); dfree(ptab); dfree(ahist);
    nslot = 0;
    ctl = nullptr;
    dctl = nullptr;
    for (auto* v : b) dfree(v);
    for (auto* v : xs) dfree(v);
    for (auto* v : xa) dfree(v);
    for (auto* v : res) dfree(v);
    b.clear(); xs.clear(); xa.clear(); res.clear();
    dfree(xb0);
    vbytes = 0;
    dfree(partial); dfree(alpha); dfree(beta); dfree(rz); dfree(rr); dfree(bn2); dfree(aux);
    dfree(active); dfree(counters);
    dfree(orr); dfree(omask); dfree(ocount); dfree(qmask);
    ncap = 0;
  }
};

}  // namespace

struct sdc_handle_s {
  int device = 0;
  int bc = SDC_BC_ROBIN;
  uint32_t nx = 0, ny = 0, nz = 0;
  double *hx = nullptr, *hy = nullptr, *hz = nullptr;
  double *bxm = nullptr, *bxp = nullptr, *bym = nullptr, *byp = nullptr, *bzm = nullptr,
         *bzp = nullptr;
  std::vector<Level> L;
  int ncoarse = 0;
  double* ainv64 = nullptr;
  float* ainv32 = nullptr;
  double *gj_a = nullptr, *gj_f = nullptr;
  size_t astride = 0;  // entries between the operator sets' dense inverses
  // operator sets: capacity of the per-set arrays, sets built by the last sdc_set_sigma_sets
  // (1 after sdc_set_sigma / sdc_set_coefficients), and during sdc_solve_sets the device map
  // column -> set that every coefficient-reading kernel follows (null otherwise)
  int nset_cap = 1;
  int nset = 1;
  int32_t* cset_dev = nullptr;
  int cset_cap = 0;
  const int32_t* cur_cset = nullptr;
  // scratch coefficient arrays for A(dsigma) (Jvec right-hand side)
  double *tcx = nullptr, *tcy = nullptr, *tcz = nullptr, *tbd = nullptr, *tdg = nullptr,
         *tidg = nullptr;
  // per-column-group partial sums of the adjoint contraction (jpart_cap doubles)
  double* jpart = nullptr;
  size_t jpart_cap = 0;
  bool sigma_set = false;
  // a mixed solve stalled on the current operator and parameters: later mixed solves run fp64
  bool mixed_stalled = false;
  sdc_params prm;
  int wlevel = 0;  // level whose coarse correction runs twice (0: none)
  int tail_level = 0;    // first level of the single-launch coarse tail (0: none)
  size_t tail_lds = 0;   // its LDS bytes per work-group
  Work<float> w32;
  Work<double> w64;
  int* h_flag = nullptr;  // pinned host copy of the convergence flag
  double* h_ctl = nullptr;  // pinned host copy of the mixed PCG control block
  int* d_ctl = nullptr;     // its device address: the mixed PCG writes it when it stops
  // recorded finest-level V-cycles (sdc_params.graph): one executable graph per input, output,
  // column count and storage variant, launched on the caller's stream
  struct GraphRec {
    const void* b;
    void* x;
    int ncol;
    int variant;
    hipGraphExec_t exec;
  };
  std::vector<GraphRec> graphs;
  // while a V-cycle is being recorded: the graph under construction and its last node
  struct Recorder {
    hipGraph_t g = nullptr;
    hipGraphNode_t last = nullptr;
    hipError_t err = hipSuccess;
    std::vector<std::shared_ptr<void>> args;  // argument values, alive until instantiation
  }* rec = nullptr;
  // event profiler for one kernel class (sdc_profile)
  uint32_t prof_mask = 0;
  struct ProfRec {
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    std::vector<double> bytes;
  } prof[SDC_PROF_NTAGS];
};

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Forget the captured V-cycles (their kernel arguments name workspaces and parameters).
void drop_graphs(sdc_handle_s* h) {
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
}

// Every kernel of the V-cycle starts here: a launch on st, or, while h->rec is set, a kernel
// node appended to the recorded graph after the previous one (the same order as on a stream).
// The graph is built node by node rather than by stream capture, so recording does not interact
// with the null-stream operations other host threads may issue meanwhile.
template <typename... P, typename... A>
void launch(sdc_handle_s* h, void (*k)(P...), dim3 grid, dim3 block, size_t shmem,
            hipStream_t st, A&&... a) {
  auto* rec = h->rec;
  if (!rec) {
    hipLaunchKernelGGL(k, grid, block, shmem, st, std::forward<A>(a)...);
    return;
  }
  auto vals = std::make_shared<std::tuple<std::decay_t<P>...>>(static_cast<std::decay_t<P>>(a)...);
  rec->args.push_back(vals);
  void* ptrs[sizeof...(P) > 0 ? sizeof...(P) : 1];
  std::apply(
      [&](auto&... v) {
        size_t i = 0;
        ((ptrs[i++] = (void*)&v), ...);
      },
      *vals);
  hipKernelNodeParams prm = {};
  prm.func = (void*)k;
  prm.gridDim = grid;
  prm.blockDim = block;
  prm.sharedMemBytes = (unsigned)shmem;
  prm.kernelParams = ptrs;
  prm.extra = nullptr;
  hipGraphNode_t node = nullptr;
  const hipError_t e = hipGraphAddKernelNode(&node, rec->g, rec->last ? &rec->last : nullptr,
                                             rec->last ? 1 : 0, &prm);
  if (e != hipSuccess) {
    if (rec->err == hipSuccess) rec->err = e;
    return;
  }
  rec->last = node;
}

// Brackets one launch of a kernel class selected by sdc_profile with events on its stream.
bool prof_on(const sdc_handle_s* h, int tag) {
  return tag > 0 && tag < SDC_PROF_NTAGS && (h->prof_mask & (1u << tag));
}

void prof_begin(sdc_handle_s* h, int tag, hipStream_t st) {
  if (!prof_on(h, tag)) return;
  auto& p = h->prof[tag];
  while (p.ev.size() < p.used + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    p.ev.push_back(e);
  }
  (void)hipEventRecord(p.ev[p.used], st);
}

void prof_end(sdc_handle_s* h, int tag, hipStream_t st, double bytes) {
  if (!prof_on(h, tag)) return;
  auto& p = h->prof[tag];
  if (p.ev.size() < p.used + 2) return;
  (void)hipEventRecord(p.ev[p.used + 1], st);
  p.used += 2;
  p.bytes.push_back(bytes);
}

template <typename T>
Sten<T> make_sten(const sdc_handle_s* h, int l, int pin_mode);

template <>
Sten<double> make_sten<double>(const sdc_handle_s* h, int l, int pin_mode) {
  const Level& v = h->L[l];
  Sten<double> s{};
  s.cx = v.cx64; s.cy = v.cy64; s.cz = v.cz64; s.dg = v.dg64; s.idg = v.idg64;
  s.n1x = v.n1x; s.n1y = v.n1y; s.n1z = v.n1z; s.nN = v.nN; s.plane = v.plane;
  s.fx = make_fastdiv(v.n1x); s.fy = make_fastdiv(v.n1y);
  s.pin0 = (h->bc == SDC_BC_NEUMANN) ? pin_mode : 0;
  s.remap = h->prm.xcd_remap != 0;
  s.cset = h->cur_cset;
  s.sstride = (uint32_t)v.sstride;
  return s;
}

template <>
Sten<float> make_sten<float>(const sdc_handle_s* h, int l, int pin_mode) {
  const Level& v = h->L[l];
  Sten<float> s{};
  s.cx = v.cx32; s.cy = v.cy32; s.cz = v.cz32; s.dg = v.dg32; s.idg = v.idg32;
  s.n1x = v.n1x; s.n1y = v.n1y; s.n1z = v.n1z; s.nN = v.nN; s.plane = v.plane;
  s.fx = make_fastdiv(v.n1x); s.fy = make_fastdiv(v.n1y);
  s.pin0 = (h->bc == SDC_BC_NEUMANN) ? pin_mode : 0;
  s.remap = h->prm.xcd_remap != 0;
  s.cset = h->cur_cset;
  s.sstride = (uint32_t)v.sstride;
  return s;
}

template <typename T>
const XferDev<T>& xfer_dev(const Level& v);
template <>
const XferDev<double>& xfer_dev<double>(const Level& v) { return v.x64; }
template <>
const XferDev<float>& xfer_dev<float>(const Level& v) { return v.x32; }

template <typename T>
Xfer<T> make_xfer(const sdc_handle_s* h, int l) {
  const Level& f = h->L[l];
  const Level& c = h->L[l + 1];
  const XferDev<T>& d = xfer_dev<T>(f);
  Xfer<T> t{};
  t.rix = d.rix; t.riy = d.riy; t.riz = d.riz;
  t.rlx = d.rlx; t.rly = d.rly; t.rlz = d.rlz;
  t.rhx = d.rhx; t.rhy = d.rhy; t.rhz = d.rhz;
  t.pix = d.pix; t.piy = d.piy; t.piz = d.piz;
  t.pwx = d.pwx; t.pwy = d.pwy; t.pwz = d.pwz;
  t.fn1x = f.n1x; t.fn1y = f.n1y; t.fplane = f.plane; t.fN = f.nN;
  t.cn1x = c.n1x; t.cn1y = c.n1y; t.cplane = c.plane; t.cN = c.nN;
  t.ffx = make_fastdiv(f.n1x); t.ffy = make_fastdiv(f.n1y);
  t.cfx = make_fastdiv(c.n1x); t.cfy = make_fastdiv(c.n1y);
  t.remap = h->prm.xcd_remap != 0;
  t.wx = h->prm.interp ? d.wx : nullptr;
  t.wy = h->prm.interp ? d.wy : nullptr;
  t.wz = h->prm.interp ? d.wz : nullptr;
  t.cset = h->cur_cset;
  t.wxs = (uint32_t)d.wxs;
  t.wys = (uint32_t)d.wys;
  t.wzs = (uint32_t)d.wzs;
  return t;
}

// deriv: geometry for a derivative kernel (dA/dsigma products), whose Robin term follows
// sdc_params.deriv_robin
Geo make_geo(const sdc_handle_s* h, bool deriv = false) {
  Geo g{};
  g.hx = h->hx; g.hy = h->hy; g.hz = h->hz;
  g.bxm = h->bxm; g.bxp = h->bxp; g.bym = h->bym; g.byp = h->byp; g.bzm = h->bzm; g.bzp = h->bzp;
  g.robin = h->bc == SDC_BC_ROBIN && (!deriv || h->prm.deriv_robin != 0);
  g.nx = h->nx; g.ny = h->ny; g.nz = h->nz;
  g.n1x = h->nx + 1; g.n1y = h->ny + 1; g.nN = h->L[0].nN; g.plane = g.n1x * g.n1y;
  g.fx = make_fastdiv(g.n1x); g.fy = make_fastdiv(g.n1y);
  g.cfx = make_fastdiv(h->nx); g.cfy = make_fastdiv(h->ny);
  return g;
}

template <typename T>
int upload(T** dst, const std::vector<T>& src) {
  CK(dmalloc(dst, src.size()));
  HIPCK(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return SDC_OK;
}

template <typename T>
int upload_xfer(XferDev<T>& d, const std::vector<int> (&ri)[3], const std::vector<double> (&rl)[3],
                const std::vector<double> (&rh)[3], const std::vector<int> (&pi)[3],
                const std::vector<double> (&pw)[3]) {
  auto cvt = [](const std::vector<double>& a) { return std::vector<T>(a.begin(), a.end()); };
  CK(upload(&d.rix, ri[0])); CK(upload(&d.riy, ri[1])); CK(upload(&d.riz, ri[2]));
  CK(upload(&d.rlx, cvt(rl[0]))); CK(upload(&d.rly, cvt(rl[1]))); CK(upload(&d.rlz, cvt(rl[2])));
  CK(upload(&d.rhx, cvt(rh[0]))); CK(upload(&d.rhy, cvt(rh[1]))); CK(upload(&d.rhz, cvt(rh[2])));
  CK(upload(&d.pix, pi[0])); CK(upload(&d.piy, pi[1])); CK(upload(&d.piz, pi[2]));
  CK(upload(&d.pwx, cvt(pw[0]))); CK(upload(&d.pwy, cvt(pw[1]))); CK(upload(&d.pwz, cvt(pw[2])));
  return SDC_OK;
}

size_t round_ld(size_t n) { return (n + 63) / 64 * 64; }

void fill_level_dims(Level& v) {
  v.n1x = v.ax[0].n + 1;
  v.n1y = v.ax[1].n + 1;
  v.n1z = v.ax[2].n + 1;
  v.plane = v.n1x * v.n1y;
  v.nN = v.plane * v.n1z;
  v.ld = round_ld(v.nN);
}

void free_level_arrays(Level& v) {
  dfree(v.cx64); dfree(v.cy64); dfree(v.cz64); dfree(v.bd64); dfree(v.dg64); dfree(v.idg64);
  dfree(v.cx32); dfree(v.cy32); dfree(v.cz32); dfree(v.dg32); dfree(v.idg32);
}

int alloc_level_arrays(Level& v, int nset) {
  free_level_arrays(v);
  v.sstride = round_ld(v.nN);
  const size_t n = v.sstride * (size_t)nset;
  CK(dmalloc(&v.cx64, n)); CK(dmalloc(&v.cy64, n)); CK(dmalloc(&v.cz64, n));
  CK(dmalloc(&v.bd64, n)); CK(dmalloc(&v.dg64, n)); CK(dmalloc(&v.idg64, n));
  CK(dmalloc(&v.cx32, n)); CK(dmalloc(&v.cy32, n)); CK(dmalloc(&v.cz32, n));
  CK(dmalloc(&v.dg32, n)); CK(dmalloc(&v.idg32, n));
  return SDC_OK;
}

// operator-dependent interpolation weights of the transfer below level f, for nset sets
int alloc_weights(Level& f, int nset) {
  const size_t k = (size_t)nset;
  dfree(f.x64.wx); dfree(f.x64.wy); dfree(f.x64.wz);
  dfree(f.x32.wx); dfree(f.x32.wy); dfree(f.x32.wz);
  CK(dmalloc(&f.x64.wx, f.x64.wxs * k)); CK(dmalloc(&f.x64.wy, f.x64.wys * k));
  CK(dmalloc(&f.x64.wz, f.x64.wzs * k));
  CK(dmalloc(&f.x32.wx, f.x32.wxs * k)); CK(dmalloc(&f.x32.wy, f.x32.wys * k));
  CK(dmalloc(&f.x32.wz, f.x32.wzs * k));
  return SDC_OK;
}

// transfer tables between fine level f and the coarse level c built from it (c.ax must
// already hold the coarse axes, fidx the fine index of every coarse node)
int make_transfer(Level& f, Level& c, std::vector<int> (&fidx)[3]) {
  std::vector<int> ri[3], pi[3];
  std::vector<double> rl[3], rh[3], pw[3];
  for (int a = 0; a < 3; ++a) {
    const int nf = f.ax[a].n, nc = c.ax[a].n;
    pi[a].assign(nf + 1, 0);
    pw[a].assign(nf + 1, 0.0);
    std::vector<char> is_c(nf + 1, 0);
    for (int I = 0; I <= nc; ++I) {
      is_c[fidx[a][I]] = 1;
      pi[a][fidx[a][I]] = std::min(I, std::max(nc - 1, 0));
      pw[a][fidx[a][I]] = (I == nc && nc > 0) ? 1.0 : 0.0;
    }
    for (int I = 0; I < nc; ++I) {
      for (int i = fidx[a][I] + 1; i < fidx[a][I + 1]; ++i) {
        pi[a][i] = I;
        pw[a][i] = (f.ax[a].X[i] - c.ax[a].X[I]) / (c.ax[a].X[I + 1] - c.ax[a].X[I]);
      }
    }
    ri[a].assign(nc + 1, 0);
    rl[a].assign(nc + 1, 0.0);
    rh[a].assign(nc + 1, 0.0);
    for (int I = 0; I <= nc; ++I) {
      const int fi = fidx[a][I];
      ri[a][I] = fi;
      if (I > 0 && fi - 1 > fidx[a][I - 1]) rl[a][I] = pw[a][fi - 1];
      if (I < nc && fi + 1 < fidx[a][I + 1]) rh[a][I] = 1.0 - pw[a][fi + 1];
    }
  }
  fill_level_dims(c);
  CK(upload_xfer(f.x64, ri, rl, rh, pi, pw));
  CK(upload_xfer(f.x32, ri, rl, rh, pi, pw));
  // operator-dependent weights: Px over the fine grid, Py over (coarse x, fine y, fine z),
  // Pz over (coarse x, coarse y, fine z)
  const size_t nwx = f.nN, nwy = (size_t)c.n1x * f.n1y * f.n1z,
               nwz = (size_t)c.n1x * c.n1y * f.n1z;
  f.x64.wxs = f.x32.wxs = round_ld(nwx);
  f.x64.wys = f.x32.wys = round_ld(nwy);
  f.x64.wzs = f.x32.wzs = round_ld(nwz);
  CK(alloc_weights(f, 1));
  f.has_coarse = true;
  return SDC_OK;
}

template <typename T>
Work<T>& work_of(sdc_handle_s* h);
template <>
Work<float>& work_of<float>(sdc_handle_s* h) { return h->w32; }
template <>
Work<double>& work_of<double>(sdc_handle_s* h) { return h->w64; }

// bytes per entry of the V-cycle vectors of the solver in precision T
template <typename T>
int vcycle_bytes(const sdc_handle_s* h) {
  return (std::is_same<T, float>::value && h->prm.vcycle_bf16) ? 2 : (int)sizeof(T);
}

template <typename T>
int ensure_work(sdc_handle_s* h, int ncol) {
  Work<T>& w = work_of<T>(h);
  const int vb = vcycle_bytes<T>(h);
  const bool want16 =
      std::is_same<T, float>::value && (h->prm.vcycle_bf16 == 3 || h->prm.vcycle_bf16 == 4);
  if (w.ncap >= ncol && w.vbytes == vb && (w.r16 != nullptr) == want16) return SDC_OK;
  ncol = std::max(ncol, w.ncap);
  drop_graphs(h);
  w.release();
  const int cap = std::max(ncol, 1);
  const size_t ld0 = h->L[0].ld;
  const size_t n0 = ld0 * (size_t)cap;
  CK(dmalloc(&w.x, n0)); CK(dmalloc(&w.r, n0)); CK(dmalloc(&w.z, n0));
  CK(dmalloc(&w.p, n0)); CK(dmalloc(&w.q, n0));
  const size_t nl = h->L.size();
  w.b.assign(nl, nullptr); w.xs.assign(nl, nullptr); w.xa.assign(nl, nullptr);
  w.res.assign(nl, nullptr);
  w.vbytes = vb;
  auto vmalloc = [](void** p, size_t bytes) -> int {
    unsigned char* q = nullptr;
    CK(dmalloc(&q, bytes));
    *p = q;
    return SDC_OK;
  };
  for (size_t l = 0; l < nl; ++l) {
    const size_t nv = h->L[l].ld * (size_t)cap * (size_t)vb;
    if (l > 0) {
      CK(vmalloc(&w.b[l], nv));
      CK(vmalloc(&w.xs[l], nv));
    }
    CK(vmalloc(&w.xa[l], nv));
    CK(vmalloc(&w.res[l], nv));
    if (l == 0) CK(vmalloc(&w.xb0, nv));
  }
  const size_t nb0 = nblk(h->L[0].nN);
  CK(dmalloc(&w.partial, nb0 * (size_t)cap));
  CK(dmalloc(&w.alpha, cap)); CK(dmalloc(&w.beta, cap)); CK(dmalloc(&w.rz, cap));
  CK(dmalloc(&w.rr, cap)); CK(dmalloc(&w.bn2, cap)); CK(dmalloc(&w.aux, cap));
  CK(dmalloc(&w.active, cap));
  CK(dmalloc(&w.counters, 8));
  HIPCK(hipMemset(w.counters, 0, 8 * sizeof(int)));
  CK(dmalloc(&w.orr, cap));
  CK(dmalloc(&w.omask, cap));
  CK(dmalloc(&w.ocount, 1));
  CK(dmalloc(&w.qmask, ld0));
  if (want16) CK(dmalloc(&w.r16, n0));
  static_assert(SDC_CTL_N * sizeof(int) == kCtlInts * sizeof(double), "control block layout");
  CK(dmalloc(&w.ctlmem, kCtlInts + SDC_DCTL_N));
  w.ctl = (int*)w.ctlmem;
  w.dctl = w.ctlmem + kCtlInts;
  CK(dmalloc(&w.rmax, cap));
  CK(dmalloc(&w.qu, 2));
  CK(dmalloc(&w.xtab, 1));
  if (std::is_same<T, float>::value) {
    w.nslot = kRingSlots;
    CK(dmalloc(&w.pr, n0 * (size_t)w.nslot));
    CK(dmalloc(&w.ahist, (size_t)cap * w.nslot));
    std::vector<T*> tab((size_t)w.nslot);
    for (int k = 0; k < w.nslot; ++k) tab[(size_t)k] = w.pr + (size_t)k * n0;
    CK(upload(&w.ptab, tab));
  }
  w.ncap = cap;
  return SDC_OK;
}

// W-cycle level: the requested one, or (wcycle_level = -1) the first level with at most 1/4 of
// the finest level's nodes (level 1 unless the first coarsening step only touched a small core
// of a padded mesh) -- a second visit there costs ~15 % of a cycle and removes an iteration on
// padded meshes with strong contrasts.  Never the coarsest level.
int pick_wlevel(const sdc_handle_s* h) {
  const int nl = (int)h->L.size();
  int l = h->prm.wcycle_level;
  if (l < 0) {
    l = 0;
    for (int k = 1; k < nl - 1; ++k)
      if ((size_t)h->L[k].nN * 4 <= (size_t)h->L[0].nN) {
        l = k;
        break;
      }
  }
  return (l >= 1 && l < nl - 1) ? l : 0;
}

// First level of the coarse tail that one work-group per column runs out of LDS: the finest
// level below the W-cycle level whose remaining levels fit (rhs + solution per level + one shared
// scratch vector, fp32) into 160 KB, with at least one level above the coarsest.
void pick_tail(sdc_handle_s* h) {
  h->tail_level = 0;
  h->tail_lds = 0;
  const int nl = (int)h->L.size();
  if (!h->prm.fused_tail) return;
  for (int lt = std::max(1, h->wlevel + 1); lt < nl - 1; ++lt) {
    if (nl - lt > SDC_TAIL_MAX) continue;
    size_t words = 0, ymax = 0;
    for (int l = lt; l < nl; ++l) {
      words += 2 * (size_t)h->L[l].nN;
      ymax = std::max(ymax, (size_t)h->L[l].nN);
    }
    const size_t bytes = (words + ymax) * sizeof(float);
    if (bytes <= 160 * 1024) {
      h->tail_level = lt;
      h->tail_lds = bytes;
      return;
    }
  }
}

// Columns per work-item of the stencil kernels: automatic = 4 (measured on MI355X: the
// smoothing steps run 4 % faster than with 8, the Krylov step the same; 2 is slower).
int pick_r(const sdc_handle_s* h, int ncol) {
  int r = h->prm.rhs_per_thread;
  if (r == 1 || r == 2 || r == 4 || r == 8) return r;
  if (ncol >= 4) return 4;
  if (ncol >= 2) return 2;
  return 1;
}

// columns per work-item of the grid-transfer kernels: automatic = 8 (their gathers share one
// index computation across the columns; 4 makes the restriction 15 % slower)
int pick_rx(const sdc_handle_s* h, int ncol) {
  const int r = h->prm.xfer_rhs_per_thread;
  if (r == 1 || r == 2 || r == 4 || r == 8) return r;
  return ncol >= 8 ? 8 : pick_r(h, ncol);
}

// Stencil kernels of the fp32 solver: R columns per work-item, or (col_loop 8 / 16) a loop over
// G single columns per work-item.  Automatic: 1 (see sdc_params.col_loop).
int pick_g(const sdc_handle_s* h, int ncol) {
  const int g = h->prm.col_loop;
  if ((g == 8 || g == 16) && ncol >= g) return g;
  return 1;
}

// DISPATCH_RG: constexpr R and G for a launch; G > 1 runs with R = 1
#define DISPATCH_RG(Rv, Gv, ...)                                        \
  if ((Gv) == 16) {                                                     \
    constexpr int R = 1, G = 16;                                        \
    __VA_ARGS__;                                                        \
  } else if ((Gv) == 8) {                                               \
    constexpr int R = 1, G = 8;                                         \
    __VA_ARGS__;                                                        \
  } else {                                                              \
    constexpr int G = 1;                                                \
    DISPATCH_R(Rv, __VA_ARGS__)                                         \
  }

#define DISPATCH_R(Rv, ...)                     \
  switch (Rv) {                                 \
    case 8: { constexpr int R = 8; __VA_ARGS__; } break; \
    case 4: { constexpr int R = 4; __VA_ARGS__; } break; \
    case 2: { constexpr int R = 2; __VA_ARGS__; } break; \
    default: { constexpr int R = 1; __VA_ARGS__; } break; \
  }

inline dim3 grid2(uint32_t nN, int ncol, int R) {
  return dim3(nblk(nN), (uint32_t)((ncol + R - 1) / R), 1);
}

// 4-nodes-per-work-item streaming kernels: (node blocks, columns)
inline uint32_t nblk4(size_t n) { return (uint32_t)((n + SDC_NODES4 - 1) / SDC_NODES4); }
inline dim3 grid4(uint32_t nN, int ncol) { return dim3(nblk4(nN), (uint32_t)ncol, 1); }

// Chebyshev step coefficients for the interval [lmax/ratio, lmax] of D^-1 A
struct Cheb {
  double c1[16], c2[16];
};

Cheb cheb_coefs(const sdc_params& p) {
  Cheb c;
  const double lmax = p.cheb_lmax, lmin = p.cheb_lmax / p.cheb_ratio;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double s1 = theta / delta;
  double rho_prev = 1.0 / s1;
  c.c1[0] = 0.0;
  c.c2[0] = 1.0 / theta;
  for (int k = 1; k < 16; ++k) {
    const double rho = 1.0 / (2.0 * s1 - rho_prev);
    c.c1[k] = rho * rho_prev;
    c.c2[k] = 2.0 * rho / delta;
    rho_prev = rho;
  }
  return c;
}

// The V-cycle of levels tail_level .. L-1 for every column in one launch (one work-group per
// column, all vectors in LDS); b and x are the tail level's blocks.
template <typename S>
int tail_vcycle(sdc_handle_s* h, const S* b, S* x, int ncol, const int* done, hipStream_t st) {
  const int nl = (int)h->L.size();
  const int lt = h->tail_level;
  TailDesc d{};
  d.nlev = nl - lt;
  uint32_t off = 0, ymax = 0;
  for (int k = 0; k < d.nlev; ++k) {
    const int l = lt + k;
    d.lv[k].s = make_sten<float>(h, l, 1);
    if (l < nl - 1) d.lv[k].t = make_xfer<float>(h, l);
    d.lv[k].offB = off;
    off += h->L[l].nN;
    d.lv[k].offX = off;
    off += h->L[l].nN;
    ymax = std::max(ymax, h->L[l].nN);
  }
  d.offY = off;
  d.ainv = h->ainv32;
  d.nn = (int)h->L[nl - 1].nN;
  d.cset = h->cur_cset;
  d.astride = (uint32_t)h->astride;
  d.nu_pre = h->prm.nu_coarse > 0 ? h->prm.nu_coarse : h->prm.nu_pre;
  d.nu_post = h->prm.nu_coarse > 0 ? h->prm.nu_coarse : h->prm.nu_post;
  const Cheb cc = cheb_coefs(h->prm);
  for (int k = 0; k < 16; ++k) {
    d.c1[k] = (float)cc.c1[k];
    d.c2[k] = (float)cc.c2[k];
  }
  const size_t bytes = (size_t)(off + ymax) * sizeof(float);
  if (bytes != h->tail_lds) return fail(SDC_ERR_STATE, "internal: tail LDS size");
  launch(h, k_tail_vcycle<S>, dim3(ncol), dim3(1024), bytes, st, d, b, x, h->L[lt].ld,
                     done);
  HIPCK(hipGetLastError());
  return SDC_OK;
}

// one V-cycle x = M^-1 b on level l; with dot_out the last kernel writes partials of b.x.
// T: arithmetic and coefficient type; S: storage type of the V-cycle vectors; B, O: types of
// b and x (T on the finest level, where b is the PCG residual and x the preconditioned one; S
// below).
template <typename T, typename S, typename B, typename O>
int vcycle(sdc_handle_s* h, int l, const B* b, O* x, int ncol, bool dot_out, const int* done,
           hipStream_t st) {
  Work<T>& W = work_of<T>(h);
  const int R = pick_r(h, ncol);
  const int G_ = pick_g(h, ncol);
  const Level& lv = h->L[l];
  const int nl = (int)h->L.size();
  if constexpr (std::is_same<T, float>::value && std::is_same<B, S>::value &&
                std::is_same<O, S>::value) {
    if (h->tail_level > 0 && l == h->tail_level) return tail_vcycle<S>(h, b, x, ncol, done, st);
  }
  if (l == nl - 1) {
    const T* ainv = (const T*)(std::is_same<T, float>::value ? (const void*)h->ainv32
                                                              : (const void*)h->ainv64);
    launch(h, (k_dense_solve<T, B, O>), dim3(nblk(lv.nN), ncol), dim3(SDC_BLOCK), 0, st,
                       ainv, (int)lv.nN, b, x, lv.ld, ncol, done, h->cur_cset, h->astride);
    if constexpr (std::is_same<B, T>::value && std::is_same<O, T>::value) {
      if (dot_out) {
        DISPATCH_R(R, launch(h, (k_dot<T, R>), grid2(lv.nN, ncol, R), dim3(SDC_BLOCK), 0,
                                         st, b, (const T*)x, lv.ld, lv.nN, ncol, W.partial, done));
      }
    }
    return SDC_OK;
  }
  const Sten<T> s = make_sten<T>(h, l, 1);
  const Cheb cc = cheb_coefs(h->prm);
  // coarse levels may use fewer symmetric smoothing steps (nu_coarse, 0 = as the finest)
  const int nuc = h->prm.nu_coarse;
  const int nu1 = (l > 0 && nuc > 0) ? nuc : h->prm.nu_pre;
  const int nu2 = (l > 0 && nuc > 0) ? nuc : h->prm.nu_post;
  // Smoothing steps ping-pong between two S buffers.  When x has the storage type (O == S) it
  // is one of them and the first output goes where the last one then lands in x; otherwise
  // (finest level of the fp32 solver with bf16 storage) the last step writes x directly.
  constexpr bool sep = !std::is_same<O, S>::value;
  // W-cycle step on one level: the coarse correction runs twice, with nu_post smoothing steps
  // between the two (the error propagator stays symmetric: S* C S C S)
  const int ngam = (h->wlevel > 0 && l == h->wlevel) ? 2 : 1;
  // smoothing steps between the two coarse corrections of a W step (any number keeps the cycle
  // symmetric: every Chebyshev polynomial in D^-1 A is self-adjoint in the A inner product)
  const int num = h->prm.wcycle_smooth > 0 ? h->prm.wcycle_smooth : nu2;
  const int swaps = (nu1 >= 2 ? nu1 - 2 : 0) + nu2 + (ngam - 1) * num;
  S* const xa = (S*)W.xa[l];
  S* cur;
  S* oth;
  if constexpr (sep) {
    cur = xa;
    oth = (S*)W.xb0;
  } else {
    cur = (swaps % 2 == 0) ? (S*)x : xa;
    oth = (cur == (S*)x) ? xa : (S*)x;
  }
  const dim3 g = grid2(lv.nN, ncol, R);
  const dim3 blk(SDC_BLOCK);
  const double nodes = (double)lv.nN, cols = (double)ncol, sT = (double)sizeof(T);
  const double sS = (double)sizeof(S), sB = (double)sizeof(B), sO = (double)sizeof(O);
  const T cz = (T)cc.c2[0];
  double* const nodot = nullptr;
  if (nu1 >= 2) {
    // steps 0 and 1 fused
    prof_begin(h, l == 0 ? SDC_PROF_PAIR : -1, st);
    DISPATCH_RG(R, G_, launch(h, (k_smooth_pair<T, B, S, R, G>), grid2(lv.nN, ncol, R * G),
                              blk, 0, st, s, b, cur, lv.ld, ncol, cz, (T)cc.c1[1],
                              (T)cc.c2[1], done));
    prof_end(h, l == 0 ? SDC_PROF_PAIR : -1, st, nodes * 5 * sT + nodes * cols * (sB + sS));
  } else {
    DISPATCH_R(R, launch(h, (k_smooth0<T, B, S, R>), g, blk, 0, st, s, b, cur, lv.ld,
                                     ncol, cz, done));
  }
  for (int k = 2; k < nu1; ++k) {
    if (k == 2) {
      DISPATCH_RG(R, G_, launch(h, (k_cheb<T, B, S, S, R, 2, false, G>),
                                grid2(lv.nN, ncol, R * G), blk, 0, st, s, b, (const S*)cur,
                                (const S*)nullptr, oth, lv.ld, ncol, (T)cc.c1[k], (T)cc.c2[k],
                                cz, nodot, done));
    } else {
      DISPATCH_RG(R, G_, launch(h, (k_cheb<T, B, S, S, R, 1, false, G>),
                                grid2(lv.nN, ncol, R * G), blk, 0, st, s, b, (const S*)cur,
                                (const S*)oth, oth, lv.ld, ncol, (T)cc.c1[k], (T)cc.c2[k], cz,
                                nodot, done));
    }
    std::swap(cur, oth);
  }
  S* const res = (S*)W.res[l];
  const Xfer<T> t = make_xfer<T>(h, l);
  const Level& cl = h->L[l + 1];
  S* const cb = (S*)W.b[l + 1];
  S* const cx = (S*)W.xs[l + 1];
  const int RX = pick_rx(h, ncol);
  auto post_smooth = [&](bool fin) {
    const int nsteps = fin ? nu2 : num;
    for (int k = 0; k < nsteps; ++k) {
      const bool last = k == nsteps - 1 && fin;
      double* part = (last && dot_out) ? W.partial : nodot;
      const T c1 = (T)(k == 0 ? 0.0 : cc.c1[k]), c2 = (T)cc.c2[k];
      const int ptag =
          (l != 0) ? -1 : last ? SDC_PROF_SMOOTH : (k == 0 && fin) ? SDC_PROF_SMOOTH0 : -1;
      prof_begin(h, ptag, st);
      // PREV = 0 on the first step (no previous iterate), 1 after
#define SDC_CHEB_STEP(PREV, OT, OUT)                                                           \
      if (part) {                                                                                \
        DISPATCH_RG(R, G_, launch(h, (k_cheb<T, B, S, OT, R, PREV, true, G>),                    \
                                  grid2(lv.nN, ncol, R * G), blk, 0, st, s, b, (const S*)cur,    \
                                  (const S*)oth, OUT, lv.ld, ncol, c1, c2, cz, part, done));     \
      } else {                                                                                   \
        DISPATCH_RG(R, G_, launch(h, (k_cheb<T, B, S, OT, R, PREV, false, G>),                   \
                                  grid2(lv.nN, ncol, R * G), blk, 0, st, s, b, (const S*)cur,    \
                                  (const S*)oth, OUT, lv.ld, ncol, c1, c2, cz, part, done));     \
      }
      if (sep && last) {
        if (k == 0) {
          SDC_CHEB_STEP(0, O, x)
        } else {
          SDC_CHEB_STEP(1, O, x)
        }
      } else {
        if (k == 0) {
          SDC_CHEB_STEP(0, S, oth)
        } else {
          SDC_CHEB_STEP(1, S, oth)
        }
      }
#undef SDC_CHEB_STEP
      prof_end(h, ptag, st,
               nodes * 5 * sT +
                   nodes * cols * (sB + sS * (k == 0 ? 1 : 2) + (sep && last ? sO : sS)));
      std::swap(cur, oth);
    }
  };
  // the separable restriction starting along x takes the residual into its first pass
  // (k_residual_rx, one work-item per node and R columns; no column loops there)
  const bool fuse_rx = SDC_FUSE_RX && h->prm.xfer_separable && cl.n1x != lv.n1x && G_ == 1;
  for (int gam = 0; gam < ngam; ++gam) {
    if (!fuse_rx) {
      prof_begin(h, l == 0 ? SDC_PROF_RESIDUAL : -1, st);
      DISPATCH_RG(R, G_, launch(h, (k_residual<T, B, S, R, G>), grid2(lv.nN, ncol, R * G), blk,
                                0, st, s, b, (const S*)cur, res, lv.ld, ncol, done));
      prof_end(h, l == 0 ? SDC_PROF_RESIDUAL : -1, st,
               nodes * 4 * sT + nodes * cols * (sB + 2 * sS));
    }
    const int rtag = l == 0 ? SDC_PROF_RESTRICT : -1;
    if (h->prm.xfer_separable) {
      // P^T one axis at a time; the intermediates ping-pong between res and the smoother buffer
      // that the post-smoothing overwrites anyway
      uint32_t dims[3] = {lv.n1x, lv.n1y, lv.n1z};
      const uint32_t cdims[3] = {cl.n1x, cl.n1y, cl.n1z};
      int axes[3], na = 0;
      for (int a = 0; a < 3; ++a)
        if (cdims[a] != dims[a]) axes[na++] = a;
      const S* src = res;
      for (int p = 0; p < na; ++p) {
        const int a = axes[p];
        const bool lastp = p == na - 1;
        S* dst = lastp ? cb : (src == res ? oth : res);
        AxisXfer ax{};
        ax.a0 = dims[0]; ax.a1 = dims[1]; ax.a2 = dims[2];
        dims[a] = cdims[a];
        ax.b0 = dims[0]; ax.b1 = dims[1]; ax.b2 = dims[2];
        ax.nout = ax.b0 * ax.b1 * ax.b2;
        ax.f0 = make_fastdiv(ax.b0);
        ax.f1 = make_fastdiv(ax.b1);
        ax.remap = h->prm.xcd_remap != 0;
        ax.cset = h->cur_cset;
        ax.wstride = a == 0 ? t.wxs : a == 1 ? t.wys : t.wzs;
        const size_t ldo = lastp ? cl.ld : lv.ld;
        const int* fi = a == 0 ? t.rix : a == 1 ? t.riy : t.riz;
        const T* wl = a == 0 ? t.rlx : a == 1 ? t.rly : t.rlz;
        const T* wh = a == 0 ? t.rhx : a == 1 ? t.rhy : t.rhz;
        const T* wop = a == 0 ? t.wx : a == 1 ? t.wy : t.wz;
        if (p == 0 && fuse_rx) {
          // residual + x pass: R columns per work-item, tiles of SDC_BLOCK - 2 nodes
          const dim3 gr((lv.nN + SDC_BLOCK - 3) / (SDC_BLOCK - 2), (uint32_t)((ncol + R - 1) / R));
          prof_begin(h, l == 0 ? SDC_PROF_RESIDUAL : -1, st);
          if (wop) {
            DISPATCH_R(R, launch(h, (k_residual_rx<T, B, S, R, true>), gr, blk, 0, st, s, b,
                                 (const S*)cur, dst, lv.ld, ldo, ncol, ax, fi, t.pix, wl, wh,
                                 wop, done));
          } else {
            DISPATCH_R(R, launch(h, (k_residual_rx<T, B, S, R, false>), gr, blk, 0, st, s, b,
                                 (const S*)cur, dst, lv.ld, ldo, ncol, ax, fi, t.pix, wl, wh,
                                 wop, done));
          }
          prof_end(h, l == 0 ? SDC_PROF_RESIDUAL : -1, st,
                   nodes * 4 * sT + nodes * cols * (sB + sS) + (double)ax.nout * cols * sS);
          src = dst;
          continue;
        }
        const dim3 ga = grid2(ax.nout, ncol, RX);
        prof_begin(h, rtag, st);
#define SDC_RESTRICT_AXIS(AXN, OPD)                                                          \
  DISPATCH_R(RX, launch(h, (k_restrict_axis<T, S, AXN, R, OPD>), ga, blk, 0, st, ax, fi, wl, \
                        wh, wop, src, dst, lv.ld, ldo, ncol, done))
        switch (a * 2 + (wop != nullptr)) {
          case 0: SDC_RESTRICT_AXIS(0, false); break;
          case 1: SDC_RESTRICT_AXIS(0, true); break;
          case 2: SDC_RESTRICT_AXIS(1, false); break;
          case 3: SDC_RESTRICT_AXIS(1, true); break;
          case 4: SDC_RESTRICT_AXIS(2, false); break;
          default: SDC_RESTRICT_AXIS(2, true); break;
        }
#undef SDC_RESTRICT_AXIS
        prof_end(h, rtag, st, ((double)ax.a0 * ax.a1 * ax.a2 + (double)ax.nout) * cols * sS);
        src = dst;
      }
    } else {
      prof_begin(h, rtag, st);
      DISPATCH_R(RX, launch(h, (k_restrict<T, S, R>), grid2(cl.nN, ncol, RX), blk, 0, st,
                                        t,
                                        (const S*)res, cb, lv.ld, cl.ld, ncol, done));
      prof_end(h, rtag, st, (nodes + (double)cl.nN) * cols * sS);
    }
    CK((vcycle<T, S, S, S>(h, l + 1, cb, cx, ncol, false, done, st)));
    prof_begin(h, l == 0 ? SDC_PROF_PROLONG : -1, st);
    if (t.wx) {
      DISPATCH_R(RX, launch(h, (k_prolong_own<T, S, R, true>), grid2(cl.nN, ncol, RX), blk, 0,
                            st, t, (const S*)cx, cur, lv.ld, cl.ld, ncol, done));
    } else {
      DISPATCH_R(RX, launch(h, (k_prolong_own<T, S, R, false>), grid2(cl.nN, ncol, RX), blk, 0,
                            st, t, (const S*)cx, cur, lv.ld, cl.ld, ncol, done));
    }
    prof_end(h, l == 0 ? SDC_PROF_PROLONG : -1, st, (2 * nodes + (double)cl.nN) * cols * sS);
    if (gam + 1 < ngam) post_smooth(false);
  }
  post_smooth(true);
  if constexpr (!sep) {
    if (cur != (S*)x) return fail(SDC_ERR_STATE, "internal: V-cycle buffer parity");
  }
  return SDC_OK;
}

// V-cycle on the finest level: b and x are PCG vectors of type T
template <typename T>
int vcycle0(sdc_handle_s* h, const T* b, T* x, int ncol, bool dot_out, const int* done,
            hipStream_t st) {
  if constexpr (std::is_same<T, float>::value) {
    if (h->prm.vcycle_bf16) return vcycle<float, __bf16, float, float>(h, 0, b, x, ncol, dot_out,
                                                                       done, st);
  }
  return vcycle<T, T, T, T>(h, 0, b, x, ncol, dot_out, done, st);
}

// Runs fn(stream) -- V-cycle launches only, no host synchronisation -- on st, through an executable
// graph recorded on the first call with the same (b, x, ncol, variant) when sdc_params.graph is
// set and none of the V-cycle's kernel classes is being profiled.
template <typename F>
int graph_run(sdc_handle_s* h, const void* b, void* x, int ncol, int variant, hipStream_t st,
              F&& fn) {
  // events bracketing the V-cycle's own kernel classes cannot sit inside the graph
  constexpr uint32_t vcycle_tags = (1u << SDC_PROF_SMOOTH) | (1u << SDC_PROF_SMOOTH0) |
                                   (1u << SDC_PROF_PAIR) | (1u << SDC_PROF_RESIDUAL) |
                                   (1u << SDC_PROF_RESTRICT) | (1u << SDC_PROF_PROLONG);
  if (!h->prm.graph || (h->prof_mask & vcycle_tags)) return fn(st);
  // kernels recorded in operator-set mode carry the column -> set map
  if (h->cur_cset) variant |= 1 << 24;
  for (const auto& g : h->graphs)
    if (g.b == b && g.x == x && g.ncol == ncol && g.variant == variant) {
      HIPCK(hipGraphLaunch(g.exec, st));
      return SDC_OK;
    }
  // bounded cache: a caller cycling through many column counts starts over past 64 graphs
  if (h->graphs.size() >= 64) drop_graphs(h);
  sdc_handle_s::Recorder rec;
  HIPCK(hipGraphCreate(&rec.g, 0));
  h->rec = &rec;
  const int rc = fn(st);
  h->rec = nullptr;
  if (rc != SDC_OK || rec.err != hipSuccess) {
    (void)hipGraphDestroy(rec.g);
    if (rc != SDC_OK) return rc;
    HIPCK(rec.err);
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, rec.g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(rec.g);
  HIPCK(ei);
  h->graphs.push_back({b, x, ncol, variant, exec});
  HIPCK(hipGraphLaunch(exec, st));
  return SDC_OK;
}

// Preconditioned CG on ncol columns: rhs in W.r (overwritten by the residual), solution in x.
// Each column stops when |r| <= tol |b|.  col_mask (device, may be null) excludes columns.
// The convergence test and the iteration count live on the device (k_fin_rr), and every kernel
// of an iteration returns at once after the stop flag is set, so check_every iterations at a
// time run without a host synchronisation -- as one recorded graph when sdc_params.graph is set.
template <typename T>
int pcg(sdc_handle_s* h, T* x, int ncol, double tol, int maxit, const int* col_mask, int* iters,
        hipStream_t st) {
  Work<T>& W = work_of<T>(h);
  const Level& l0 = h->L[0];
  const int R = pick_r(h, ncol);
  const int G_ = pick_g(h, ncol);
  const dim3 g = grid2(l0.nN, ncol, R);
  const dim3 blk(SDC_BLOCK);
  const uint32_t nbx = nblk(l0.nN);
  const size_t len = l0.ld * (size_t)ncol;
  int* done = W.counters + 3;
  const int K = std::max(1, (int)h->prm.check_every);
  // k_fin_rr sets the stop flag; the other kernels check it (a device-scope load at their start)
  // only when several iterations run between two host reads, since with one the kernels after
  // the test change nothing that the solve returns (x is final once r passed the test)
  const int* vdone = K > 1 ? done : nullptr;
  hipLaunchKernelGGL(k_zero<T>, dim3(nblk(len)), blk, 0, st, x, len);
  hipLaunchKernelGGL(k_pcg_init<T>, dim3(1), dim3(64), 0, st, W.counters, W.ctl, W.xtab, x,
                     tol * tol, maxit);
  DISPATCH_R(R, hipLaunchKernelGGL((k_dot<T, R>), g, blk, 0, st, (const T*)W.r, (const T*)W.r,
                                   l0.ld, l0.nN, ncol, W.partial, (const int*)nullptr));
  hipLaunchKernelGGL(k_fin_norm, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx, W.bn2,
                     W.active, 1, (const int*)nullptr);
  if (col_mask)
    hipLaunchKernelGGL(k_and_active, dim3(nblk(ncol)), blk, 0, st, W.active, col_mask, ncol);
  CK(graph_run(h, W.r, W.z, ncol, 8 + (int)sizeof(T), st, [&](hipStream_t gs) {
    return vcycle0<T>(h, W.r, W.z, ncol, true, nullptr, gs);
  }));
  hipLaunchKernelGGL(k_fin_beta, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx, W.rz,
                     W.beta, (const int*)W.active, 1, (const int*)nullptr);
  hipLaunchKernelGGL((k_update_p4<T, true>), grid4(l0.nN, ncol), blk, 0, st, (const T*)W.z, W.p,
                     (const double*)W.beta, l0.ld, l0.nN, (const int*)nullptr);
  const Sten<T> s = make_sten<T>(h, 0, 1);
  // one iteration: q = A p, alpha, x += alpha p and r -= alpha q, the convergence test, z = M r,
  // beta, p = z + beta p -- every launch through launch() so that it can be recorded
  auto iteration = [&](hipStream_t s2) -> int {
    prof_begin(h, SDC_PROF_APPLY, s2);
    DISPATCH_RG(R, G_, launch(h, (k_apply<T, R, true, G>), grid2(l0.nN, ncol, R * G), blk, 0,
                              s2, s, (const T*)W.p, W.q, l0.ld, ncol, W.partial, vdone));
    prof_end(h, SDC_PROF_APPLY, s2,
             (double)l0.nN * 4 * sizeof(T) + (double)l0.nN * ncol * 2 * sizeof(T));
    launch(h, k_fin_alpha, dim3(ncol), blk, 0, s2, (const double*)W.partial, nbx,
           (const double*)W.rz, W.alpha, (const int*)W.active, vdone, (double*)nullptr);
    prof_begin(h, SDC_PROF_UPDATE, s2);
    launch(h, k_update_xr4<T>, grid4(l0.nN, ncol), blk, 0, s2, (T* const*)W.xtab, (const T*)W.p,
           W.r, (const T*)W.q, (const double*)W.alpha, l0.ld, l0.nN, W.partial,
           vdone);
    prof_end(h, SDC_PROF_UPDATE, s2, (double)l0.nN * ncol * 6 * sizeof(T));
    launch(h, k_fin_rr, dim3(ncol), blk, 0, s2, (const double*)W.partial, nblk4(l0.nN),
           (const double*)W.bn2, (const int*)W.ctl, W.rr, W.active, W.counters, done,
           (int*)nullptr);
    CK(vcycle0<T>(h, W.r, W.z, ncol, true, vdone, s2));
    launch(h, k_fin_beta, dim3(ncol), blk, 0, s2, (const double*)W.partial, nbx, W.rz, W.beta,
           (const int*)W.active, 0, vdone);
    prof_begin(h, SDC_PROF_UPDATE_P, s2);
    launch(h, k_update_p4<T, false>, grid4(l0.nN, ncol), blk, 0, s2, (const T*)W.z, W.p,
           (const double*)W.beta, l0.ld, l0.nN, vdone);
    prof_end(h, SDC_PROF_UPDATE_P, s2, (double)l0.nN * ncol * 3 * sizeof(T));
    return SDC_OK;
  };
  const bool rec = h->prm.graph && h->prof_mask == 0;
  const int variant = 2000 + 16 * K + (int)sizeof(T);
  for (int it = 0;; it += K) {
    auto block = [&](hipStream_t s2) -> int {
      for (int k = 0; k < K; ++k) CK(iteration(s2));
      return SDC_OK;
    };
    if (rec)
      CK(graph_run(h, W.counters, (void*)W.partial, ncol, variant, st, block));
    else
      CK(block(st));
    HIPCK(hipMemcpyAsync(h->h_flag, W.counters + 3, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (h->h_flag[0]) break;
    // the device stops at maxit by itself; this bound only guards the host loop
    if (it > maxit + K) return fail(SDC_ERR_STATE, "internal: PCG did not stop");
  }
  HIPCK(hipGetLastError());
  if (iters) *iters = std::min(h->h_flag[1], maxit);
  return SDC_OK;
}

// fp32 PCG with reliable updates: the fp32 partial solution is folded into the fp64 solution u
// and the fp32 residual is replaced by the fp64 true residual q - A u whenever every active
// column's fp32 residual passed the tolerance, or some column's residual dropped by
// reliable_delta since its last replacement.  Columns stop on the true residual only.
// Every decision is taken on the device (k_norm_decide, k_true_after on the control block
// W.ctl): an iteration is a fixed sequence of launches whose kernels return at once when the
// control block says the solve stopped or no update is due, so check_every iterations at a time
// run without a host synchronisation -- as one recorded graph when sdc_params.graph is set.
int pcg_reliable(sdc_handle_s* h, const double* q, double* u, int ncol, const sdc_solve_opts* o,
                 sdc_solve_stats& S, hipStream_t st) {
  CK(ensure_work<float>(h, ncol));
  Work<float>& W = h->w32;
  const Level& l0 = h->L[0];
  const int R = pick_r(h, ncol);
  const int G_ = pick_g(h, ncol);
  const dim3 g = grid2(l0.nN, ncol, R);
  const dim3 blk(SDC_BLOCK);
  const uint32_t nbx = nblk(l0.nN);
  // 4 node blocks per work-group of the fp64 residual: its empty launches (no reliable update)
  // are cheap (cfg 2: 9.7 us).  With 8 its full pass took 686-702 against 705-733 us but
  // fetched 3.84 against 2.77 GB from HBM (5.5 TB/s), too close to the bandwidth that
  // concurrent streams share.
  constexpr int kIrChunks = 4;
  const uint32_t nbx_ir = (nbx + kIrChunks - 1) / kIrChunks;
  // r update: kUpdChunks chunks of SDC_NODES4 nodes per work-group, so that k_norm_decide sums
  // few partials per column
  constexpr int kUpdChunks = 2;
  const dim3 gupd((nblk4(l0.nN) + kUpdChunks - 1) / kUpdChunks, (uint32_t)ncol);
  const size_t len = l0.ld * (size_t)ncol;
  PinFix pf;
  pf.on = h->bc == SDC_BC_NEUMANN;
  pf.n1x = l0.n1x;
  pf.plane = l0.plane;
  pf.cx = l0.cx64;
  pf.cy = l0.cy64;
  pf.cz = l0.cz64;
  const Sten<double> s2 = make_sten<double>(h, 0, 2);
  Sten<float> s1 = make_sten<float>(h, 0, 1);
  if (s1.remap) s1.remap = kApplyBand;
  const double tol2 = o->tol * o->tol;
  const double delta2 = h->prm.reliable_delta * h->prm.reliable_delta;
  const int maxit = o->max_iter;
  // Stall guard: a reliable update is forced after kStallWindow iterations without one, and the
  // solve gives up (converged = 0, the caller continues in fp64) when an update fails to halve
  // the largest true relative residual of the previous one.  This happens when the fp32
  // recursion no longer tracks the true residual (conductivity contrasts of 1e4 and more).
  constexpr int kStallWindow = 25;
  // the device writes the control block to h_ctl when it stops (d_ctl), and the host reads the
  // stop flag there after each block of iterations: cleared before the first launch
  ((int*)h->h_ctl)[SDC_CTL_STOP] = 0;
  // the right-hand-side node mask cleared and the control block set up in one launch; u is not
  // cleared: the first fold writes it (and its column padding) without reading it
  hipLaunchKernelGGL(k_mixed_init, dim3(nblk(l0.nN)), blk, 0, st, W.qmask, l0.nN, W.ctl, W.qu, q,
                     u, tol2, delta2, maxit, kStallWindow);
  // vcycle_bf16 = 3, 4: the multigrid reads a bfloat16 copy of r, written with r
  __bf16* const r16 =
      ((h->prm.vcycle_bf16 == 3 || h->prm.vcycle_bf16 == 4) && h->L.size() > 1) ? W.r16 : nullptr;
  DISPATCH_R(R, hipLaunchKernelGGL((k_ir_residual<float, R, true>), g, blk, 0, st, s2, q,
                                   (const double*)u, W.r, l0.ld, ncol, pf, W.partial, W.qmask, r16,
                                   (const int*)nullptr, (const double* const*)nullptr));
  // aux = |q|^2 and active = |q| > 0; rmax = |q|^2 (the residual at the last replacement)
  hipLaunchKernelGGL(k_fin_qnorm, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx, W.aux,
                     W.rmax, W.active);
  std::vector<double> h_qn(ncol);
  HIPCK(hipMemcpyAsync(h_qn.data(), W.aux, ncol * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCK(hipStreamSynchronize(st));
  int nact = 0;
  for (int c = 0; c < ncol; ++c) nact += h_qn[c] > 0.0;
  S.outer = 0;
  S.iters = 0;
  S.max_rel_res = 0.0;
  S.converged = nact == 0;
  if (nact == 0) {
    hipLaunchKernelGGL(k_zero<double>, dim3(nblk(len)), blk, 0, st, u, len);
    return SDC_OK;
  }
  // vcycle_bf16 = 1, 4: the preconditioned residual z is bfloat16
  const bool zb = (h->prm.vcycle_bf16 == 1 || h->prm.vcycle_bf16 == 4) && h->L.size() > 1;
  __bf16* const z16 = (__bf16*)W.xb0;
  const int K = std::max(1, (int)h->prm.check_every);
  // Early exits after the stop flag only matter when several iterations run between two host
  // reads of it (check_every > 1); the flag is read with a device-scope load at the start of
  // every kernel that checks it, which costs latency, so with one iteration per read the
  // kernels that cannot change u skip the check (the update kernels always read the block).
  const int* stop = K > 1 ? W.ctl + SDC_CTL_STOP : nullptr;
  const double bz = zb ? 2.0 : 4.0;
  // fp64 stencil of the true residual: refine_rhs_per_thread columns per work-item (default 8;
  // its coefficients are fp64, so more columns per work-item save more traffic than in the fp32
  // kernels)
  const int R64 = h->prm.refine_rhs_per_thread > 0 ? h->prm.refine_rhs_per_thread
                  : ncol >= 8                      ? 8
                                                   : R;
  // Iteration: z = M r (z kept in the V-cycle storage type), beta, p = z + beta p and
  // q = A z + beta q in one pass (p into the next ring slot), alpha, r -= alpha q, the update
  // decision, and the predicated fold of the pending directions into u (reliable update or
  // full ring) and fp64 true residual (reliable update).
  // Every launch goes through launch() so that the whole iteration can be recorded.
  bool fold_fresh = true;  // profiled runs only: the next fold writes u without reading it
  int fold_pend = 0;       // profiled runs only: directions the next fold reads
  const int S_ = W.nslot;
  const size_t slot_len = l0.ld * (size_t)W.ncap;
  // iteration `it` (0: the first) writes its search direction to ring slot it % S_
  auto iteration = [&](bool first, hipStream_t s, int it) -> int {
    const int phi = it % S_;
    float* const p_out = W.pr + (size_t)phi * slot_len;
    const float* const p_prev = W.pr + (size_t)((it + S_ - 1) % S_) * slot_len;
    int rc = SDC_OK;
    if (zb && r16)
      rc = vcycle<float, __bf16, __bf16, __bf16>(h, 0, r16, z16, ncol, true, stop, s);
    else if (zb)
      rc = vcycle<float, __bf16, float, __bf16>(h, 0, W.r, z16, ncol, true, stop, s);
    else if (r16)
      rc = vcycle<float, __bf16, __bf16, float>(h, 0, r16, W.z, ncol, true, stop, s);
    else
      rc = vcycle0<float>(h, W.r, W.z, ncol, true, stop, s);
    CK(rc);
    launch(h, k_fin_beta, dim3(ncol), blk, 0, s, (const double*)W.partial, nbx, W.rz, W.beta,
           (const int*)W.active, first ? 1 : 0, stop);
    // the class covers the steady-state launches (the first one skips p, q and x)
    const int atag = first ? -1 : SDC_PROF_APPLY;
    prof_begin(h, atag, s);
#define SDC_APPLY_PQ(ZT, ZP)                                                                   \
    if (first) {                                                                               \
      DISPATCH_RG(R, G_, launch(h, (k_apply_pq<ZT, R, true, G>), grid2(l0.nN, ncol, R * G),    \
                                blk, 0, s, s1, ZP, p_prev, p_out, W.q, (const double*)W.beta,  \
                                l0.ld, ncol, W.partial, (const int*)W.ctl, K > 1 ? 1 : 0));    \
    } else {                                                                                   \
      DISPATCH_RG(R, G_, launch(h, (k_apply_pq<ZT, R, false, G>), grid2(l0.nN, ncol, R * G),   \
                                blk, 0, s, s1, ZP, p_prev, p_out, W.q, (const double*)W.beta,  \
                                l0.ld, ncol, W.partial, (const int*)W.ctl, K > 1 ? 1 : 0));    \
    }
    if (zb) {
      SDC_APPLY_PQ(__bf16, (const __bf16*)z16)
    } else {
      SDC_APPLY_PQ(float, (const float*)W.z)
    }
#undef SDC_APPLY_PQ
    // algorithmic bytes of the steady state: z (stencil), p_prev and q read, p and q written
    prof_end(h, atag, s, (double)l0.nN * 16 + (double)l0.nN * ncol * (bz + 16.0));
    launch(h, k_fin_alpha, dim3(ncol), blk, 0, s, (const double*)W.partial, nbx,
           (const double*)W.rz, W.alpha, (const int*)W.active, stop,
           W.ahist + (size_t)phi * W.ncap);
    prof_begin(h, SDC_PROF_UPDATE, s);
    if (r16)
      launch(h, k_update_r4<true, kUpdChunks>, gupd, blk, 0, s, W.r, (const float*)W.q,
             (const double*)W.alpha, l0.ld, l0.nN, W.partial, r16, stop);
    else
      launch(h, k_update_r4<false, kUpdChunks>, gupd, blk, 0, s, W.r, (const float*)W.q,
             (const double*)W.alpha, l0.ld, l0.nN, W.partial, r16, stop);
    prof_end(h, SDC_PROF_UPDATE, s, (double)l0.nN * ncol * (r16 ? 14 : 12));
    launch(h, k_norm_decide, dim3(1), dim3(SDC_FIN_THREADS), 0, s, (const double*)W.partial,
           gupd.x, W.rr, W.ctl, (const double*)W.aux, (const double*)W.rmax,
           (const int*)W.active, ncol, S_);
    // reliable update (every kernel below returns at once unless k_norm_decide asked for it).  A
    // profiled run brackets the fold and the fp64 residual only when they do work, so it reads
    // the decision on the host (graphs are off while profiling).
    bool upd = true;
    if (prof_on(h, SDC_PROF_REFINE) || prof_on(h, SDC_PROF_FOLD)) {
      HIPCK(hipMemcpyAsync(h->h_ctl, W.ctlmem, SDC_CTL_N * sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCK(hipStreamSynchronize(s));
      const int* hc = (const int*)h->h_ctl;
      upd = hc[SDC_CTL_STOP] == 0 && hc[SDC_CTL_UPD] != 0;
      fold_fresh = hc[SDC_CTL_UFRESH] != 0;
      fold_pend = hc[SDC_CTL_PEND];
    }
    prof_begin(h, upd ? SDC_PROF_FOLD : -1, s);
    launch(h, k_ir_fold_ring,
           dim3((nblk4(l0.ld) + SDC_FOLD_CHUNKS - 1) / SDC_FOLD_CHUNKS, (uint32_t)ncol), blk, 0, s,
           (const double* const*)W.qu, (const float* const*)W.ptab, (const double*)W.ahist, S_,
           phi, (int)W.ncap, l0.ld, l0.nN, (const int*)W.ctl);
    // algorithmic bytes: the pending directions read, u read (unless fresh) and written
    prof_end(h, upd ? SDC_PROF_FOLD : -1, s,
             (double)l0.nN * ncol * (4.0 * fold_pend + (fold_fresh ? 8 : 16)));
    prof_begin(h, upd ? SDC_PROF_REFINE : -1, s);
    // kIrChunks node blocks per work-group: most launches return at once (no update)
    DISPATCH_RG(R64, G_, launch(h, (k_ir_residual<float, R, false, G, kIrChunks>),
                                dim3(nbx_ir, (uint32_t)((ncol + R * G - 1) / (R * G))), blk, 0,
                                s, s2, (const double*)nullptr, (const double*)nullptr, W.r,
                                l0.ld, ncol, pf, W.partial, (unsigned*)W.qmask, r16,
                                (const int*)W.ctl, (const double* const*)W.qu));
    // algorithmic bytes: fp64 coefficients, per column u64 read and r32 write; q is read on
    // source nodes only (not counted)
    prof_end(h, upd ? SDC_PROF_REFINE : -1, s, (double)l0.nN * 32 + (double)l0.nN * ncol * 12);
    launch(h, k_true_after, dim3(1), dim3(SDC_FIN_THREADS), 0, s, (const double*)W.partial,
           nbx_ir, (const double*)W.aux, W.orr, W.rmax, W.active, W.ctl, ncol, h->d_ctl);
    return SDC_OK;
  };
  // the first iteration (FIRST variants) is launched; later ones run check_every at a time,
  // then the host reads the control block
  CK(iteration(true, st, 0));
  // profiled kernel classes are bracketed by events, which cannot sit inside a graph
  const bool rec = h->prm.graph && h->prof_mask == 0;
  const int variant = 1000 + 16 * K + (zb ? 1 : 0) + (r16 ? 2 : 0) + 4 * (R64 == 8);
  const int* h_ctl = (const int*)h->h_ctl;
  for (int done_it = 1;; done_it += K) {
    auto block = [&](hipStream_t s) -> int {
      for (int k = 0; k < K; ++k) CK(iteration(false, s, done_it + k));
      return SDC_OK;
    };
    // a recorded block fixes its ring slots: one graph per starting slot
    if (rec)
      CK(graph_run(h, W.ctl, (void*)W.partial, ncol, variant + ((done_it % S_) << 16), st,
                   block));
    else
      CK(block(st));
    if (!h->d_ctl)
      HIPCK(hipMemcpyAsync(h->h_ctl, W.ctlmem, (kCtlInts + SDC_DCTL_N) * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (__atomic_load_n(h_ctl + SDC_CTL_STOP, __ATOMIC_ACQUIRE)) break;
    // the device stops at maxit by itself; this bound only guards the host loop
    if (done_it > maxit + K) return fail(SDC_ERR_STATE, "internal: mixed PCG did not stop");
  }
  S.iters = std::min(h_ctl[SDC_CTL_IT], maxit);
  S.outer = h_ctl[SDC_CTL_OUTER];
  S.max_rel_res = h->h_ctl[kCtlInts + SDC_DCTL_MREL];
  S.converged = h_ctl[SDC_CTL_REASON] == 1;
  HIPCK(hipGetLastError());
  return SDC_OK;
}

}  // namespace

namespace {
int build_levels(sdc_handle_s* h, hipStream_t st, int nset);

// dense inverses (and the Gauss-Jordan scratch) of the coarsest level for nset operator sets
int alloc_coarse_inverse(sdc_handle_s* h, int nset) {
  dfree(h->ainv64); dfree(h->ainv32); dfree(h->gj_a); dfree(h->gj_f);
  const size_t k = (size_t)nset, nc = (size_t)h->ncoarse;
  CK(dmalloc(&h->ainv64, h->astride * k));
  CK(dmalloc(&h->ainv32, h->astride * k));
  CK(dmalloc(&h->gj_a, h->astride * k));
  CK(dmalloc(&h->gj_f, nc * k));
  return SDC_OK;
}

// Room for nset operator sets: the per-set arrays (coefficients of every level, interpolation
// weights, coarsest inverses) are reallocated when the capacity grows; their contents are then
// undefined until the next sdc_set_sigma_sets.
int ensure_sets(sdc_handle_s* h, int nset) {
  if (nset <= h->nset_cap) return SDC_OK;
  HIPCK(hipDeviceSynchronize());
  drop_graphs(h);
  for (size_t l = 0; l < h->L.size(); ++l) {
    CK(alloc_level_arrays(h->L[l], nset));
    if (l + 1 < h->L.size()) CK(alloc_weights(h->L[l], nset));
  }
  CK(alloc_coarse_inverse(h, nset));
  h->nset_cap = nset;
  h->sigma_set = false;
  return SDC_OK;
}

// adjoint contraction: columns per work-item group, and the partial-sum buffer for ng groups
constexpr int kJtvecGroupCols = SDC_JT_COLS;
int jtvec_groups(int ncol) { return (ncol + kJtvecGroupCols - 1) / kJtvecGroupCols; }

// Device address of `scratch` on each device, resolved once per device with that device
// current (sdc_create); null until then.
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_dev;

int scratch_resolve(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0) return fail(SDC_ERR_ARG, "negative device ordinal");
  if ((size_t)device >= g_scratch_dev.size()) g_scratch_dev.resize((size_t)device + 1, nullptr);
  if (g_scratch_dev[device]) return SDC_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(SDC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_dev[device] = (double*)p;
  return SDC_OK;
}

double* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && (size_t)device < g_scratch_dev.size() ? g_scratch_dev[device] : nullptr;
}

int jtvec_partials(sdc_handle_s* h, int ng, size_t nC, double** part) {
  const size_t need = (size_t)ng * nC;
  if (h->jpart_cap < need) {
    dfree(h->jpart);
    h->jpart_cap = 0;
    CK(dmalloc(&h->jpart, need));
    h->jpart_cap = need;
  }
  *part = h->jpart;
  return SDC_OK;
}
}  // namespace

// ============================================================================================
// C ABI

extern "C" {

const char* sdc_version(void) { return "simpeg_dc 0.1 (gfx950)"; }

const char* sdc_last_error(void) { return g_err.c_str(); }

int sdc_default_params(sdc_params* p) {
  if (!p) return fail(SDC_ERR_ARG, "null params");
  p->nu_pre = 2;
  p->nu_post = 2;
  p->cheb_lmax = 2.0;
  p->cheb_ratio = 8.0;
  p->rhs_per_thread = 0;
  p->xfer_rhs_per_thread = 0;
  p->xcd_remap = 1;
  p->vcycle_bf16 = 1;
  p->xfer_separable = 1;
  p->wcycle_level = -1;
  p->fused_tail = 1;
  p->refine_rhs_per_thread = 0;
  p->graph = 1;
  p->interp = 1;
  p->wcycle_smooth = 1;
  p->nu_coarse = 0;
  p->deriv_robin = 1;
  p->col_loop = 0;
  p->check_every = 1;
  p->reliable_delta = 1e-4;
  return SDC_OK;
}

int sdc_create(const sdc_mesh_desc* d, sdc_handle* out) {
  if (!d || !out) return fail(SDC_ERR_ARG, "null argument");
  *out = nullptr;
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(SDC_ERR_ARG, "mesh needs >= 1 cell per axis");
  if (!d->hx || !d->hy || !d->hz) return fail(SDC_ERR_ARG, "null width array");
  if (d->bc_type != SDC_BC_ROBIN && d->bc_type != SDC_BC_NEUMANN)
    return fail(SDC_ERR_ARG, "bc_type must be SDC_BC_ROBIN or SDC_BC_NEUMANN");
  const uint64_t nN = (uint64_t)(d->nx + 1) * (d->ny + 1) * (d->nz + 1);
  if (nN >= (1ull << 31)) return fail(SDC_ERR_ARG, "mesh too large (nodes >= 2^31)");
  for (int i = 0; i < d->nx; ++i)
    if (!(d->hx[i] > 0)) return fail(SDC_ERR_ARG, "cell widths must be positive");
  for (int i = 0; i < d->ny; ++i)
    if (!(d->hy[i] > 0)) return fail(SDC_ERR_ARG, "cell widths must be positive");
  for (int i = 0; i < d->nz; ++i)
    if (!(d->hz[i] > 0)) return fail(SDC_ERR_ARG, "cell widths must be positive");
  if (d->bc_type == SDC_BC_ROBIN &&
      (!d->beta_xmin || !d->beta_xmax || !d->beta_ymin || !d->beta_ymax || !d->beta_zmin ||
       !d->beta_zmax))
    return fail(SDC_ERR_ARG, "Robin boundary needs the six beta arrays");
  DeviceGuard dg(d->device);
  if (int rc0 = scratch_resolve(d->device)) return rc0;
  sdc_handle_s* h = new sdc_handle_s();
  h->device = d->device;
  h->bc = d->bc_type;
  h->nx = d->nx; h->ny = d->ny; h->nz = d->nz;
  sdc_default_params(&h->prm);
  auto bail = [&](int rc) {
    sdc_destroy(h);
    return rc;
  };
  std::vector<double> hx(d->hx, d->hx + d->nx), hy(d->hy, d->hy + d->ny), hz(d->hz, d->hz + d->nz);
  int rc;
  if ((rc = upload(&h->hx, hx)) || (rc = upload(&h->hy, hy)) || (rc = upload(&h->hz, hz)))
    return bail(rc);
  const size_t nyz = (size_t)d->ny * d->nz, nxz = (size_t)d->nx * d->nz, nxy = (size_t)d->nx * d->ny;
  auto up_or_zero = [&](double** dst, const double* src, size_t n) {
    std::vector<double> v(n, 0.0);
    if (src && d->bc_type == SDC_BC_ROBIN) v.assign(src, src + n);
    return upload(dst, v);
  };
  if ((rc = up_or_zero(&h->bxm, d->beta_xmin, nyz)) || (rc = up_or_zero(&h->bxp, d->beta_xmax, nyz)) ||
      (rc = up_or_zero(&h->bym, d->beta_ymin, nxz)) || (rc = up_or_zero(&h->byp, d->beta_ymax, nxz)) ||
      (rc = up_or_zero(&h->bzm, d->beta_zmin, nxy)) || (rc = up_or_zero(&h->bzp, d->beta_zmax, nxy)))
    return bail(rc);
  // multigrid hierarchy
  Level l0;
  const std::vector<double>* w[3] = {&hx, &hy, &hz};
  for (int a = 0; a < 3; ++a) {
    l0.ax[a].n = (int)w[a]->size();
    l0.ax[a].X.assign(w[a]->size() + 1, 0.0);
    for (size_t i = 0; i < w[a]->size(); ++i) l0.ax[a].X[i + 1] = l0.ax[a].X[i] + (*w[a])[i];
  }
  fill_level_dims(l0);
  h->L.push_back(l0);
  // coarsest level: at most 64 nodes by default (its dense inverse is rebuilt by one work-group
  // at every sdc_set_sigma, O(n^3): 128 nodes took 240 us, 64 nodes ~40 us)
  const uint32_t cmax = d->coarse_max > 0 ? (uint32_t)d->coarse_max : 64u;
  double hmin = 1e300;
  for (int a = 0; a < 3; ++a)
    for (double v : *w[a]) hmin = std::min(hmin, v);
  // target width of coarsening step s: hmin * 2^s (with a small tolerance for round-off)
  for (int step = 0; h->L.back().nN > cmax && h->L.size() < 32 && step < 64; ++step) {
    Level& f = h->L.back();
    if (f.ax[0].n < 2 && f.ax[1].n < 2 && f.ax[2].n < 2) break;
    const double target = hmin * std::ldexp(1.0, step) * (1.0 + 1e-9);
    Level c;
    std::vector<int> fidx[3];
    bool changed = false;
    for (int a = 0; a < 3; ++a) {
      coarsen_axis(f.ax[a], c.ax[a], fidx[a], target);
      changed |= c.ax[a].n != f.ax[a].n;
    }
    if (!changed) continue;
    if ((rc = make_transfer(f, c, fidx))) return bail(rc);
    h->L.push_back(c);
  }
  if (h->L.back().nN > 4096) return bail(fail(SDC_ERR_ARG, "coarsest multigrid level too large"));
  h->wlevel = pick_wlevel(h);
  pick_tail(h);
  for (auto& v : h->L)
    if ((rc = alloc_level_arrays(v, 1))) return bail(rc);
  h->ncoarse = (int)h->L.back().nN;
  h->astride = round_ld((size_t)h->ncoarse * h->ncoarse);
  if ((rc = alloc_coarse_inverse(h, 1))) return bail(rc);
  const size_t n0 = h->L[0].nN;
  if ((rc = dmalloc(&h->tcx, n0)) || (rc = dmalloc(&h->tcy, n0)) || (rc = dmalloc(&h->tcz, n0)) ||
      (rc = dmalloc(&h->tbd, n0)) || (rc = dmalloc(&h->tdg, n0)) || (rc = dmalloc(&h->tidg, n0)))
    return bail(rc);
  if (hipHostMalloc((void**)&h->h_flag, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    h->h_flag = nullptr;
    return bail(fail(SDC_ERR_ALLOC, "hipHostMalloc failed"));
  }
  if (hipHostMalloc((void**)&h->h_ctl, (kCtlInts + SDC_DCTL_N) * sizeof(double), hipHostMallocDefault) !=
      hipSuccess) {
    h->h_ctl = nullptr;
    return bail(fail(SDC_ERR_ALLOC, "hipHostMalloc failed"));
  }
  // pinned host memory is mapped for the device; without a device address the mixed PCG copies
  // its control block to the host after every block of iterations instead
  if (hipHostGetDevicePointer((void**)&h->d_ctl, h->h_ctl, 0) != hipSuccess) {
    h->d_ctl = nullptr;
    (void)hipGetLastError();
  }
  // the fused coarse tail and the coarsest inverse may use all 160 KB of LDS (set once,
  // outside any graph capture)
  if (hipFuncSetAttribute((const void*)k_tail_vcycle<__bf16>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_tail_vcycle<float>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_coarse_inverse<true>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCoarseInvLds) !=
          hipSuccess)
    return bail(fail(SDC_ERR_HIP, "hipFuncSetAttribute failed"));
  *out = h;
  return SDC_OK;
}

int sdc_destroy(sdc_handle h) {
  if (!h) return SDC_OK;
  DeviceGuard dg(h->device);
  (void)hipDeviceSynchronize();
  dfree(h->hx); dfree(h->hy); dfree(h->hz);
  dfree(h->bxm); dfree(h->bxp); dfree(h->bym); dfree(h->byp); dfree(h->bzm); dfree(h->bzp);
  for (auto& v : h->L) v.release();
  dfree(h->ainv64); dfree(h->ainv32); dfree(h->gj_a); dfree(h->gj_f);
  dfree(h->tcx); dfree(h->tcy); dfree(h->tcz); dfree(h->tbd); dfree(h->tdg); dfree(h->tidg);
  dfree(h->jpart);
  dfree(h->cset_dev);
  h->w32.release();
  h->w64.release();
  if (h->h_flag) (void)hipHostFree(h->h_flag);
  if (h->h_ctl) (void)hipHostFree(h->h_ctl);
  drop_graphs(h);
  for (auto& p : h->prof)
    for (auto e : p.ev) (void)hipEventDestroy(e);
  delete h;
  return SDC_OK;
}

int64_t sdc_ld(sdc_handle h) { return h ? (int64_t)h->L[0].ld : -1; }
int64_t sdc_num_nodes(sdc_handle h) { return h ? (int64_t)h->L[0].nN : -1; }
int64_t sdc_num_cells(sdc_handle h) { return h ? (int64_t)h->nx * h->ny * h->nz : -1; }
int32_t sdc_num_levels(sdc_handle h) { return h ? (int32_t)h->L.size() : -1; }

int sdc_set_params(sdc_handle h, const sdc_params* p) {
  if (!h || !p) return fail(SDC_ERR_ARG, "null argument");
  if (p->nu_pre < 1 || p->nu_pre > 15 || p->nu_post < 1 || p->nu_post > 15)
    return fail(SDC_ERR_ARG, "smoothing steps must be in [1, 15]");
  if (!(p->cheb_lmax > 0) || !(p->cheb_ratio > 1))
    return fail(SDC_ERR_ARG, "cheb_lmax must be > 0 and cheb_ratio > 1");
  if (p->rhs_per_thread != 0 && p->rhs_per_thread != 1 && p->rhs_per_thread != 2 &&
      p->rhs_per_thread != 4 && p->rhs_per_thread != 8)
    return fail(SDC_ERR_ARG, "rhs_per_thread must be 0, 1, 2, 4 or 8");
  if (p->xfer_rhs_per_thread != 0 && p->xfer_rhs_per_thread != 1 &&
      p->xfer_rhs_per_thread != 2 && p->xfer_rhs_per_thread != 4 && p->xfer_rhs_per_thread != 8)
    return fail(SDC_ERR_ARG, "xfer_rhs_per_thread must be 0, 1, 2, 4 or 8");
  if (p->xcd_remap != 0 && p->xcd_remap != 1) return fail(SDC_ERR_ARG, "xcd_remap must be 0 or 1");
  if (p->vcycle_bf16 < 0 || p->vcycle_bf16 > 4)
    return fail(SDC_ERR_ARG, "vcycle_bf16 must be 0, 1, 2, 3 or 4");
  if (p->xfer_separable != 0 && p->xfer_separable != 1)
    return fail(SDC_ERR_ARG, "xfer_separable must be 0 or 1");
  if (p->wcycle_level < -1) return fail(SDC_ERR_ARG, "wcycle_level must be >= -1");
  if (p->fused_tail != 0 && p->fused_tail != 1) return fail(SDC_ERR_ARG, "fused_tail must be 0 or 1");
  if (p->refine_rhs_per_thread != 0 && p->refine_rhs_per_thread != 1 &&
      p->refine_rhs_per_thread != 2 && p->refine_rhs_per_thread != 4 &&
      p->refine_rhs_per_thread != 8)
    return fail(SDC_ERR_ARG, "refine_rhs_per_thread must be 0, 1, 2, 4 or 8");
  if (!(p->reliable_delta > 0) || !(p->reliable_delta < 1))
    return fail(SDC_ERR_ARG, "reliable_delta must be in (0, 1)");
  if (p->graph != 0 && p->graph != 1) return fail(SDC_ERR_ARG, "graph must be 0 or 1");
  if (p->interp != 0 && p->interp != 1) return fail(SDC_ERR_ARG, "interp must be 0 or 1");
  if (p->wcycle_smooth < 0 || p->wcycle_smooth > 15)
    return fail(SDC_ERR_ARG, "wcycle_smooth must be in [0, 15]");
  if (p->nu_coarse < 0 || p->nu_coarse > 15) return fail(SDC_ERR_ARG, "nu_coarse must be in [0, 15]");
  if (p->deriv_robin != 0 && p->deriv_robin != 1) return fail(SDC_ERR_ARG, "deriv_robin must be 0 or 1");
  if (p->col_loop != 0 && p->col_loop != 1 && p->col_loop != 8 && p->col_loop != 16)
    return fail(SDC_ERR_ARG, "col_loop must be 0, 1, 8 or 16");
  DeviceGuard dg(h->device);
  drop_graphs(h);
  h->prm = *p;
  h->mixed_stalled = false;
  if (h->prm.check_every < 1) h->prm.check_every = 1;
  h->wlevel = pick_wlevel(h);
  pick_tail(h);
  return SDC_OK;
}

int sdc_set_sigma(sdc_handle h, const double* sigma, void* stream) {
  return sdc_set_sigma_sets(h, 1, sigma, stream);
}

int sdc_set_sigma_sets(sdc_handle h, int nset, const double* sigma, void* stream) {
  if (!h || !sigma) return fail(SDC_ERR_ARG, "null argument");
  if (nset < 1 || nset > 4096) return fail(SDC_ERR_ARG, "nset must be in [1, 4096]");
  DeviceGuard dg(h->device);
  CK(ensure_sets(h, nset));
  hipStream_t st = (hipStream_t)stream;
  const Geo g = make_geo(h);
  Level& l0 = h->L[0];
  const size_t nC = (size_t)h->nx * h->ny * h->nz;
  hipLaunchKernelGGL(k_coef_from_sigma, dim3(nblk(l0.nN), 1, (uint32_t)nset), dim3(SDC_BLOCK), 0,
                     st, g, sigma, l0.cx64, l0.cy64, l0.cz64, l0.bd64, 1.0, nC, l0.sstride);
  return build_levels(h, st, nset);
}

int sdc_set_coefficients(sdc_handle h, const double* cx, const double* cy, const double* cz,
                         const double* bd, void* stream) {
  if (!h || !cx || !cy || !cz || !bd) return fail(SDC_ERR_ARG, "null argument");
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  Level& l0 = h->L[0];
  const size_t bytes = (size_t)l0.nN * sizeof(double);
  HIPCK(hipMemcpyAsync(l0.cx64, cx, bytes, hipMemcpyDeviceToDevice, st));
  HIPCK(hipMemcpyAsync(l0.cy64, cy, bytes, hipMemcpyDeviceToDevice, st));
  HIPCK(hipMemcpyAsync(l0.cz64, cz, bytes, hipMemcpyDeviceToDevice, st));
  HIPCK(hipMemcpyAsync(l0.bd64, bd, bytes, hipMemcpyDeviceToDevice, st));
  return build_levels(h, st, 1);
}

}  // extern "C"

namespace {

// Everything that follows the finest level's edge conductances and Robin diagonal: coarse
// conductances, diagonals, fp32 copies and the coarsest dense inverse.
int build_levels(sdc_handle_s* h, hipStream_t st, int nset) {
  // every launch covers all nset operator sets (grid z, or one work-group per set)
  const uint32_t z = (uint32_t)nset;
  // coarse conductances and, in the same launch, the operator-dependent interpolation weights
  // of every transfer (used when sdc_params.interp = 1), which read the fine conductances
  for (size_t l = 0; l + 1 < h->L.size(); ++l) {
    Level& f = h->L[l];
    Level& c = h->L[l + 1];
    const Xfer<double> t = make_xfer<double>(h, (int)l);
    InterpGrids iw{};
    const uint32_t g[3][3] = {{f.n1x, f.n1y, f.n1z}, {c.n1x, f.n1y, f.n1z}, {c.n1x, c.n1y, f.n1z}};
    double* const w64[3] = {f.x64.wx, f.x64.wy, f.x64.wz};
    float* const w32[3] = {f.x32.wx, f.x32.wy, f.x32.wz};
    const size_t ws[3] = {f.x64.wxs, f.x64.wys, f.x64.wzs};
    size_t nb = nblk(c.nN);
    for (int a = 0; a < 3; ++a) {
      iw.g0[a] = g[a][0];
      iw.g1[a] = g[a][1];
      iw.g2[a] = g[a][2];
      iw.w64[a] = w64[a];
      iw.w32[a] = w32[a];
      iw.wstride[a] = ws[a];
      nb = std::max(nb, (size_t)nblk((size_t)g[a][0] * g[a][1] * g[a][2]));
    }
    hipLaunchKernelGGL(k_coarsen_coef, dim3((uint32_t)nb, 4, z), dim3(SDC_BLOCK), 0, st, t,
                       (const double*)f.cx64, (const double*)f.cy64, (const double*)f.cz64,
                       (const double*)f.bd64, c.cx64, c.cy64, c.cz64, c.bd64, c.n1z, f.sstride,
                       c.sstride, iw);
  }
  // diagonals and fp32 copies of all levels, SDC_DIAG_LEVELS levels per launch
  const int pin = h->bc == SDC_BC_NEUMANN ? 1 : 0;
  for (size_t l0 = 0; l0 < h->L.size(); l0 += SDC_DIAG_LEVELS) {
    DiagLevels tab{};
    uint32_t nb = 0;
    for (size_t l = l0; l < h->L.size() && tab.n < SDC_DIAG_LEVELS; ++l) {
      Level& v = h->L[l];
      tab.lv[tab.n++] = DiagLevel{v.n1x, v.n1y, v.nN, nb, make_fastdiv(v.n1x),
                                  make_fastdiv(v.n1y), v.cx64, v.cy64, v.cz64, v.bd64, v.dg64,
                                  v.idg64, v.sstride,
                                  Coef32{v.cx32, v.cy32, v.cz32, v.dg32, v.idg32}, pin};
      nb += (uint32_t)nblk(v.nN);
    }
    hipLaunchKernelGGL(k_coef_diag, dim3(nb, 1, z), dim3(SDC_BLOCK), 0, st, tab);
  }
  // coarsest level: dense inverse
  const int lc = (int)h->L.size() - 1;
  const Sten<double> sc = make_sten<double>(h, lc, 1);
  const size_t nc = h->ncoarse;
  if (sc.nN != nc) return fail(SDC_ERR_STATE, "internal: coarsest level size");
  // the elimination in LDS when a, inv and the factors fit (nn <= 99 with 160 KB)
  const size_t cinv_lds = (2 * nc * nc + nc) * sizeof(double);
  if (cinv_lds <= kCoarseInvLds)
    hipLaunchKernelGGL(k_coarse_inverse<true>, dim3(z), dim3(1024), cinv_lds, st, sc, h->gj_a,
                       h->ainv64, h->ainv32, h->gj_f, h->astride);
  else
    hipLaunchKernelGGL(k_coarse_inverse<false>, dim3(z), dim3(1024), 0, st, sc, h->gj_a,
                       h->ainv64, h->ainv32, h->gj_f, h->astride);
  HIPCK(hipGetLastError());
  h->nset = nset;
  h->sigma_set = true;
  h->mixed_stalled = false;
  return SDC_OK;
}

}  // namespace

extern "C" {

int sdc_apply(sdc_handle h, int dtype, const void* x, void* y, int ncol, void* stream) {
  if (!h || !x || !y || ncol < 1) return fail(SDC_ERR_ARG, "bad argument");
  if (!h->sigma_set) return fail(SDC_ERR_STATE, "no operator set: call sdc_set_sigma or sdc_set_coefficients first");
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  const Level& l0 = h->L[0];
  const int R = pick_r(h, ncol);
  const dim3 g = grid2(l0.nN, ncol, R);
  if (dtype == SDC_F64) {
    const Sten<double> s = make_sten<double>(h, 0, 2);
    DISPATCH_R(R, hipLaunchKernelGGL((k_apply<double, R, false>), g, dim3(SDC_BLOCK), 0, st, s,
                                     (const double*)x, (double*)y, l0.ld, ncol, (double*)nullptr,
                                     (const int*)nullptr));
  } else if (dtype == SDC_F32) {
    const Sten<float> s = make_sten<float>(h, 0, 2);
    DISPATCH_R(R, hipLaunchKernelGGL((k_apply<float, R, false>), g, dim3(SDC_BLOCK), 0, st, s,
                                     (const float*)x, (float*)y, l0.ld, ncol, (double*)nullptr,
                                     (const int*)nullptr));
  } else {
    return fail(SDC_ERR_ARG, "dtype must be SDC_F32 or SDC_F64");
  }
  HIPCK(hipGetLastError());
  return SDC_OK;
}

int sdc_vcycle(sdc_handle h, int dtype, const void* b, void* z, int ncol, void* stream) {
  if (!h || !b || !z || ncol < 1) return fail(SDC_ERR_ARG, "bad argument");
  if (!h->sigma_set) return fail(SDC_ERR_STATE, "no operator set: call sdc_set_sigma or sdc_set_coefficients first");
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  if (dtype == SDC_F64) {
    CK(ensure_work<double>(h, ncol));
    CK(vcycle0<double>(h, (const double*)b, (double*)z, ncol, false, nullptr, st));
  } else if (dtype == SDC_F32) {
    CK(ensure_work<float>(h, ncol));
    CK(vcycle0<float>(h, (const float*)b, (float*)z, ncol, false, nullptr, st));
  } else {
    return fail(SDC_ERR_ARG, "dtype must be SDC_F32 or SDC_F64");
  }
  HIPCK(hipGetLastError());
  return SDC_OK;
}

int sdc_solve(sdc_handle h, const double* q, double* u, int ncol, const sdc_solve_opts* o,
              sdc_solve_stats* stats, void* stream) {
  if (!h || !q || !u || !o || ncol < 1) return fail(SDC_ERR_ARG, "bad argument");
  if (!h->sigma_set) return fail(SDC_ERR_STATE, "no operator set: call sdc_set_sigma or sdc_set_coefficients first");
  if (o->mode != SDC_MODE_MIXED && o->mode != SDC_MODE_FP64 && o->mode != SDC_MODE_MIXED_IR)
    return fail(SDC_ERR_ARG, "mode must be SDC_MODE_MIXED, SDC_MODE_MIXED_IR or SDC_MODE_FP64");
  if (!(o->tol > 0) || o->max_iter < 1) return fail(SDC_ERR_ARG, "tol must be > 0, max_iter >= 1");
  if (((uintptr_t)u & 31u) || ((uintptr_t)q & 31u))
    return fail(SDC_ERR_ARG, "q and u must be 32-byte aligned");
  if (o->mode == SDC_MODE_MIXED_IR && (!(o->inner_tol > 0) || o->max_outer < 1))
    return fail(SDC_ERR_ARG, "mixed mode needs inner_tol > 0 and max_outer >= 1");
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  const Level& l0 = h->L[0];
  const int R = pick_r(h, ncol);
  const dim3 g = grid2(l0.nN, ncol, R);
  const dim3 blk(SDC_BLOCK);
  const uint32_t nbx = nblk(l0.nN);
  const size_t len = l0.ld * (size_t)ncol;
  PinFix pf;
  pf.on = h->bc == SDC_BC_NEUMANN;
  pf.n1x = l0.n1x;
  pf.plane = l0.plane;
  pf.cx = l0.cx64;
  pf.cy = l0.cy64;
  pf.cz = l0.cz64;
  const Sten<double> s2 = make_sten<double>(h, 0, 2);
  sdc_solve_stats S;
  S.outer = 0;
  S.iters = 0;
  S.max_rel_res = 0.0;
  S.converged = 0;
  S.fp64_iters = 0;
  std::vector<double> h_rr(ncol), h_qn(ncol);
  auto rel_max = [&](const double* d_rr, const double* d_qn) -> int {
    HIPCK(hipMemcpyAsync(h_rr.data(), d_rr, ncol * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCK(hipMemcpyAsync(h_qn.data(), d_qn, ncol * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    double m = 0.0;
    for (int c = 0; c < ncol; ++c)
      if (h_qn[c] > 0) m = std::max(m, std::sqrt(h_rr[c] / h_qn[c]));
    S.max_rel_res = m;
    return SDC_OK;
  };
  // plain fp64 PCG from u = 0
  auto fp64_solve = [&](int* iters) -> int {
    CK(ensure_work<double>(h, ncol));
    Work<double>& W = h->w64;
    hipLaunchKernelGGL(k_zero<unsigned>, dim3(nblk(l0.nN)), blk, 0, st, W.qmask, (size_t)l0.nN);
    DISPATCH_R(R, hipLaunchKernelGGL((k_ir_residual<double, R, true>), g, blk, 0, st, s2, q,
                                     (const double*)u, W.r, l0.ld, ncol, pf, W.partial, W.qmask,
                                     (__bf16*)nullptr, (const int*)nullptr, (const double* const*)nullptr));
    hipLaunchKernelGGL(k_fin_norm, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx, W.aux,
                       W.active, 0, (const int*)nullptr);
    CK(pcg<double>(h, u, ncol, o->tol, o->max_iter, nullptr, iters, st));
    S.outer = 1;
    DISPATCH_R(R, hipLaunchKernelGGL((k_ir_residual<double, R, false>), g, blk, 0, st, s2, q,
                                     (const double*)u, W.z, l0.ld, ncol, pf, W.partial,
                                     W.qmask, (__bf16*)nullptr, (const int*)nullptr, (const double* const*)nullptr));
    hipLaunchKernelGGL(k_fin_norm, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx, W.orr,
                       W.active, 0, (const int*)nullptr);
    CK(rel_max(W.orr, W.aux));
    // converged when every column met the PCG test (recursive residual)
    HIPCK(hipMemcpyAsync(h->h_flag, W.counters + 2, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    S.converged = (h->h_flag[0] == 0) ? 1 : 0;
    return SDC_OK;
  };
  if (o->mode == SDC_MODE_MIXED && h->mixed_stalled) {
    // an earlier mixed solve on this operator stalled: fp64 from the start
    int it = 0;
    CK(fp64_solve(&it));
    S.fp64_iters = it;
  } else if (o->mode == SDC_MODE_MIXED) {
    CK(pcg_reliable(h, q, u, ncol, o, S, st));
    if (!S.converged) {
      // fp64 continuation from the fp32 result: e solves A e = q - A u in fp64 PCG (to the
      // residual reduction that brings the worst column to tol), then u += e
      h->mixed_stalled = true;
      CK(ensure_work<double>(h, ncol));
      Work<double>& W = h->w64;
      const Work<float>& W32 = h->w32;
      DISPATCH_R(R, hipLaunchKernelGGL((k_ir_residual<double, R, false>), g, blk, 0, st, s2, q,
                                       (const double*)u, W.r, l0.ld, ncol, pf, W.partial,
                                       W32.qmask, (__bf16*)nullptr, (const int*)nullptr, (const double* const*)nullptr));
      const double tin = std::min(0.5, 0.5 * o->tol / std::max(S.max_rel_res, 1e-300));
      int it = 0;
      CK(pcg<double>(h, W.x, ncol, tin, o->max_iter, nullptr, &it, st));
      S.fp64_iters = it;
      DISPATCH_R(R, hipLaunchKernelGGL((k_ir_update<double, R>), g, blk, 0, st, u,
                                       (const double*)W.x, l0.ld, l0.nN, ncol));
      DISPATCH_R(R, hipLaunchKernelGGL((k_ir_residual<double, R, false>), g, blk, 0, st, s2, q,
                                       (const double*)u, W.z, l0.ld, ncol, pf, W.partial,
                                       W32.qmask, (__bf16*)nullptr, (const int*)nullptr, (const double* const*)nullptr));
      hipLaunchKernelGGL(k_fin_norm, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx, W.orr,
                         W.active, 0, (const int*)nullptr);
      CK(rel_max(W.orr, W32.aux));
      S.outer += 1;
      S.converged = S.max_rel_res <= o->tol ? 1 : 0;
    }
  } else if (o->mode == SDC_MODE_FP64) {
    int it = 0;
    CK(fp64_solve(&it));
    S.iters = it;
  } else {
    CK(ensure_work<float>(h, ncol));
    Work<float>& W = h->w32;
    hipLaunchKernelGGL(k_zero<double>, dim3(nblk(len)), blk, 0, st, u, len);
    hipLaunchKernelGGL(k_zero<unsigned>, dim3(nblk(l0.nN)), blk, 0, st, W.qmask, (size_t)l0.nN);
    DISPATCH_R(R, hipLaunchKernelGGL((k_ir_residual<float, R, true>), g, blk, 0, st, s2, q,
                                     (const double*)u, W.r, l0.ld, ncol, pf, W.partial, W.qmask,
                                     (__bf16*)nullptr, (const int*)nullptr, (const double* const*)nullptr));
    hipLaunchKernelGGL(k_fin_norm, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx, W.aux,
                       W.active, 0, (const int*)nullptr);
    hipLaunchKernelGGL(k_fin_norm, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx, W.orr,
                       W.active, 0, (const int*)nullptr);
    const double tol2 = o->tol * o->tol;
    for (int outer = 0;; ++outer) {
      if (outer > 0) {
        prof_begin(h, SDC_PROF_REFINE, st);
        DISPATCH_R(R, hipLaunchKernelGGL((k_ir_residual<float, R, false>), g, blk, 0, st, s2, q,
                                         (const double*)u, W.r, l0.ld, ncol, pf, W.partial,
                                         W.qmask, (__bf16*)nullptr, (const int*)nullptr, (const double* const*)nullptr));
        prof_end(h, SDC_PROF_REFINE, st, (double)l0.nN * 32 + (double)l0.nN * ncol * 20);
        hipLaunchKernelGGL(k_fin_norm, dim3(ncol), blk, 0, st, (const double*)W.partial, nbx,
                           W.orr, W.active, 0, (const int*)nullptr);
      }
      hipLaunchKernelGGL(k_zero<int>, dim3(1), blk, 0, st, W.ocount, (size_t)1);
      hipLaunchKernelGGL(k_outer_check, dim3(nblk(ncol)), blk, 0, st, (const double*)W.orr,
                         (const double*)W.aux, tol2, W.omask, W.ocount, ncol);
      HIPCK(hipMemcpyAsync(h->h_flag, W.ocount, sizeof(int), hipMemcpyDeviceToHost, st));
      CK(rel_max(W.orr, W.aux));
      const int open = h->h_flag[0];
      S.outer = outer;
      if (open == 0) {
        S.converged = 1;
        break;
      }
      if (outer == o->max_outer) break;
      int it = 0;
      CK(pcg<float>(h, W.x, ncol, o->inner_tol, o->max_iter, W.omask, &it, st));
      S.iters += it;
      DISPATCH_R(R, hipLaunchKernelGGL((k_ir_update<float, R>), g, blk, 0, st, u, (const float*)W.x,
                                       l0.ld, l0.nN, ncol));
    }
  }
  HIPCK(hipGetLastError());
  if (stats) *stats = S;
  if (!S.converged) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "solver stopped at relative residual %.3e (tol %.3e)",
                  S.max_rel_res, o->tol);
    return fail(SDC_ERR_NOCONV, buf);
  }
  return SDC_OK;
}

int sdc_solve_sets(sdc_handle h, const double* q, double* u, int ncol, const int32_t* colset,
                   const sdc_solve_opts* o, sdc_solve_stats* stats, void* stream) {
  if (!h || !colset) return fail(SDC_ERR_ARG, "null argument");
  if (ncol < 1) return fail(SDC_ERR_ARG, "ncol must be >= 1");
  if (!h->sigma_set) return fail(SDC_ERR_STATE, "sdc_set_sigma_sets must be called first");
  for (int c = 0; c < ncol; ++c) {
    if (colset[c] < 0 || colset[c] >= h->nset)
      return fail(SDC_ERR_ARG, "colset[" + std::to_string(c) + "] = " + std::to_string(colset[c]) +
                                   " outside the " + std::to_string(h->nset) + " operator sets");
    if (colset[c] != colset[c & ~7])
      return fail(SDC_ERR_ARG, "every aligned group of 8 columns must use one operator set "
                               "(column " + std::to_string(c) + ")");
  }
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  if (h->cset_cap < ncol) {
    // recorded iterations captured the old array's address
    HIPCK(hipStreamSynchronize(st));
    drop_graphs(h);
    dfree(h->cset_dev);
    h->cset_cap = 0;
    CK(dmalloc(&h->cset_dev, (size_t)ncol));
    h->cset_cap = ncol;
  }
  HIPCK(hipMemcpyAsync(h->cset_dev, colset, (size_t)ncol * sizeof(int32_t), hipMemcpyHostToDevice,
                       st));
  h->cur_cset = h->cset_dev;
  const int rc = sdc_solve(h, q, u, ncol, o, stats, stream);
  h->cur_cset = nullptr;
  return rc;
}

int sdc_jvec_rhs(sdc_handle h, const double* dsig, const double* u, double* out, int ncol,
                 void* stream) {
  if (!h || !dsig || !u || !out || ncol < 1) return fail(SDC_ERR_ARG, "bad argument");
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  const Level& l0 = h->L[0];
  const Geo geo = make_geo(h, true);
  hipLaunchKernelGGL(k_coef_from_sigma, dim3(nblk(l0.nN)), dim3(SDC_BLOCK), 0, st, geo, dsig,
                     h->tcx, h->tcy, h->tcz, h->tbd, -1.0, (size_t)0, (size_t)0);
  DiagLevels tab{};
  tab.n = 1;
  tab.lv[0] = DiagLevel{l0.n1x, l0.n1y, l0.nN, 0, make_fastdiv(l0.n1x), make_fastdiv(l0.n1y),
                        h->tcx, h->tcy, h->tcz, h->tbd, h->tdg, h->tidg, 0, Coef32{}, 0};
  hipLaunchKernelGGL(k_coef_diag, dim3(nblk(l0.nN)), dim3(SDC_BLOCK), 0, st, tab);
  Sten<double> s{};
  s.cx = h->tcx; s.cy = h->tcy; s.cz = h->tcz; s.dg = h->tdg; s.idg = h->tidg;
  s.n1x = l0.n1x; s.n1y = l0.n1y; s.n1z = l0.n1z; s.nN = l0.nN; s.plane = l0.plane;
  s.fx = make_fastdiv(l0.n1x); s.fy = make_fastdiv(l0.n1y);
  s.pin0 = 0;
  s.remap = h->prm.xcd_remap != 0;
  const int R = pick_r(h, ncol);
  DISPATCH_R(R, hipLaunchKernelGGL((k_apply<double, R, false>), grid2(l0.nN, ncol, R),
                                   dim3(SDC_BLOCK), 0, st, s, u, out, l0.ld, ncol,
                                   (double*)nullptr, (const int*)nullptr));
  HIPCK(hipGetLastError());
  return SDC_OK;
}

int sdc_jtvec_contract(sdc_handle h, const double* u, const double* lam, int ncol,
                       const int32_t* ucol, int per_column, double* out, int64_t ldo,
                       void* stream) {
  if (!h || !u || !lam || !out || ncol < 1) return fail(SDC_ERR_ARG, "bad argument");
  if (per_column < 0 || per_column > 1) return fail(SDC_ERR_ARG, "per_column must be 0 or 1");
  const int64_t nC = (int64_t)h->nx * h->ny * h->nz;
  if (per_column && ldo < nC) return fail(SDC_ERR_ARG, "ldo must be >= number of cells");
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  const Geo geo = make_geo(h, true);
  const Level& l0 = h->L[0];
  const int ng = jtvec_groups(ncol);
  double* part = nullptr;
  if (!per_column && ng > 1) CK(jtvec_partials(h, ng, (size_t)nC, &part));
  const dim3 grid(nblk(nC), (uint32_t)ng);
  prof_begin(h, SDC_PROF_JTVEC, st);
  if (per_column)
    hipLaunchKernelGGL(k_jtvec<1>, grid, dim3(SDC_BLOCK), 0, st, geo, u, lam, l0.ld, ncol,
                       (const int*)ucol, out, (size_t)ldo, (const double*)nullptr,
                       (double*)nullptr);
  else
    hipLaunchKernelGGL(k_jtvec<0>, grid, dim3(SDC_BLOCK), 0, st, geo, u, lam, l0.ld, ncol,
                       (const int*)ucol, out, (size_t)ldo, (const double*)nullptr, part);
  prof_end(h, SDC_PROF_JTVEC, st,
           (double)l0.nN * ncol * 16 + (double)nC * 8 * (per_column ? ncol : 1));
  if (part)
    hipLaunchKernelGGL(k_jtvec_reduce<0>, dim3(nblk(nC)), dim3(SDC_BLOCK), 0, st,
                       (const double*)part, ng, (uint32_t)nC, out);
  HIPCK(hipGetLastError());
  return SDC_OK;
}

int sdc_jtvec_contract_noalloc(sdc_handle h, const double* u, const double* lam, int ncol,
                               const int32_t* ucol, double* out, void* stream) {
  if (!h || !u || !lam || !out || ncol < 1) return fail(SDC_ERR_ARG, "bad argument");
  const int64_t nC = (int64_t)h->nx * h->ny * h->nz;
  if (nC > (int64_t)SDC_SCRATCH_DOUBLES)
    return fail(SDC_ERR_ARG, "sdc_jtvec_contract_noalloc: the mesh has " + std::to_string(nC) +
                                 " cells, more than the " + std::to_string(SDC_SCRATCH_DOUBLES) +
                                 " its fixed scratch buffer holds (use sdc_jtvec_contract)");
  DeviceGuard dg(h->device);
  double* scr = scratch_on(h->device);
  if (!scr) return fail(SDC_ERR_STATE, "scratch buffer not resolved on the handle's device");
  hipStream_t st = (hipStream_t)stream;
  const Geo geo = make_geo(h, true);
  const size_t ld = h->L[0].ld;
  const int ng = jtvec_groups(ncol);
  if (ng == 1) {
    hipLaunchKernelGGL(k_jtvec<0>, dim3(nblk(nC), 1), dim3(SDC_BLOCK), 0, st, geo, u, lam, ld,
                       ncol, (const int*)ucol, out, (size_t)nC, (const double*)nullptr,
                       (double*)nullptr);
    HIPCK(hipGetLastError());
    return SDC_OK;
  }
  // as many column groups per pass as the scratch holds partials for; every pass overwrites
  // the partials of all its groups, and its reduction continues the running sum in out
  const int per_pass = (int)std::min<int64_t>(ng, (int64_t)SDC_SCRATCH_DOUBLES / nC);
  for (int g0 = 0; g0 < ng; g0 += per_pass) {
    const int gn = std::min(per_pass, ng - g0);
    const int cbeg = g0 * kJtvecGroupCols;
    const int nc = std::min(ncol - cbeg, gn * kJtvecGroupCols);
    const double* up = ucol ? u : u + (size_t)cbeg * ld;
    const int32_t* uc = ucol ? ucol + cbeg : nullptr;
    hipLaunchKernelGGL(k_jtvec<0>, dim3(nblk(nC), (uint32_t)gn), dim3(SDC_BLOCK), 0, st, geo, up,
                       lam + (size_t)cbeg * ld, ld, nc, (const int*)uc, out, (size_t)nC,
                       (const double*)nullptr, scr);
    if (g0 == 0)
      hipLaunchKernelGGL(k_jtvec_reduce<0>, dim3(nblk(nC)), dim3(SDC_BLOCK), 0, st,
                         (const double*)scr, gn, (uint32_t)nC, out);
    else
      hipLaunchKernelGGL(k_jtvec_reduce<2>, dim3(nblk(nC)), dim3(SDC_BLOCK), 0, st,
                         (const double*)scr, gn, (uint32_t)nC, out);
  }
  HIPCK(hipGetLastError());
  return SDC_OK;
}

int sdc_jtj_diag_accumulate(sdc_handle h, const double* u, const double* lam, int ncol,
                            const int32_t* ucol, const double* wsq, double* out, void* stream) {
  if (!h || !u || !lam || !out || !wsq || ncol < 1) return fail(SDC_ERR_ARG, "bad argument");
  const int64_t nC = (int64_t)h->nx * h->ny * h->nz;
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  const Geo geo = make_geo(h, true);
  const int ng = jtvec_groups(ncol);
  double* part = nullptr;
  if (ng > 1) CK(jtvec_partials(h, ng, (size_t)nC, &part));
  hipLaunchKernelGGL(k_jtvec<2>, dim3(nblk(nC), (uint32_t)ng), dim3(SDC_BLOCK), 0, st, geo, u,
                     lam, h->L[0].ld, ncol, (const int*)ucol, out, (size_t)nC, wsq, part);
  if (part)
    hipLaunchKernelGGL(k_jtvec_reduce<1>, dim3(nblk(nC)), dim3(SDC_BLOCK), 0, st,
                       (const double*)part, ng, (uint32_t)nC, out);
  HIPCK(hipGetLastError());
  return SDC_OK;
}

int sdc_profile(sdc_handle h, uint32_t mask) {
  if (!h) return fail(SDC_ERR_ARG, "null handle");
  if (mask & ~(((1u << SDC_PROF_NTAGS) - 1u) & ~1u)) return fail(SDC_ERR_ARG, "unknown profile tag");
  h->prof_mask = mask;
  for (auto& p : h->prof) {
    p.used = 0;
    p.bytes.clear();
  }
  return SDC_OK;
}

int sdc_profile_read(sdc_handle h, int tag, int64_t* count, double* total_ms, double* total_bytes) {
  if (!h || !count || !total_ms || !total_bytes) return fail(SDC_ERR_ARG, "null argument");
  if (tag <= 0 || tag >= SDC_PROF_NTAGS) return fail(SDC_ERR_ARG, "unknown profile tag");
  DeviceGuard dg(h->device);
  auto& p = h->prof[tag];
  double ms = 0.0, by = 0.0;
  const size_t n = p.used / 2;
  for (size_t i = 0; i < n; ++i) {
    HIPCK(hipEventSynchronize(p.ev[2 * i + 1]));
    float t = 0.f;
    HIPCK(hipEventElapsedTime(&t, p.ev[2 * i], p.ev[2 * i + 1]));
    ms += t;
    by += p.bytes[i];
  }
  *count = (int64_t)n;
  *total_ms = ms;
  *total_bytes = by;
  p.used = 0;
  p.bytes.clear();
  return SDC_OK;
}

int sdc_gather(sdc_handle h, int nrow, const int32_t* rowptr, const int64_t* idx, const double* w,
               const double* u, double* d, void* stream) {
  if (!h || nrow < 0) return fail(SDC_ERR_ARG, "bad argument");
  if (nrow == 0) return SDC_OK;
  if (!rowptr || !idx || !w || !u || !d) return fail(SDC_ERR_ARG, "null pointer");
  DeviceGuard dg(h->device);
  hipLaunchKernelGGL(k_gather, dim3(nblk(nrow)), dim3(SDC_BLOCK), 0, (hipStream_t)stream, nrow,
                     (const int*)rowptr, idx, w, u, d);
  HIPCK(hipGetLastError());
  return SDC_OK;
}

int sdc_scatter(sdc_handle h, int nrow, const int64_t* dst, const int32_t* rowptr,
                const int32_t* src, const double* w, const double* v, double* out, int ncol,
                void* stream) {
  if (!h || nrow < 0 || ncol < 0 || !out) return fail(SDC_ERR_ARG, "bad argument");
  DeviceGuard dg(h->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t len = h->L[0].ld * (size_t)ncol;
  if (len) hipLaunchKernelGGL(k_zero<double>, dim3(nblk(len)), dim3(SDC_BLOCK), 0, st, out, len);
  if (nrow > 0) {
    if (!dst || !rowptr || !src || !w || !v) return fail(SDC_ERR_ARG, "null pointer");
    hipLaunchKernelGGL(k_scatter, dim3(nblk(nrow)), dim3(SDC_BLOCK), 0, st, nrow, dst,
                       (const int*)rowptr, (const int*)src, w, v, out);
  }
  HIPCK(hipGetLastError());
  return SDC_OK;
}

}  // extern "C"
