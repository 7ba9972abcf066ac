// nvector_hip.hip -- HIP N_Vector for gfx950 with fused ops and
// deterministic reductions.
//
// Replaces NVECTOR_CUDA and the reference's fused layer
// (src/irregular/2d_fft/2d_i4/deferred_nvector.cu:95-235,
//  src/2d/2d_heter/deferred_nvector.cu:135-151,209-236).
//
// Design (MI355X): every vector carries the stream its ops launch on.
// Element-wise ops are grid-stride kernels over 64-wide wavefronts.
// A reduction is a streaming launch whose workgroups each reduce a slice
// (independent accumulators, wave shuffles, LDS across the 4 waves) into
// a fixed partial slot, followed by a one-workgroup launch that sums the
// partials in slot order and stores the K results straight into pinned
// host memory; the host then waits once.  The streaming launch has no
// device-scope fence or atomic: on the multi-XCD part those write back
// and invalidate the XCD's L2, which costs more than the second launch.
// (A single-launch form, where the last workgroup to take an atomic
// ticket does the final sum, is kept selectable.)  The result does not
// depend on workgroup scheduling, unlike the reference's atomicAdd
// accumulation (deferred_nvector.cu:127-129).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "llg/nvector_hip.h"

// Floating-point contraction is off in this file: every a*b+c that should
// round once is written as fma(), everything else rounds each operation.
// The arithmetic of each op is therefore fixed by the source, and a fused
// kernel that repeats an op sequence reproduces its bits.
#pragma clang fp contract(off)

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMaxRedBlocks = 4096;
constexpr int kMaxK = NVHIP_FUSED_MAX;

struct NvContent {
  sunindextype n = 0;
  double* d = nullptr;
  double* h = nullptr;
  bool own_d = false;
  hipStream_t stream = nullptr;
};

inline NvContent* C(N_Vector v) { return static_cast<NvContent*>(v->content); }
inline double* D(N_Vector v) { return C(v)->d; }
inline sunindextype LEN(N_Vector v) { return C(v)->n; }
inline hipStream_t S(N_Vector v) { return C(v)->stream; }

// ----------------------------------------------------- reduction scratch
// Result slots: pinned host memory holding kMaxK results + a sequence
// word each.  Slot 0 serves every reduction; slot 1 holds the multi-dot of
// a Gram-Schmidt projection while its combination + norm uses slot 0, so
// one host wait covers both.  `dres` is a device copy of slot 1 that the
// combination kernel reads its coefficients from.
constexpr int kSlot = kMaxK + 1;
struct RedWs {
  double* partials = nullptr;   // kMaxRedBlocks * kMaxK
  unsigned* ticket = nullptr;   // zero between launches
  double* host = nullptr;       // pinned, 2 slots of kSlot doubles
  double* host_dev = nullptr;   // device alias of host
  double* dres = nullptr;       // device memory, kMaxK doubles
  unsigned seq = 0;             // last sequence number issued on this stream
  void* sp = nullptr;           // device state of a device-resident SPGMR solve
  // gate (NVHip_GateOpen/Release): pinned mailbox, device copies of its
  // coefficients and of its stop flag, last gate number, whether one is open
  struct GateBox* gbox = nullptr;
  struct GateBox* gbox_dev = nullptr;   // device alias of gbox
  double* gcoef = nullptr;
  int* gskip = nullptr;
  unsigned gseq = 0;
  bool gopen = false;
  // device-driven lagged SPGMR (NVHip_Lag*): device state, pinned published
  // result and its device alias, the current solve's number and parameters
  struct LagDev* lag = nullptr;
  struct LagPub* lagpub = nullptr;
  struct LagPub* lagpub_dev = nullptr;
  unsigned lagseq = 0;
  int lagmaxl = 0;
  double lagdelta = 0.0, lagbfloor = -1.0, lagnlen = 1.0;
};

// Mailbox of a gate in pinned host memory, written by the host only.
struct GateBox {
  double coef[kMaxK];
  int stop;
  unsigned seq;   // written last (release): the gate number being released
};

std::mutex g_mu;
std::unordered_map<hipStream_t, RedWs*> g_ws;
hipStream_t g_default_stream = nullptr;
bool g_default_stream_made = false;
std::atomic<long> g_syncs{0};
int g_sync_mode = 0;   // 0: hipStreamSynchronize, 1: spin on the pinned sequence word
int g_red_blocks = 1024;  // partial workgroups per reduction (upper bound)
int g_red_mode = 1;       // 0: single launch with ticket, 1: partial + finish launches

RedWs* workspace(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_ws.find(s);
  if (it != g_ws.end()) return it->second;
  RedWs* w = new RedWs();
  if (hipMalloc(&w->partials, sizeof(double) * kMaxRedBlocks * kMaxK) != hipSuccess ||
      hipMalloc(&w->ticket, sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&w->dres, sizeof(double) * kMaxK) != hipSuccess ||
      hipHostMalloc(&w->host, sizeof(double) * 2 * kSlot, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&w->host_dev, w->host, 0) != hipSuccess) {
    fprintf(stderr, "[nvector_hip] reduction workspace allocation failed\n");
    abort();
  }
  hipMemset(w->ticket, 0, sizeof(unsigned));
  std::memset(w->host, 0, sizeof(double) * 2 * kSlot);
  hipDeviceSynchronize();
  g_ws[s] = w;
  return w;
}

inline int ew_grid(sunindextype n) {
  sunindextype g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return (int)(g < 1 ? 1 : g);
}
inline int red_grid(sunindextype n) {
  // at least ~8 elements per thread, at most g_red_blocks workgroups
  sunindextype g = (n + 8 * kBlock - 1) / (8 * kBlock);
  if (g > g_red_blocks) g = g_red_blocks;
  return (int)(g < 1 ? 1 : g);
}

void check_launch(const char* what) {
  hipError_t e = hipPeekAtLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "[nvector_hip] %s launch failed: %s\n", what, hipGetErrorString(e));
  }
}

// ------------------------------------------------------- element-wise ops
#define GRID_STRIDE(i, n) \
  for (sunindextype i = (sunindextype)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (sunindextype)gridDim.x * blockDim.x)

__global__ void k_linearsum(double a, const double* __restrict__ x, double b,
                            const double* __restrict__ y, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = fma(a, x[i], b * y[i]);
}
__global__ void k_const(double c, double* z, sunindextype n) { GRID_STRIDE(i, n) z[i] = c; }
__global__ void k_prod(const double* x, const double* y, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = x[i] * y[i];
}
__global__ void k_div(const double* x, const double* y, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = x[i] / y[i];
}
__global__ void k_scale(double c, const double* x, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = c * x[i];
}
__global__ void k_abs(const double* x, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = fabs(x[i]);
}
__global__ void k_inv(const double* x, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = 1.0 / x[i];
}
__global__ void k_addconst(const double* x, double b, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = x[i] + b;
}
__global__ void k_compare(double c, const double* x, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = (fabs(x[i]) >= c) ? 1.0 : 0.0;
}

struct PtrPack { const double* p[kMaxK]; };
struct OutPack { double* p[kMaxK]; };
struct CoefPack { double c[kMaxK]; };

// sum_j c_j X_j[i] in the order every linear-combination kernel uses
__device__ inline double lincomb_elem(int nv, const CoefPack& c, const PtrPack& X,
                                      sunindextype i) {
  double s = c.c[0] * X.p[0][i];
  for (int j = 1; j < nv; j++) s = fma(c.c[j], X.p[j][i], s);
  return s;
}
// z = sum_j c_j X_j   (z may alias any X_j: each element read before write)
__global__ void k_lincomb(int nv, CoefPack c, PtrPack X, double* z, sunindextype n) {
  GRID_STRIDE(i, n) z[i] = lincomb_elem(nv, c, X, i);
}
// x = (0 + sum_j c_j X_j) / s: a linear combination into a zeroed vector
// followed by N_VDiv, the SPGMR solution update
__global__ void k_lincombdiv(int nv, CoefPack c, PtrPack X, const double* __restrict__ sv,
                             double* x, sunindextype n) {
  GRID_STRIDE(i, n) {
    double acc = 1.0 * 0.0;
    for (int j = 0; j < nv; j++) acc = fma(c.c[j], X.p[j][i], acc);
    x[i] = acc / sv[i];
  }
}
// xo = c x, vo = xo / s  (the Krylov basis normalisation fused with the
// next iteration's unscaling)
__global__ void k_scalediv(double c, const double* x, const double* __restrict__ sv, double* xo,
                           double* vo, sunindextype n) {
  GRID_STRIDE(i, n) {
    const double a = c * x[i];
    xo[i] = a;
    vo[i] = a / sv[i];
  }
}
// Z_j = a_j x + Y_j
__global__ void k_scaleaddmulti(int nv, CoefPack a, const double* x, PtrPack Y, OutPack Z,
                                sunindextype n) {
  GRID_STRIDE(i, n) {
    const double xi = x[i];
    for (int j = 0; j < nv; j++) Z.p[j][i] = fma(a.c[j], xi, Y.p[j][i]);
  }
}
// Z_j = a X_j + b Y_j
__global__ void k_lsumarray(int nv, double a, PtrPack X, double b, PtrPack Y, OutPack Z,
                            sunindextype n) {
  GRID_STRIDE(i, n) {
    for (int j = 0; j < nv; j++) Z.p[j][i] = fma(a, X.p[j][i], b * Y.p[j][i]);
  }
}
// Z_j = c_j X_j
__global__ void k_scalearray(int nv, CoefPack c, PtrPack X, OutPack Z, sunindextype n) {
  GRID_STRIDE(i, n) {
    for (int j = 0; j < nv; j++) Z.p[j][i] = c.c[j] * X.p[j][i];
  }
}
__global__ void k_constarray(int nv, double c, OutPack Z, sunindextype n) {
  GRID_STRIDE(i, n) {
    for (int j = 0; j < nv; j++) Z.p[j][i] = c;
  }
}

// Nordsieck predict (sign +1) / restore (-1): for k=1..q, j=q..k:
// z[j-1] += sign*z[j]; z[q] unchanged.  Same per-element operations as the
// q(q+1)/2 N_VLinearSum calls, one pass over memory.
template <int Q>
__global__ void k_pascal(OutPack Z, double sign, sunindextype n) {
  GRID_STRIDE(i, n) {
    double z[Q + 1];
#pragma unroll
    for (int j = 0; j <= Q; j++) z[j] = Z.p[j][i];
#pragma unroll
    for (int k = 1; k <= Q; k++)
#pragma unroll
      for (int j = Q; j >= k; j--) z[j - 1] = fma(sign, z[j], z[j - 1]);
#pragma unroll
    for (int j = 0; j < Q; j++) Z.p[j][i] = z[j];
  }
}

// the predictor followed by the Newton iteration's start: acor = 0 and
// y = zn[0] + acor, as N_VConst and N_VLinearSum would write them; with w
// set, first the error weights of the step from the zn[0] being replaced,
// as k_ewt forms them
// COMP: first the previous step's completion z_j = fma(l_j, acor, z_j)
// (N_VScaleAddMulti, j = 0..Q) and the step-size rescale z_j = c_j z_j
// (N_VScaleVectorArray, j = 1..Q; c_j = 1 leaves the value unchanged), then
// the rest as above; zn[Q] is written too
template <int Q, bool COMP>
__global__ void k_pascal_init(OutPack Z, double* __restrict__ acor, double* __restrict__ y,
                              sunindextype n, double rtol, double atol,
                              const double* __restrict__ av, double* __restrict__ w,
                              CoefPack lc, CoefPack sc, double* __restrict__ save) {
  GRID_STRIDE(i, n) {
    double z[Q + 1];
#pragma unroll
    for (int j = 0; j <= Q; j++) z[j] = Z.p[j][i];
    if constexpr (COMP) {
      const double a = acor[i];
      if (save) save[i] = a;   // the completed step's correction kept (N_VScale(1, acor, save))
#pragma unroll
      for (int j = 0; j <= Q; j++) z[j] = fma(lc.c[j], a, z[j]);
#pragma unroll
      for (int j = 1; j <= Q; j++) z[j] = sc.c[j] * z[j];
    }
    if (w) {
      const double t = fabs(z[0]);
      const double d = av ? fma(rtol, t, av[i]) : rtol * t + atol;
      w[i] = 1.0 / d;
    }
#pragma unroll
    for (int k = 1; k <= Q; k++)
#pragma unroll
      for (int j = Q; j >= k; j--) z[j - 1] = fma(1.0, z[j], z[j - 1]);
#pragma unroll
    for (int j = 0; j < (COMP ? Q + 1 : Q); j++) Z.p[j][i] = z[j];
    const double zero = 0.0;
    acor[i] = zero;
    y[i] = fma(1.0, z[0], 1.0 * zero);
  }
}

__global__ void k_ewt(const double* __restrict__ y, double rtol, double atol,
                      const double* __restrict__ av, double* __restrict__ w, sunindextype n) {
  GRID_STRIDE(i, n) {
    const double t = fabs(y[i]);
    // vector atol: N_VLinearSum(rtol, |y|, 1, atol) = fma(rtol, |y|, atol);
    // scalar atol: N_VScale then N_VAddConst (two roundings)
    const double d = av ? fma(rtol, t, av[i]) : rtol * t + atol;
    w[i] = 1.0 / d;
  }
}

// ------------------------------------------------------------ reductions
enum RedKind { RED_SUM = 0, RED_MAX = 1, RED_MIN = 2 };

__device__ inline double combine(int kind, double a, double b) {
  return kind == RED_SUM ? a + b : (kind == RED_MAX ? fmax(a, b) : fmin(a, b));
}
__device__ inline double identity(int kind) {
  return kind == RED_SUM ? 0.0 : (kind == RED_MAX ? -INFINITY : INFINITY);
}

// element contribution functors
struct FDot { const double *x, *y; __device__ double operator()(sunindextype i) const { return x[i] * y[i]; } };
struct FWsq { const double *x, *w; __device__ double operator()(sunindextype i) const { const double p = x[i] * w[i]; return p * p; } };
struct FWsqMask {
  const double *x, *w, *id;
  __device__ double operator()(sunindextype i) const {
    const double p = x[i] * w[i];
    return id[i] > 0.0 ? p * p : 0.0;
  }
};
struct FAbs { const double* x; __device__ double operator()(sunindextype i) const { return fabs(x[i]); } };
struct FVal { const double* x; __device__ double operator()(sunindextype i) const { return x[i]; } };
struct FQuot {
  const double *a, *b;
  __device__ double operator()(sunindextype i) const { return b[i] != 0.0 ? a[i] / b[i] : INFINITY; }
};
// invtest: z = 1/x where x != 0; contribution 1 when x == 0
struct FInvTest {
  const double* x; double* z;
  __device__ double operator()(sunindextype i) const {
    const double xi = x[i];
    if (xi == 0.0) return 1.0;
    z[i] = 1.0 / xi;
    return 0.0;
  }
};
struct FConstr {
  const double *c, *x; double* m;
  __device__ double operator()(sunindextype i) const {
    const double ci = c[i], xi = x[i];
    double bad = 0.0;
    if (ci != 0.0) {
      const bool b = (fabs(ci) > 1.5 && xi * ci <= 0.0) || (fabs(ci) > 0.5 && xi * ci < 0.0);
      bad = b ? 1.0 : 0.0;
    }
    m[i] = bad;
    return bad;
  }
};

// z = sum_j c_j X_j (written), contribution z_i^2: the linear combination
// and the squared norm of its result in one pass
struct FLincombSq {
  int nv; CoefPack c; PtrPack X; double* z;
  __device__ double operator()(sunindextype i) const {
    const double zi = lincomb_elem(nv, c, X, i);
    z[i] = zi;
    return zi * zi;
  }
};

// Reductions accumulate rounded terms with separate adds, so a fused
// producer that hands the same terms to the same accumulation loop gets
// the same bits.

// z = s .* t (written), contribution z_i * y_i (y = z when self): the
// product and one dot in one pass, in the single-dot accumulation order
struct FProdDot {
  const double *s, *t, *y; double* z; bool self;
  __device__ double operator()(sunindextype i) const {
    const double zi = s[i] * t[i];
    z[i] = zi;
    return zi * (self ? zi : y[i]);
  }
};
// z = X_0 - sum_{j>=1} d_{j-1} X_j with the coefficients d read from device
// memory (a Gram-Schmidt multi-dot's results), contribution z_i^2.  Same
// arithmetic as lincomb_elem with c = {1, -d_0, -d_1, ...}.
struct FGsLincombSq {
  int nv; const double* d; PtrPack X; double* z;
  __device__ double operator()(sunindextype i) const {
    double zi = 1.0 * X.p[0][i];
    for (int j = 1; j < nv; j++) zi = fma(-d[j - 1], X.p[j][i], zi);
    z[i] = zi;
    return zi * zi;
  }
};

// u = a w - sum_{j<nb} b_j U_j (written to out when set), contribution
// u_i^2: the projection norm of a Gram-Schmidt step whose coefficients
// come from the host
struct FProjSq {
  int nb; double a; CoefPack b; const double* w; PtrPack U; double* out;
  __device__ double operator()(sunindextype i) const {
    double u = a * w[i];
    for (int j = 0; j < nb; j++) u = fma(-b.c[j], U.p[j][i], u);
    if (out) out[i] = u;
    return u * u;
  }
};

// first Newton iteration: ycor += del, y = zn0 + ycor, contribution
// (del w)^2 -- the FWsq term of N_VWrmsNorm(del, w)
struct FNewtonUpd0 {
  const double *del, *w, *zn0; double *ycor, *y;
  __device__ double operator()(sunindextype i) const {
    const double d = del[i];
    const double c = fma(1.0, ycor[i], 1.0 * d);
    ycor[i] = c;
    y[i] = fma(1.0, zn0[i], 1.0 * c);
    const double p = d * w[i];
    return p * p;
  }
};

// first Newton iteration with the SPGMR solution formed where read:
// del = (0 + sum_j c_j X_j) / s (as k_lincombdiv), then FNewtonUpd0 (del
// itself is not written)
struct FNewtonUpdComb0 {
  int nv; CoefPack c; PtrPack X; const double *s, *w, *zn0; double *ycor, *y;
  __device__ double operator()(sunindextype i) const {
    double acc = 1.0 * 0.0;
    for (int j = 0; j < nv; j++) acc = fma(c.c[j], X.p[j][i], acc);
    const double d = acc / s[i];
    const double cy = fma(1.0, ycor[i], 1.0 * d);
    ycor[i] = cy;
    y[i] = fma(1.0, zn0[i], 1.0 * cy);
    const double p = d * w[i];
    return p * p;
  }
};

// Block-wide reduction of K values per thread over W waves; result valid
// in thread 0.
template <int K, int W = kWaves>
__device__ void block_reduce(int kind, double (&v)[K], double (*sm)[W]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the K cross-lane chains interleaved step by step, so their shuffle
  // latencies overlap (each value sees the same combines in the same order)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = combine(kind, v[k], __shfl_xor(v[k], off, 64));
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) sm[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      double a = sm[k][0];
      for (int w = 1; w < W; w++) a = combine(kind, a, sm[k][w]);
      v[k] = a;
    }
  }
}

// Final pass: partials[k*nblk + b] for b < nblk (one row of nblk per
// result, so a wave's loads are contiguous).  SAME_KERNEL: run by the last
// workgroup of the kernel that wrote them, so each load is an agent-scope
// atomic that no stale per-CU L1 line can serve; otherwise (a kernel of its
// own) plain loads, which the hardware keeps in flight together.  Same
// order of summation either way.
// sum K rows of nblk partials in slot order; thread 0 ends with the sums
template <int K, bool SAME_KERNEL, int W = kWaves>
__device__ void sum_partials(int kind, int nblk, const double* partials, double (*sm)[W],
                             double (&v)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = identity(kind);
  int b = threadIdx.x;
  if (!SAME_KERNEL) {
    // U blocks' loads issued before any is summed (one memory round trip
    // per U blocks instead of per block)
    constexpr int U = K <= 8 ? 4 : (K <= 12 ? 2 : 1);
    for (; b + (U - 1) * (int)blockDim.x < nblk; b += U * blockDim.x) {
      double p[U][K];
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int k = 0; k < K; k++) p[u][k] = partials[k * nblk + b + u * blockDim.x];
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = combine(kind, v[k], p[u][k]);
    }
  }
  for (; b < nblk; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double p = SAME_KERNEL ? __hip_atomic_load(partials + k * nblk + b,
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                   : partials[k * nblk + b];
      v[k] = combine(kind, v[k], p);
    }
  }
  __syncthreads();
  block_reduce<K, W>(kind, v, sm);
}

template <int K, bool SAME_KERNEL, int W = kWaves>
__device__ void finish(int kind, int nblk, const double* partials, double* out, double* dout,
                       double (*sm)[W], unsigned seq) {
  double v[K];
  sum_partials<K, SAME_KERNEL, W>(kind, nblk, partials, sm, v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = v[k];
    if (dout) {
#pragma unroll
      for (int k = 0; k < K; k++) dout[k] = v[k];
    }
    // results first, then the sequence word the host may spin on
    __threadfence_system();
    __hip_atomic_store(reinterpret_cast<unsigned*>(out + kMaxK), seq, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// the workgroup's share of a one-value reduction of f over [0, n), valid
// in thread 0 (the accumulation every k_reduce1 form uses)
template <class F>
__device__ double reduce1_block(int kind, const F& f, sunindextype n, double (*sm)[kWaves]) {
  // four independent accumulators so four loads per thread are in flight
  double a0 = identity(kind), a1 = a0, a2 = a0, a3 = a0;
  const sunindextype stride = (sunindextype)gridDim.x * blockDim.x;
  sunindextype i = (sunindextype)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const double x0 = f(i), x1 = f(i + stride), x2 = f(i + 2 * stride), x3 = f(i + 3 * stride);
    a0 = combine(kind, a0, x0);
    a1 = combine(kind, a1, x1);
    a2 = combine(kind, a2, x2);
    a3 = combine(kind, a3, x3);
  }
  for (; i < n; i += stride) a0 = combine(kind, a0, f(i));
  double v[1] = {combine(kind, combine(kind, a0, a1), combine(kind, a2, a3))};
  block_reduce<1>(kind, v, sm);
  return v[0];
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_reduce1(int kind, F f, sunindextype n,
                                                    double* partials, unsigned* ticket,
                                                    double* out, double* dout, unsigned seq,
                                                    const int* skip) {
  if (skip && *skip) return;   // a device-resident solve that has finished
  __shared__ double sm[1][kWaves];
  __shared__ bool last;
  double v[1] = {reduce1_block(kind, f, n, sm)};
  if (!ticket) {
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
    return;
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = v[0];
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (last) {
    __threadfence();
    finish<1, true>(kind, gridDim.x, partials, out, dout, sm, seq);
    if (threadIdx.x == 0) *ticket = 0u;
  }
}

// Two reductions in one pass, each accumulated exactly as k_reduce1
// accumulates it (same accumulators, same order), so each result equals
// the single reduction's bit for bit.  Two-launch form only.
template <class F1, class F2>
__global__ __launch_bounds__(kBlock) void k_reduce2(int kind, F1 f1, F2 f2, sunindextype n,
                                                    double* partials) {
  __shared__ double sm[2][kWaves];
  double a0 = identity(kind), a1 = a0, a2 = a0, a3 = a0;
  double b0 = a0, b1 = a0, b2 = a0, b3 = a0;
  const sunindextype stride = (sunindextype)gridDim.x * blockDim.x;
  sunindextype i = (sunindextype)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const double x0 = f1(i), x1 = f1(i + stride), x2 = f1(i + 2 * stride),
                 x3 = f1(i + 3 * stride);
    const double y0 = f2(i), y1 = f2(i + stride), y2 = f2(i + 2 * stride),
                 y3 = f2(i + 3 * stride);
    a0 = combine(kind, a0, x0);
    a1 = combine(kind, a1, x1);
    a2 = combine(kind, a2, x2);
    a3 = combine(kind, a3, x3);
    b0 = combine(kind, b0, y0);
    b1 = combine(kind, b1, y1);
    b2 = combine(kind, b2, y2);
    b3 = combine(kind, b3, y3);
  }
  for (; i < n; i += stride) {
    a0 = combine(kind, a0, f1(i));
    b0 = combine(kind, b0, f2(i));
  }
  double v[2] = {combine(kind, combine(kind, a0, a1), combine(kind, a2, a3)),
                 combine(kind, combine(kind, b0, b1), combine(kind, b2, b3))};
  block_reduce<2>(kind, v, sm);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = v[0];
    partials[gridDim.x + blockIdx.x] = v[1];
  }
}

// (a y + b z) w, squared: the FWsq term of N_VLinearSum(a, y, b, z, t)
// followed by N_VWrmsNorm(t, w), without writing t
// (fma(c, a, x) w)^2: the WRMS term of the completed zn[q] (N_VScaleAddMulti's
// value) before the completion is written
struct FWsqComp {
  double c; const double *x, *a, *w;
  __device__ double operator()(sunindextype i) const {
    const double p = fma(c, a[i], x[i]) * w[i];
    return p * p;
  }
};
struct FLinWsq {
  double a, b; const double *y, *z, *w;
  __device__ double operator()(sunindextype i) const {
    const double t = fma(a, y[i], b * z[i]);
    const double p = t * w[i];
    return p * p;
  }
};

// K dot products x . Y_k (and K weighted squares when W given)
// x_i = (prod ? s_i * t_i : x_i); dots_k = sum_i x_i Y_k[i], where a Y_k
// flagged in `self` is x itself.  With prod set, x is also written (the
// Krylov vector z = s .* t and its Gram-Schmidt dots in one pass).  With
// sidx >= 0, Y_sidx is scaled on read, a = sc * Y_sidx[i], and a written
// back: the normalisation of the newest basis vector done in this pass.
template <int K>
__global__ __launch_bounds__(kBlock) void k_dotmulti(const double* x, PtrPack Y, sunindextype n,
                                                     double* partials, unsigned* ticket,
                                                     double* out, double* dout, unsigned seq,
                                                     const double* sv, const double* tv,
                                                     double* xw, unsigned self, int sidx,
                                                     double sc, const double* scd,
                                                     const int* skip) {
  if (skip && *skip) return;   // a device-resident solve that has finished
  if (scd) sc = *scd;          // the scale of Y_sidx from device memory
  __shared__ double sm[K][kWaves];
  __shared__ bool last;
  double v[K], u[K];
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = u[k] = 0.0;
  const sunindextype stride = (sunindextype)gridDim.x * blockDim.x;
  sunindextype i = (sunindextype)blockIdx.x * blockDim.x + threadIdx.x;
  auto xval = [&](sunindextype j) -> double {
    if (sv) {
      const double a = sv[j] * tv[j];
      xw[j] = a;
      return a;
    }
    return x[j];
  };
  auto yval = [&](int k, sunindextype j, double xj) -> double {
    if ((self >> k) & 1u) return xj;
    if (k == sidx) {
      const double a = sc * Y.p[k][j];
      const_cast<double*>(Y.p[k])[j] = a;
      return a;
    }
    return Y.p[k][j];
  };
  for (; i + stride < n; i += 2 * stride) {
    const double x0 = xval(i), x1 = xval(i + stride);
#pragma unroll
    for (int k = 0; k < K; k++) {
      v[k] += x0 * yval(k, i, x0);
      u[k] += x1 * yval(k, i + stride, x1);
    }
  }
  for (; i < n; i += stride) {
    const double x0 = xval(i);
#pragma unroll
    for (int k = 0; k < K; k++) v[k] += x0 * yval(k, i, x0);
  }
#pragma unroll
  for (int k = 0; k < K; k++) v[k] += u[k];
  block_reduce<K>(RED_SUM, v, sm);
  if (!ticket) {
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < K; k++) partials[k * gridDim.x + blockIdx.x] = v[k];
    }
    return;
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partials[k * gridDim.x + blockIdx.x] = v[k];
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (last) {
    __threadfence();
    finish<K, true>(RED_SUM, gridDim.x, partials, out, dout, sm, seq);
    if (threadIdx.x == 0) *ticket = 0u;
  }
}

// Newton update fused into a 2-vector WRMS pass: X_1 = ycor + X_0 is
// formed (ycor overwritten) and y = zn0 + X_1 written, as the two
// N_VLinearSum calls it replaces would.
struct NewtonUpd {
  double* ycor = nullptr;
  const double* zn0 = nullptr;
  double* y = nullptr;
};

template <int K>
__global__ __launch_bounds__(kBlock) void k_wsqarray(PtrPack X, PtrPack W, const double* id,
                                                     sunindextype n, double* partials,
                                                     unsigned* ticket, double* out,
                                                     double* dout, unsigned seq, NewtonUpd nu) {
  __shared__ double sm[K][kWaves];
  __shared__ bool last;
  double v[K], u[K];
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = u[k] = 0.0;
  const sunindextype stride = (sunindextype)gridDim.x * blockDim.x;
  sunindextype i = (sunindextype)blockIdx.x * blockDim.x + threadIdx.x;
  auto xv = [&](int k, sunindextype j) -> double {
    if (nu.ycor && k == 1) {
      const double c = fma(1.0, nu.ycor[j], 1.0 * X.p[0][j]);
      nu.ycor[j] = c;
      nu.y[j] = fma(1.0, nu.zn0[j], 1.0 * c);
      return c;
    }
    return X.p[k][j];
  };
  for (; i + stride < n; i += 2 * stride) {
    const bool on0 = id ? (id[i] > 0.0) : true;
    const bool on1 = id ? (id[i + stride] > 0.0) : true;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double p0 = xv(k, i) * W.p[k][i];
      const double p1 = xv(k, i + stride) * W.p[k][i + stride];
      v[k] += on0 ? p0 * p0 : 0.0;
      u[k] += on1 ? p1 * p1 : 0.0;
    }
  }
  for (; i < n; i += stride) {
    const bool on = id ? (id[i] > 0.0) : true;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double p = xv(k, i) * W.p[k][i];
      v[k] += on ? p * p : 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) v[k] += u[k];
  block_reduce<K>(RED_SUM, v, sm);
  if (!ticket) {
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < K; k++) partials[k * gridDim.x + blockIdx.x] = v[k];
    }
    return;
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partials[k * gridDim.x + blockIdx.x] = v[k];
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (last) {
    __threadfence();
    finish<K, true>(RED_SUM, gridDim.x, partials, out, dout, sm, seq);
    if (threadIdx.x == 0) *ticket = 0u;
  }
}

// Second launch of the two-launch form: one workgroup sums the nblk
// partials (slot order) and publishes the K results.
// one workgroup of 256 threads (1024 measured slower: 10-16 us against
// 6-10 us for 1024 partials of 4-7 results)
constexpr int kFinishBlock = kBlock;
template <int K>
__global__ __launch_bounds__(kFinishBlock) void k_finish(int kind, int nblk,
                                                         const double* partials, double* out,
                                                         double* dout, unsigned seq,
                                                         const int* skip) {
  if (skip && *skip) return;
  __shared__ double sm[K][kFinishBlock / 64];
  finish<K, false, kFinishBlock / 64>(kind, nblk, partials, out, dout, sm, seq);
}

// dout: the device copy of the results (slot 1: the workspace's dres)
template <int K>
void launch_finish(hipStream_t s, int kind, int nblk, const RedWs* w, unsigned seq,
                   int slot = 0, const int* skip = nullptr, double* dout = nullptr) {
  hipLaunchKernelGGL(k_finish<K>, dim3(1), dim3(kFinishBlock), 0, s, kind, nblk, w->partials,
                     w->host_dev + slot * kSlot, dout ? dout : (slot ? w->dres : nullptr), seq,
                     skip);
}
void finish_dispatch(int K, hipStream_t s, int kind, int nblk, const RedWs* w, unsigned seq) {
  switch (K) {
    case 1: launch_finish<1>(s, kind, nblk, w, seq); break;
    case 2: launch_finish<2>(s, kind, nblk, w, seq); break;
    case 3: launch_finish<3>(s, kind, nblk, w, seq); break;
    case 4: launch_finish<4>(s, kind, nblk, w, seq); break;
    case 5: launch_finish<5>(s, kind, nblk, w, seq); break;
    case 6: launch_finish<6>(s, kind, nblk, w, seq); break;
    case 7: launch_finish<7>(s, kind, nblk, w, seq); break;
    case 8: launch_finish<8>(s, kind, nblk, w, seq); break;
    case 9: launch_finish<9>(s, kind, nblk, w, seq); break;
    case 10: launch_finish<10>(s, kind, nblk, w, seq); break;
    case 11: launch_finish<11>(s, kind, nblk, w, seq); break;
    case 12: launch_finish<12>(s, kind, nblk, w, seq); break;
    case 13: launch_finish<13>(s, kind, nblk, w, seq); break;
    case 14: launch_finish<14>(s, kind, nblk, w, seq); break;
    case 15: launch_finish<15>(s, kind, nblk, w, seq); break;
    default: launch_finish<16>(s, kind, nblk, w, seq); break;
  }
}

// Wait for the reduction tagged `seq` on stream s (result slot `slot`).
// Spin mode polls the pinned sequence word (bounded: after 2 s it falls
// back to a stream sync, which also surfaces any device error; a later
// sequence number in the slot ends the wait too); otherwise
// hipStreamSynchronize.
void wait_result(hipStream_t s, RedWs* w, unsigned seq, int slot = 0) {
  g_syncs++;
  if (g_sync_mode == 1 || g_sync_mode == 2) {
    const unsigned* word = reinterpret_cast<const unsigned*>(w->host + slot * kSlot + kMaxK);
    auto t0 = std::chrono::steady_clock::now();
    for (unsigned it = 0;; it++) {
      const unsigned got = __atomic_load_n(word, __ATOMIC_ACQUIRE);
      if (got == seq) return;
      if ((int)(got - seq) > 0) break;
      if (g_sync_mode == 2 && (it & 15u) == 15u) std::this_thread::yield();
      if ((it & 4095u) == 4095u &&
          std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2))
        break;
    }
  }
  hipStreamSynchronize(s);
}

// One thread waits until the host releases gate `expected` (or `limit`
// ticks of the 100 MHz wall clock pass: then it stops); then the lanes copy
// the coefficients to device memory, one each, so the reads of host memory
// overlap, and the skip flag that the kernels queued behind the gate test
// is set.  Every path ends the kernel.
__global__ __launch_bounds__(64) void k_gate(const GateBox* box, unsigned expected,
                                              double* dcoef, int* skip,
                                              unsigned long long limit) {
  __shared__ int stop_s;
  if (threadIdx.x == 0) {
    const unsigned long long t0 = wall_clock64();
    int stop = 1;
    for (;;) {
      const unsigned q =
          __hip_atomic_load(&box->seq, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
      if (q == expected) {
        stop = __hip_atomic_load(&box->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        break;
      }
      if (wall_clock64() - t0 > limit) break;
      __builtin_amdgcn_s_sleep(2);
    }
    stop_s = stop;
  }
  __syncthreads();
  const int stop = stop_s;
  if (!stop && threadIdx.x < kMaxK)
    dcoef[threadIdx.x] =
        __hip_atomic_load(&box->coef[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  if (threadIdx.x == 0) *skip = stop;
}

// wait_result that reports whether the result arrived (a finish launch
// skipped behind a stopped gate never publishes)
bool wait_result_checked(hipStream_t s, RedWs* w, unsigned seq) {
  wait_result(s, w, seq);
  const unsigned* word = reinterpret_cast<const unsigned*>(w->host + kMaxK);
  return __atomic_load_n(word, __ATOMIC_ACQUIRE) == seq;
}

template <class F>
double reduce1(N_Vector v, int kind, F f) {
  const sunindextype n = LEN(v);
  hipStream_t s = S(v);
  RedWs* w = workspace(s);
  const int g = red_grid(n);
  const unsigned seq = ++w->seq;
  hipLaunchKernelGGL(k_reduce1<F>, dim3(g), dim3(kBlock), 0, s, kind, f, n, w->partials,
                     g_red_mode == 0 ? w->ticket : nullptr, w->host_dev, (double*)nullptr, seq,
                     (const int*)nullptr);
  if (g_red_mode != 0) finish_dispatch(1, s, kind, g, w, seq);
  check_launch("reduce");
  wait_result(s, w, seq);
  return w->host[0];
}

struct ProdIn {
  const double* s = nullptr;   // when set, x = s .* t is formed and written to xw
  const double* t = nullptr;
  double* xw = nullptr;
  unsigned self = 0;           // bit k: Y_k is x itself
  int slot = 0;                // result slot (1 also fills the device copy)
  int sidx = -1;               // Y_sidx scaled by sc on read and written back
  double sc = 1.0;
  const double* scd = nullptr; // when set, sc is read from device memory
  const int* skip = nullptr;   // when set and nonzero on the device, no work
};

template <int K>
void launch_dotmulti(hipStream_t s, const double* x, const PtrPack& Y, sunindextype n,
                     RedWs* w, const ProdIn& pi) {
  const int g = red_grid(n);
  hipLaunchKernelGGL(k_dotmulti<K>, dim3(g), dim3(kBlock), 0, s, x, Y, n,
                     w->partials, g_red_mode == 0 ? w->ticket : nullptr,
                     w->host_dev + pi.slot * kSlot, pi.slot ? w->dres : nullptr, w->seq,
                     pi.s, pi.t, pi.xw, pi.self, pi.sidx, pi.sc, pi.scd, pi.skip);
  if (g_red_mode != 0) launch_finish<K>(s, RED_SUM, g, w, w->seq, pi.slot, pi.skip);
}

void dotmulti_dispatch(int K, hipStream_t s, const double* x, const PtrPack& Y, sunindextype n,
                       RedWs* w, const ProdIn& pi = ProdIn{}) {
  switch (K) {
    case 1: launch_dotmulti<1>(s, x, Y, n, w, pi); break;
    case 2: launch_dotmulti<2>(s, x, Y, n, w, pi); break;
    case 3: launch_dotmulti<3>(s, x, Y, n, w, pi); break;
    case 4: launch_dotmulti<4>(s, x, Y, n, w, pi); break;
    case 5: launch_dotmulti<5>(s, x, Y, n, w, pi); break;
    case 6: launch_dotmulti<6>(s, x, Y, n, w, pi); break;
    case 7: launch_dotmulti<7>(s, x, Y, n, w, pi); break;
    case 8: launch_dotmulti<8>(s, x, Y, n, w, pi); break;
    case 9: launch_dotmulti<9>(s, x, Y, n, w, pi); break;
    case 10: launch_dotmulti<10>(s, x, Y, n, w, pi); break;
    case 11: launch_dotmulti<11>(s, x, Y, n, w, pi); break;
    case 12: launch_dotmulti<12>(s, x, Y, n, w, pi); break;
    case 13: launch_dotmulti<13>(s, x, Y, n, w, pi); break;
    case 14: launch_dotmulti<14>(s, x, Y, n, w, pi); break;
    case 15: launch_dotmulti<15>(s, x, Y, n, w, pi); break;
    default: launch_dotmulti<16>(s, x, Y, n, w, pi); break;
  }
}

template <int K>
void launch_wsq(hipStream_t s, const PtrPack& X, const PtrPack& W, const double* id,
                sunindextype n, RedWs* w, const NewtonUpd& nu = NewtonUpd{}) {
  const int g = red_grid(n);
  hipLaunchKernelGGL(k_wsqarray<K>, dim3(g), dim3(kBlock), 0, s, X, W, id, n,
                     w->partials, g_red_mode == 0 ? w->ticket : nullptr, w->host_dev,
                     (double*)nullptr, w->seq, nu);
  if (g_red_mode != 0) launch_finish<K>(s, RED_SUM, g, w, w->seq);
}
void wsq_dispatch(int K, hipStream_t s, const PtrPack& X, const PtrPack& W, const double* id,
                  sunindextype n, RedWs* w) {
  switch (K) {
    case 1: launch_wsq<1>(s, X, W, id, n, w); break;
    case 2: launch_wsq<2>(s, X, W, id, n, w); break;
    case 3: launch_wsq<3>(s, X, W, id, n, w); break;
    case 4: launch_wsq<4>(s, X, W, id, n, w); break;
    case 5: launch_wsq<5>(s, X, W, id, n, w); break;
    case 6: launch_wsq<6>(s, X, W, id, n, w); break;
    case 7: launch_wsq<7>(s, X, W, id, n, w); break;
    case 8: launch_wsq<8>(s, X, W, id, n, w); break;
    default: launch_wsq<16>(s, X, W, id, n, w); break;
  }
}

void install_ops(N_Vector v, bool fused);

N_Vector make_vector(sunindextype n, hipStream_t s, bool alloc, SUNContext sunctx) {
  N_Vector v = (N_Vector)calloc(1, sizeof(*v));
  if (!v) return nullptr;
  v->ops = (N_Vector_Ops)calloc(1, sizeof(struct _generic_N_Vector_Ops));
  NvContent* c = new NvContent();
  c->n = n;
  c->stream = s;
  if (alloc && n > 0) {
    if (hipMalloc(&c->d, sizeof(double) * (size_t)n) != hipSuccess) {
      delete c;
      free(v->ops);
      free(v);
      return nullptr;
    }
    c->own_d = true;
  }
  v->content = c;
  v->sunctx = sunctx;
  install_ops(v, false);
  return v;
}

}  // namespace

// ----------------------------------------------- allocation-free multi-dot
// Device scratch of N_VDotProdMultiNoAlloc_Hip: the per-workgroup partial
// sums of at most kNoAllocBlocks workgroups for up to kMaxK products, then
// one result slot (kMaxK results and the sequence word).  A file-scope
// device array: every GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` is only a symbol; its device address
// comes from hipGetSymbolAddress, once per device (noalloc_scratch).
constexpr int kNoAllocBlocks = 1024;
constexpr int kNoAllocPartials = kNoAllocBlocks * kMaxK;
constexpr int kNoAllocDoubles = kNoAllocPartials + kSlot;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
double* g_na_ptr[kMaxDevices] = {};
std::atomic<unsigned> g_na_seq{0};

// the current device's address of `scratch` (looked up on its first use)
double* noalloc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "[nvector_hip] hipGetSymbolAddress(scratch) on device %d: %s\n", dev,
              hipGetErrorString(e));
      return nullptr;
    }
    g_na_ptr[dev] = static_cast<double*>(p);
  }
  return g_na_ptr[dev];
}

template <int K>
void launch_dotmulti_noalloc(hipStream_t s, const double* x, const PtrPack& Y, sunindextype n,
                             int g, double* ws, unsigned seq) {
  double* out = ws + kNoAllocPartials;
  hipLaunchKernelGGL(k_dotmulti<K>, dim3(g), dim3(kBlock), 0, s, x, Y, n, ws, (unsigned*)nullptr,
                     out, (double*)nullptr, seq, (const double*)nullptr, (const double*)nullptr,
                     (double*)nullptr, 0u, -1, 1.0, (const double*)nullptr, (const int*)nullptr);
  hipLaunchKernelGGL(k_finish<K>, dim3(1), dim3(kFinishBlock), 0, s, (int)RED_SUM, g,
                     (const double*)ws, out, (double*)nullptr, seq, (const int*)nullptr);
}
}  // namespace

// =================================================================== C API
extern "C" {

void* NVHip_DefaultStream(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_default_stream_made) {
    if (hipStreamCreateWithFlags(&g_default_stream, hipStreamNonBlocking) != hipSuccess)
      g_default_stream = nullptr;
    g_default_stream_made = true;
  }
  return (void*)g_default_stream;
}

long NVHip_NumSyncs(void) { return g_syncs.load(); }

int NVHip_SetSyncMode(int mode) {
  if (mode < 0 || mode > 2) return -1;
  g_sync_mode = mode;
  return 0;
}

int NVHip_SetReduceConfig(int max_blocks, int mode) {
  if (max_blocks < 1 || max_blocks > kMaxRedBlocks || (mode != 0 && mode != 1)) return -1;
  g_red_blocks = max_blocks;
  g_red_mode = mode;
  return 0;
}

N_Vector N_VNewWithStream_Hip(sunindextype length, void* stream, SUNContext sunctx) {
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)NVHip_DefaultStream();
  return make_vector(length, s, true, sunctx);
}
N_Vector N_VNew_Hip(sunindextype length, SUNContext sunctx) {
  return N_VNewWithStream_Hip(length, nullptr, sunctx);
}
N_Vector N_VNewEmpty_Hip(SUNContext sunctx) {
  return make_vector(0, (hipStream_t)NVHip_DefaultStream(), false, sunctx);
}
N_Vector N_VMake_Hip(sunindextype length, sunrealtype* d_data, void* stream, SUNContext sunctx) {
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)NVHip_DefaultStream();
  N_Vector v = make_vector(length, s, false, sunctx);
  if (v) C(v)->d = d_data;
  return v;
}

N_Vector_ID N_VGetVectorID_Hip(N_Vector v) { (void)v; return SUNDIALS_NVEC_HIP; }

N_Vector N_VCloneEmpty_Hip(N_Vector w) {
  N_Vector v = make_vector(LEN(w), S(w), false, w->sunctx);
  if (!v) return nullptr;
  std::memcpy(v->ops, w->ops, sizeof(struct _generic_N_Vector_Ops));
  return v;
}
N_Vector N_VClone_Hip(N_Vector w) {
  N_Vector v = make_vector(LEN(w), S(w), true, w->sunctx);
  if (!v) return nullptr;
  std::memcpy(v->ops, w->ops, sizeof(struct _generic_N_Vector_Ops));
  return v;
}
void N_VDestroy_Hip(N_Vector v) {
  if (!v) return;
  NvContent* c = C(v);
  if (c) {
    if (c->own_d && c->d) hipFree(c->d);
    if (c->h) hipHostFree(c->h);
    delete c;
  }
  free(v->ops);
  free(v);
}
void N_VSpace_Hip(N_Vector v, sunindextype* lrw, sunindextype* liw) {
  *lrw = LEN(v);
  *liw = 2;
}
sunrealtype* N_VGetDeviceArrayPointer_Hip(N_Vector v) { return D(v); }
sunrealtype* N_VGetHostArrayPointer_Hip(N_Vector v) {
  NvContent* c = C(v);
  if (!c->h && c->n > 0) {
    if (hipHostMalloc(&c->h, sizeof(double) * (size_t)c->n, hipHostMallocDefault) != hipSuccess)
      c->h = nullptr;
  }
  return c->h;
}
static void nv_setarray(sunrealtype* d, N_Vector v) { C(v)->d = d; }
sunindextype N_VGetLength_Hip(N_Vector v) { return LEN(v); }
void* N_VGetStream_Hip(N_Vector v) { return (void*)S(v); }

void N_VCopyToDevice_Hip(N_Vector v) {
  double* h = N_VGetHostArrayPointer_Hip(v);
  if (!h) return;
  hipMemcpyAsync(D(v), h, sizeof(double) * (size_t)LEN(v), hipMemcpyHostToDevice, S(v));
  hipStreamSynchronize(S(v));
}
void N_VCopyFromDevice_Hip(N_Vector v) {
  double* h = N_VGetHostArrayPointer_Hip(v);
  if (!h) return;
  hipMemcpyAsync(h, D(v), sizeof(double) * (size_t)LEN(v), hipMemcpyDeviceToHost, S(v));
  hipStreamSynchronize(S(v));
}
int N_VSetFromHost_Hip(N_Vector v, const sunrealtype* src) {
  if (hipMemcpyAsync(D(v), src, sizeof(double) * (size_t)LEN(v), hipMemcpyHostToDevice, S(v)) !=
      hipSuccess)
    return -1;
  return hipStreamSynchronize(S(v)) == hipSuccess ? 0 : -1;
}
int N_VGetToHost_Hip(N_Vector v, sunrealtype* dst) {
  if (hipMemcpyAsync(dst, D(v), sizeof(double) * (size_t)LEN(v), hipMemcpyDeviceToHost, S(v)) !=
      hipSuccess)
    return -1;
  return hipStreamSynchronize(S(v)) == hipSuccess ? 0 : -1;
}

#define EW_LAUNCH(kern, z, ...)                                                        \
  do {                                                                                 \
    const sunindextype n_ = LEN(z);                                                    \
    if (n_ > 0) hipLaunchKernelGGL(kern, dim3(ew_grid(n_)), dim3(kBlock), 0, S(z), __VA_ARGS__); \
    check_launch(#kern);                                                               \
  } while (0)

void N_VLinearSum_Hip(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z) {
  EW_LAUNCH(k_linearsum, z, a, D(x), b, D(y), D(z), LEN(z));
}
void N_VConst_Hip(sunrealtype c, N_Vector z) { EW_LAUNCH(k_const, z, c, D(z), LEN(z)); }
void N_VProd_Hip(N_Vector x, N_Vector y, N_Vector z) { EW_LAUNCH(k_prod, z, D(x), D(y), D(z), LEN(z)); }
void N_VDiv_Hip(N_Vector x, N_Vector y, N_Vector z) { EW_LAUNCH(k_div, z, D(x), D(y), D(z), LEN(z)); }
void N_VScale_Hip(sunrealtype c, N_Vector x, N_Vector z) {
  if (x == z && c == 1.0) return;
  EW_LAUNCH(k_scale, z, c, D(x), D(z), LEN(z));
}
void N_VAbs_Hip(N_Vector x, N_Vector z) { EW_LAUNCH(k_abs, z, D(x), D(z), LEN(z)); }
void N_VInv_Hip(N_Vector x, N_Vector z) { EW_LAUNCH(k_inv, z, D(x), D(z), LEN(z)); }
void N_VAddConst_Hip(N_Vector x, sunrealtype b, N_Vector z) {
  EW_LAUNCH(k_addconst, z, D(x), b, D(z), LEN(z));
}
void N_VCompare_Hip(sunrealtype c, N_Vector x, N_Vector z) {
  EW_LAUNCH(k_compare, z, c, D(x), D(z), LEN(z));
}

sunrealtype N_VDotProd_Hip(N_Vector x, N_Vector y) {
  return reduce1(x, RED_SUM, FDot{D(x), D(y)});
}
sunrealtype N_VMaxNorm_Hip(N_Vector x) { return reduce1(x, RED_MAX, FAbs{D(x)}); }
sunrealtype N_VWSqrSumLocal_Hip(N_Vector x, N_Vector w) {
  return reduce1(x, RED_SUM, FWsq{D(x), D(w)});
}
sunrealtype N_VWSqrSumMaskLocal_Hip(N_Vector x, N_Vector w, N_Vector id) {
  return reduce1(x, RED_SUM, FWsqMask{D(x), D(w), D(id)});
}
sunrealtype N_VWrmsNorm_Hip(N_Vector x, N_Vector w) {
  return std::sqrt(N_VWSqrSumLocal_Hip(x, w) / (double)LEN(x));
}
sunrealtype N_VWrmsNormMask_Hip(N_Vector x, N_Vector w, N_Vector id) {
  return std::sqrt(N_VWSqrSumMaskLocal_Hip(x, w, id) / (double)LEN(x));
}
sunrealtype N_VMin_Hip(N_Vector x) { return reduce1(x, RED_MIN, FVal{D(x)}); }
sunrealtype N_VWL2Norm_Hip(N_Vector x, N_Vector w) {
  return std::sqrt(N_VWSqrSumLocal_Hip(x, w));
}
sunrealtype N_VL1Norm_Hip(N_Vector x) { return reduce1(x, RED_SUM, FAbs{D(x)}); }
sunbooleantype N_VInvTest_Hip(N_Vector x, N_Vector z) {
  const double nz = reduce1(z, RED_SUM, FInvTest{D(x), D(z)});
  return nz > 0.0 ? SUNFALSE : SUNTRUE;
}
sunbooleantype N_VConstrMask_Hip(N_Vector c, N_Vector x, N_Vector m) {
  const double nb = reduce1(m, RED_SUM, FConstr{D(c), D(x), D(m)});
  return nb > 0.0 ? SUNFALSE : SUNTRUE;
}
sunrealtype N_VMinQuotient_Hip(N_Vector num, N_Vector denom) {
  const double q = reduce1(num, RED_MIN, FQuot{D(num), D(denom)});
  return std::isinf(q) ? 1.7976931348623157e308 : q;
}

SUNErrCode N_VLinearCombination_Hip(int nvec, sunrealtype* c, N_Vector* X, N_Vector z) {
  if (nvec < 1) return SUN_SUCCESS;
  if (nvec == 1) { N_VScale_Hip(c[0], X[0], z); return SUN_SUCCESS; }
  // groups of kMaxK: first group writes z, later groups accumulate into z
  int done = 0;
  while (done < nvec) {
    CoefPack cp{};
    PtrPack xp{};
    int k = 0;
    if (done > 0) { cp.c[0] = 1.0; xp.p[0] = D(z); k = 1; }
    while (k < kMaxK && done < nvec) { cp.c[k] = c[done]; xp.p[k] = D(X[done]); k++; done++; }
    EW_LAUNCH(k_lincomb, z, k, cp, xp, D(z), LEN(z));
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VScaleAddMulti_Hip(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y, N_Vector* Z) {
  for (int base = 0; base < nvec; base += kMaxK) {
    const int k = (nvec - base < kMaxK) ? nvec - base : kMaxK;
    CoefPack cp{};
    PtrPack yp{};
    OutPack zp{};
    for (int j = 0; j < k; j++) {
      cp.c[j] = a[base + j];
      yp.p[j] = D(Y[base + j]);
      zp.p[j] = D(Z[base + j]);
    }
    EW_LAUNCH(k_scaleaddmulti, x, k, cp, D(x), yp, zp, LEN(x));
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VDotProdMulti_Hip(int nvec, N_Vector x, N_Vector* Y, sunrealtype* dotprods) {
  if (nvec < 1) return SUN_SUCCESS;
  if (nvec == 1) {   // one dot: the single-dot kernel (as SUNDIALS' generic layer)
    dotprods[0] = N_VDotProd_Hip(x, Y[0]);
    return SUN_SUCCESS;
  }
  hipStream_t s = S(x);
  RedWs* w = workspace(s);
  for (int base = 0; base < nvec; base += kMaxK) {
    const int k = (nvec - base < kMaxK) ? nvec - base : kMaxK;
    PtrPack yp{};
    for (int j = 0; j < k; j++) yp.p[j] = D(Y[base + j]);
    ++w->seq;
    dotmulti_dispatch(k, s, D(x), yp, LEN(x), w);
    check_launch("dotmulti");
    wait_result(s, w, w->seq);
    for (int j = 0; j < k; j++) dotprods[base + j] = w->host[j];
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VDotProdMultiNoAlloc_Hip(int nvec, N_Vector x, N_Vector* Y, sunrealtype* dotprods) {
  if (nvec < 1) return SUN_SUCCESS;
  if (nvec > kMaxK || !x || !Y || !dotprods) {
    fprintf(stderr, "[nvector_hip] N_VDotProdMultiNoAlloc_Hip: nvec = %d outside 1..%d or a "
                    "NULL argument (the device scratch holds %d products)\n",
            nvec, kMaxK, kMaxK);
    return -1;
  }
  const sunindextype n = LEN(x);
  PtrPack yp{};
  for (int j = 0; j < nvec; j++) {
    if (!Y[j] || LEN(Y[j]) != n) {
      fprintf(stderr, "[nvector_hip] N_VDotProdMultiNoAlloc_Hip: vector %d missing or of "
                      "another length\n", j);
      return -1;
    }
    yp.p[j] = D(Y[j]);
  }
  double* ws = noalloc_scratch();
  if (!ws) return -1;
  // the same launch shape as N_VDotProdMulti_Hip up to kNoAllocBlocks
  // workgroups (the library default), hence the same bits; every partial
  // the finish reads is written by this call, so nothing is cleared
  int g = red_grid(n);
  if (g > kNoAllocBlocks) g = kNoAllocBlocks;
  const unsigned seq = ++g_na_seq;
  hipStream_t s = S(x);
  switch (nvec) {
    case 1: launch_dotmulti_noalloc<1>(s, D(x), yp, n, g, ws, seq); break;
    case 2: launch_dotmulti_noalloc<2>(s, D(x), yp, n, g, ws, seq); break;
    case 3: launch_dotmulti_noalloc<3>(s, D(x), yp, n, g, ws, seq); break;
    case 4: launch_dotmulti_noalloc<4>(s, D(x), yp, n, g, ws, seq); break;
    case 5: launch_dotmulti_noalloc<5>(s, D(x), yp, n, g, ws, seq); break;
    case 6: launch_dotmulti_noalloc<6>(s, D(x), yp, n, g, ws, seq); break;
    case 7: launch_dotmulti_noalloc<7>(s, D(x), yp, n, g, ws, seq); break;
    case 8: launch_dotmulti_noalloc<8>(s, D(x), yp, n, g, ws, seq); break;
    case 9: launch_dotmulti_noalloc<9>(s, D(x), yp, n, g, ws, seq); break;
    case 10: launch_dotmulti_noalloc<10>(s, D(x), yp, n, g, ws, seq); break;
    case 11: launch_dotmulti_noalloc<11>(s, D(x), yp, n, g, ws, seq); break;
    case 12: launch_dotmulti_noalloc<12>(s, D(x), yp, n, g, ws, seq); break;
    case 13: launch_dotmulti_noalloc<13>(s, D(x), yp, n, g, ws, seq); break;
    case 14: launch_dotmulti_noalloc<14>(s, D(x), yp, n, g, ws, seq); break;
    case 15: launch_dotmulti_noalloc<15>(s, D(x), yp, n, g, ws, seq); break;
    default: launch_dotmulti_noalloc<16>(s, D(x), yp, n, g, ws, seq); break;
  }
  if (hipPeekAtLastError() != hipSuccess) {
    check_launch("dotmulti (no-alloc)");
    return -1;
  }
  g_syncs++;
  if (hipMemcpyAsync(dotprods, ws + kNoAllocPartials, sizeof(double) * nvec,
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1;
  return SUN_SUCCESS;
}

SUNErrCode N_VLinearSumVectorArray_Hip(int nvec, sunrealtype a, N_Vector* X, sunrealtype b,
                                       N_Vector* Y, N_Vector* Z) {
  for (int base = 0; base < nvec; base += kMaxK) {
    const int k = (nvec - base < kMaxK) ? nvec - base : kMaxK;
    PtrPack xp{}, yp{};
    OutPack zp{};
    for (int j = 0; j < k; j++) {
      xp.p[j] = D(X[base + j]);
      yp.p[j] = D(Y[base + j]);
      zp.p[j] = D(Z[base + j]);
    }
    EW_LAUNCH(k_lsumarray, Z[0], k, a, xp, b, yp, zp, LEN(Z[0]));
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VScaleVectorArray_Hip(int nvec, sunrealtype* c, N_Vector* X, N_Vector* Z) {
  for (int base = 0; base < nvec; base += kMaxK) {
    const int k = (nvec - base < kMaxK) ? nvec - base : kMaxK;
    CoefPack cp{};
    PtrPack xp{};
    OutPack zp{};
    for (int j = 0; j < k; j++) {
      cp.c[j] = c[base + j];
      xp.p[j] = D(X[base + j]);
      zp.p[j] = D(Z[base + j]);
    }
    EW_LAUNCH(k_scalearray, Z[0], k, cp, xp, zp, LEN(Z[0]));
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VConstVectorArray_Hip(int nvec, sunrealtype c, N_Vector* Z) {
  for (int base = 0; base < nvec; base += kMaxK) {
    const int k = (nvec - base < kMaxK) ? nvec - base : kMaxK;
    OutPack zp{};
    for (int j = 0; j < k; j++) zp.p[j] = D(Z[base + j]);
    EW_LAUNCH(k_constarray, Z[0], k, c, zp, LEN(Z[0]));
  }
  return SUN_SUCCESS;
}

static SUNErrCode wsq_array(int nvec, N_Vector* X, N_Vector* W, N_Vector id, sunrealtype* nrm) {
  if (nvec < 1) return SUN_SUCCESS;
  hipStream_t s = S(X[0]);
  RedWs* w = workspace(s);
  const sunindextype n = LEN(X[0]);
  for (int base = 0; base < nvec; base += 8) {
    const int k = (nvec - base < 8) ? nvec - base : 8;
    PtrPack xp{}, wp{};
    for (int j = 0; j < k; j++) {
      xp.p[j] = D(X[base + j]);
      wp.p[j] = D(W[base + j]);
    }
    ++w->seq;
    wsq_dispatch(k, s, xp, wp, id ? D(id) : nullptr, n, w);
    check_launch("wsqarray");
    wait_result(s, w, w->seq);
    for (int j = 0; j < k; j++) nrm[base + j] = std::sqrt(w->host[j] / (double)n);
  }
  return SUN_SUCCESS;
}

SUNErrCode N_VWrmsNormVectorArray_Hip(int nvec, N_Vector* X, N_Vector* W, sunrealtype* nrm) {
  return wsq_array(nvec, X, W, nullptr, nrm);
}
SUNErrCode N_VWrmsNormMaskVectorArray_Hip(int nvec, N_Vector* X, N_Vector* W, N_Vector id,
                                          sunrealtype* nrm) {
  return wsq_array(nvec, X, W, id, nrm);
}

static sunrealtype nv_dot_local(N_Vector x, N_Vector y) { return N_VDotProd_Hip(x, y); }
static SUNComm nv_comm(N_Vector v) { (void)v; return SUN_COMM_NULL; }
static void nv_print(N_Vector v) {
  double* h = N_VGetHostArrayPointer_Hip(v);
  N_VCopyFromDevice_Hip(v);
  for (sunindextype i = 0; i < LEN(v) && h; i++) printf("%.16e\n", h[i]);
}
static void nv_printfile(N_Vector v, FILE* f) {
  double* h = N_VGetHostArrayPointer_Hip(v);
  N_VCopyFromDevice_Hip(v);
  for (sunindextype i = 0; i < LEN(v) && h; i++) fprintf(f, "%.16e\n", h[i]);
}

int N_VEnableFusedOps_Hip(N_Vector v, sunbooleantype tf) {
  if (!v || !v->ops) return -1;
  N_Vector_Ops o = v->ops;
  o->nvlinearcombination = tf ? N_VLinearCombination_Hip : nullptr;
  o->nvscaleaddmulti = tf ? N_VScaleAddMulti_Hip : nullptr;
  o->nvdotprodmulti = tf ? N_VDotProdMulti_Hip : nullptr;
  o->nvlinearsumvectorarray = tf ? N_VLinearSumVectorArray_Hip : nullptr;
  o->nvscalevectorarray = tf ? N_VScaleVectorArray_Hip : nullptr;
  o->nvconstvectorarray = tf ? N_VConstVectorArray_Hip : nullptr;
  o->nvwrmsnormvectorarray = tf ? N_VWrmsNormVectorArray_Hip : nullptr;
  o->nvwrmsnormmaskvectorarray = tf ? N_VWrmsNormMaskVectorArray_Hip : nullptr;
  return 0;
}

int NVHip_Pascal(int q, N_Vector* zn, sunrealtype sign) {
  if (q < 1 || q > 5) return -1;
  for (int j = 0; j <= q; j++)
    if (!zn[j] || zn[j]->ops->nvgetvectorid != N_VGetVectorID_Hip) return -1;
  OutPack zp{};
  for (int j = 0; j <= q; j++) zp.p[j] = D(zn[j]);
  const sunindextype n = LEN(zn[0]);
  hipStream_t s = S(zn[0]);
  switch (q) {
    case 1: hipLaunchKernelGGL(k_pascal<1>, dim3(ew_grid(n)), dim3(kBlock), 0, s, zp, sign, n); break;
    case 2: hipLaunchKernelGGL(k_pascal<2>, dim3(ew_grid(n)), dim3(kBlock), 0, s, zp, sign, n); break;
    case 3: hipLaunchKernelGGL(k_pascal<3>, dim3(ew_grid(n)), dim3(kBlock), 0, s, zp, sign, n); break;
    case 4: hipLaunchKernelGGL(k_pascal<4>, dim3(ew_grid(n)), dim3(kBlock), 0, s, zp, sign, n); break;
    default: hipLaunchKernelGGL(k_pascal<5>, dim3(ew_grid(n)), dim3(kBlock), 0, s, zp, sign, n); break;
  }
  check_launch("pascal");
  return 0;
}

static int predict_launch(int q, N_Vector* zn, N_Vector acor, const double* lc,
                          const double* sc, N_Vector y, double rtol, double atol,
                          N_Vector abstol, N_Vector ewt, N_Vector save = nullptr) {
  if (q < 1 || q > 5 || !acor || !y) return -1;
  if (save && (!lc || save->ops->nvgetvectorid != N_VGetVectorID_Hip ||
               LEN(save) != LEN(zn[0]) || D(save) == D(acor) || D(save) == D(y) ||
               (ewt && D(save) == D(ewt))))
    return -1;
  if (save)
    for (int j = 0; j <= q; j++)
      if (zn[j] && D(zn[j]) == D(save)) return -1;
  if (ewt && (ewt->ops->nvgetvectorid != N_VGetVectorID_Hip || LEN(ewt) != LEN(zn[0])))
    return -1;
  if (abstol && LEN(abstol) != LEN(zn[0])) return -1;
  for (int j = 0; j <= q; j++)
    if (!zn[j] || zn[j]->ops->nvgetvectorid != N_VGetVectorID_Hip) return -1;
  if (acor->ops->nvgetvectorid != N_VGetVectorID_Hip ||
      y->ops->nvgetvectorid != N_VGetVectorID_Hip)
    return -1;
  const sunindextype n = LEN(zn[0]);
  if (LEN(acor) != n || LEN(y) != n) return -1;
  double* dw = ewt ? D(ewt) : nullptr;
  for (int j = 0; j <= q; j++)
    if (D(zn[j]) == D(acor) || D(zn[j]) == D(y) || D(zn[j]) == dw) return -1;
  if (D(acor) == D(y) || D(acor) == dw || D(y) == dw) return -1;
  OutPack zp{};
  for (int j = 0; j <= q; j++) zp.p[j] = D(zn[j]);
  hipStream_t s = S(zn[0]);
  const dim3 g(ew_grid(n)), b(kBlock);
  const double* av = abstol ? D(abstol) : nullptr;
  CoefPack lp{}, sp{};
  for (int j = 0; j <= q; j++) {
    lp.c[j] = lc ? lc[j] : 0.0;
    sp.c[j] = (sc && j > 0) ? sc[j] : 1.0;
  }
  double* const dsave = save ? D(save) : nullptr;
  switch (q * 2 + (lc ? 1 : 0)) {
#define NVHIP_PI_CASE(k)                                                                      \
    case 2 * k: hipLaunchKernelGGL((k_pascal_init<k, false>), g, b, 0, s, zp, D(acor), D(y), n, \
                                   rtol, atol, av, dw, lp, sp, (double*)nullptr); break;      \
    case 2 * k + 1: hipLaunchKernelGGL((k_pascal_init<k, true>), g, b, 0, s, zp, D(acor), D(y), \
                                       n, rtol, atol, av, dw, lp, sp, dsave); break;
    NVHIP_PI_CASE(1) NVHIP_PI_CASE(2) NVHIP_PI_CASE(3) NVHIP_PI_CASE(4) NVHIP_PI_CASE(5)
#undef NVHIP_PI_CASE
    default: return -1;
  }
  check_launch("pascal_init");
  return 0;
}

int NVHip_PredictInit(int q, N_Vector* zn, N_Vector acor, N_Vector y, sunrealtype rtol,
                      sunrealtype atol, N_Vector abstol, N_Vector ewt) {
  return predict_launch(q, zn, acor, nullptr, nullptr, y, rtol, atol, abstol, ewt);
}

int NVHip_PredictComplete(int q, N_Vector* zn, N_Vector acor, const sunrealtype* l,
                          const sunrealtype* c, N_Vector y, sunrealtype rtol, sunrealtype atol,
                          N_Vector abstol, N_Vector ewt) {
  if (!l) return -1;
  return predict_launch(q, zn, acor, l, c, y, rtol, atol, abstol, ewt);
}

int NVHip_PredictCompleteSave(int q, N_Vector* zn, N_Vector acor, const sunrealtype* l,
                              const sunrealtype* c, N_Vector y, sunrealtype rtol,
                              sunrealtype atol, N_Vector abstol, N_Vector ewt, N_Vector save) {
  if (!l || !save) return -1;
  return predict_launch(q, zn, acor, l, c, y, rtol, atol, abstol, ewt, save);
}

// x_comp: nrm[0] is the norm of fma(cx, x_comp, x) (x + cx x_comp as
// N_VScaleAddMulti forms it) when given
// the two sums of a WRMS pair queued on x's stream, results in `slot`
static int wrms_pair_queue(N_Vector x, double cx, N_Vector xc, double a, N_Vector y, double b,
                           N_Vector z, N_Vector w, int slot, unsigned* seq_out) {
  if (!x || !y || !z || !w) return -1;
  N_Vector vs[5] = {x, y, z, w, xc ? xc : x};
  for (N_Vector v : vs)
    if (v->ops->nvgetvectorid != N_VGetVectorID_Hip || LEN(v) != LEN(x)) return -1;
  const sunindextype n = LEN(x);
  hipStream_t s = S(x);
  RedWs* ws = workspace(s);
  const int g = red_grid(n);
  const unsigned seq = ++ws->seq;
  if (xc)
    hipLaunchKernelGGL((k_reduce2<FWsqComp, FLinWsq>), dim3(g), dim3(kBlock), 0, s, RED_SUM,
                       FWsqComp{cx, D(x), D(xc), D(w)}, FLinWsq{a, b, D(y), D(z), D(w)}, n,
                       ws->partials);
  else
    hipLaunchKernelGGL((k_reduce2<FWsq, FLinWsq>), dim3(g), dim3(kBlock), 0, s, RED_SUM,
                       FWsq{D(x), D(w)}, FLinWsq{a, b, D(y), D(z), D(w)}, n, ws->partials);
  launch_finish<2>(s, RED_SUM, g, ws, seq, slot);
  check_launch("wrms pair");
  *seq_out = seq;
  return 0;
}

static int wrms_pair(N_Vector x, double cx, N_Vector xc, double a, N_Vector y, double b,
                     N_Vector z, N_Vector w, double* nrm) {
  unsigned seq = 0;
  if (!nrm || wrms_pair_queue(x, cx, xc, a, y, b, z, w, 0, &seq) != 0) return -1;
  const sunindextype n = LEN(x);
  RedWs* ws = workspace(S(x));
  wait_result(S(x), ws, seq);
  nrm[0] = std::sqrt(ws->host[0] / (double)n);
  nrm[1] = std::sqrt(ws->host[1] / (double)n);
  return 0;
}

int NVHip_WrmsNormPair(N_Vector x, sunrealtype a, N_Vector y, sunrealtype b, N_Vector z,
                       N_Vector w, sunrealtype* nrm) {
  return wrms_pair(x, 0.0, nullptr, a, y, b, z, w, nrm);
}

int NVHip_WrmsNormPairComp(N_Vector x, sunrealtype cx, N_Vector xc, sunrealtype a, N_Vector y,
                           sunrealtype b, N_Vector z, N_Vector w, sunrealtype* nrm) {
  if (!xc) return -1;
  return wrms_pair(x, cx, xc, a, y, b, z, w, nrm);
}

int NVHip_WrmsNormPairCompQueue(N_Vector x, sunrealtype cx, N_Vector xc, sunrealtype a,
                                N_Vector y, sunrealtype b, N_Vector z, N_Vector w,
                                unsigned* ticket) {
  if (!xc || !ticket || g_red_mode == 0) return -1;
  return wrms_pair_queue(x, cx, xc, a, y, b, z, w, 1, ticket);
}

int NVHip_WrmsNormPairWait(N_Vector x, unsigned ticket, sunrealtype* nrm) {
  if (!x || !nrm || x->ops->nvgetvectorid != N_VGetVectorID_Hip) return -1;
  hipStream_t s = S(x);
  RedWs* ws = workspace(s);
  wait_result(s, ws, ticket, 1);
  const unsigned* word = reinterpret_cast<const unsigned*>(ws->host + kSlot + kMaxK);
  if (__atomic_load_n(word, __ATOMIC_ACQUIRE) != ticket) return -1;
  const double n = (double)LEN(x);
  nrm[0] = std::sqrt(ws->host[kSlot] / n);
  nrm[1] = std::sqrt(ws->host[kSlot + 1] / n);
  return 0;
}

int NVHip_EwtSet(N_Vector y, sunrealtype rtol, sunrealtype atol, N_Vector abstol, N_Vector w) {
  if (y->ops->nvgetvectorid != N_VGetVectorID_Hip) return -1;
  EW_LAUNCH(k_ewt, w, D(y), rtol, atol, abstol ? (const double*)D(abstol) : nullptr, D(w), LEN(w));
  return 0;
}

int NVHip_ScaleDiv(sunrealtype c, N_Vector x, N_Vector sv, N_Vector xout, N_Vector vout) {
  if (!x || !sv || !xout || !vout || LEN(x) != LEN(sv) || LEN(x) != LEN(xout) ||
      LEN(x) != LEN(vout) || D(sv) == D(vout) || D(x) == D(vout))
    return -1;
  EW_LAUNCH(k_scalediv, x, c, D(x), D(sv), D(xout), D(vout), LEN(x));
  return 0;
}

int NVHip_ProdDotMulti(N_Vector sv, N_Vector t, N_Vector z, int nvec, N_Vector* Y,
                       sunrealtype* dots) {
  if (!sv || !t || !z || nvec < 1 || nvec > kMaxK || LEN(sv) != LEN(z) || LEN(t) != LEN(z))
    return -1;
  PtrPack yp{};
  ProdIn pi;
  pi.s = D(sv);
  pi.t = D(t);
  pi.xw = D(z);
  for (int j = 0; j < nvec; j++) {
    if (!Y[j] || LEN(Y[j]) != LEN(z)) return -1;
    yp.p[j] = D(Y[j]);
    if (D(Y[j]) == D(z)) pi.self |= 1u << j;
  }
  if (nvec == 1) {
    dots[0] = reduce1(z, RED_SUM, FProdDot{pi.s, pi.t, yp.p[0], pi.xw, pi.self != 0});
    return 0;
  }
  hipStream_t s = S(z);
  RedWs* w = workspace(s);
  ++w->seq;
  dotmulti_dispatch(nvec, s, D(z), yp, LEN(z), w, pi);
  check_launch("proddotmulti");
  wait_result(s, w, w->seq);
  for (int j = 0; j < nvec; j++) dots[j] = w->host[j];
  return 0;
}

int NVHip_LinCombDiv(int nvec, sunrealtype* c, N_Vector* X, N_Vector sv, N_Vector x) {
  if (nvec < 1 || nvec > kMaxK || !sv || !x || LEN(sv) != LEN(x)) return -1;
  CoefPack cp{};
  PtrPack xp{};
  for (int j = 0; j < nvec; j++) {
    if (!X[j] || LEN(X[j]) != LEN(x)) return -1;
    cp.c[j] = c[j];
    xp.p[j] = D(X[j]);
  }
  EW_LAUNCH(k_lincombdiv, x, nvec, cp, xp, D(sv), D(x), LEN(x));
  return 0;
}

int NVHip_NewtonUpdate(int iter, N_Vector ycor, N_Vector del, N_Vector zn0, N_Vector w,
                       N_Vector y, sunrealtype* dnorm, sunrealtype* ynorm) {
  if (!ycor || !del || !zn0 || !w || !y || !dnorm) return -1;
  const sunindextype n = LEN(ycor);
  if (LEN(del) != n || LEN(zn0) != n || LEN(w) != n || LEN(y) != n) return -1;
  if (D(y) == D(ycor) || D(y) == D(del) || D(y) == D(zn0) || D(ycor) == D(del)) return -1;
  if (iter == 0) {
    const double sq = reduce1(ycor, RED_SUM, FNewtonUpd0{D(del), D(w), D(zn0), D(ycor), D(y)});
    *dnorm = std::sqrt(sq / (double)n);
    return 0;
  }
  if (!ynorm) return -1;
  hipStream_t s = S(ycor);
  RedWs* wk = workspace(s);
  PtrPack xp{}, wp{};
  xp.p[0] = D(del);
  xp.p[1] = D(ycor);
  wp.p[0] = wp.p[1] = D(w);
  NewtonUpd nu;
  nu.ycor = D(ycor);
  nu.zn0 = D(zn0);
  nu.y = D(y);
  ++wk->seq;
  launch_wsq<2>(s, xp, wp, nullptr, n, wk, nu);
  check_launch("newtonupdate");
  wait_result(s, wk, wk->seq);
  *dnorm = std::sqrt(wk->host[0] / (double)n);
  *ynorm = std::sqrt(wk->host[1] / (double)n);
  return 0;
}

int NVHip_NewtonUpdateComb(N_Vector ycor, int nvec, const sunrealtype* c, N_Vector* X,
                           N_Vector sv, N_Vector zn0, N_Vector w, N_Vector y,
                           sunrealtype* dnorm) {
  if (!ycor || !c || !X || !sv || !zn0 || !w || !y || !dnorm || nvec < 1 || nvec > kMaxK)
    return -1;
  const sunindextype n = LEN(ycor);
  if (LEN(sv) != n || LEN(zn0) != n || LEN(w) != n || LEN(y) != n) return -1;
  if (D(y) == D(ycor) || D(y) == D(zn0) || D(y) == D(sv) || D(y) == D(w)) return -1;
  FNewtonUpdComb0 f{};
  f.nv = nvec;
  for (int j = 0; j < nvec; j++) {
    if (!X[j] || LEN(X[j]) != n || D(X[j]) == D(y) || D(X[j]) == D(ycor)) return -1;
    f.c.c[j] = c[j];
    f.X.p[j] = D(X[j]);
  }
  f.s = D(sv);
  f.w = D(w);
  f.zn0 = D(zn0);
  f.ycor = D(ycor);
  f.y = D(y);
  const double sq = reduce1(ycor, RED_SUM, f);
  *dnorm = std::sqrt(sq / (double)n);
  return 0;
}

}  // extern "C"

namespace {
// Gram-Schmidt projection of z against Y (Y[nvec-1] == z): dots, then
// z -= sum d_j Y_j with its squared norm, one host wait.  With sv set,
// z = sv .* t is formed first; with sidx >= 0, Y_sidx is scaled by sc
// (and written back) as the dots read it.
int gs_project(N_Vector sv, N_Vector t, N_Vector z, int nvec, N_Vector* Y, int sidx, double sc,
               sunrealtype* dots, sunrealtype* zz) {
  PtrPack yp{};
  ProdIn pi;
  pi.s = sv ? D(sv) : nullptr;
  pi.t = sv ? D(t) : nullptr;
  pi.xw = sv ? D(z) : nullptr;
  pi.slot = 1;
  pi.sidx = sidx;
  pi.sc = sc;
  FGsLincombSq f{};
  f.nv = nvec;
  f.z = D(z);
  f.X.p[0] = D(z);
  for (int j = 0; j < nvec; j++) {
    if (!Y[j] || LEN(Y[j]) != LEN(z)) return -1;
    yp.p[j] = D(Y[j]);
    if (D(Y[j]) == D(z)) pi.self |= 1u << j;
    if (j < nvec - 1) {
      if (D(Y[j]) == D(z)) return -1;
      f.X.p[j + 1] = D(Y[j]);
    }
  }
  const sunindextype n = LEN(z);
  hipStream_t s = S(z);
  RedWs* w = workspace(s);
  // 1) z = s .* t and its dots with Y (slot 1 + device copy), no host wait
  ++w->seq;
  dotmulti_dispatch(nvec, s, D(z), yp, n, w, pi);
  const unsigned seq_dots = w->seq;
  f.d = w->dres;
  // 2) z -= sum d_j Y_j and <z, z> (slot 0); the one wait covers both
  const int g = red_grid(n);
  const unsigned seq = ++w->seq;
  hipLaunchKernelGGL(k_reduce1<FGsLincombSq>, dim3(g), dim3(kBlock), 0, s, RED_SUM, f, n,
                     w->partials, g_red_mode == 0 ? w->ticket : nullptr, w->host_dev,
                     (double*)nullptr, seq, (const int*)nullptr);
  if (g_red_mode != 0) finish_dispatch(1, s, RED_SUM, g, w, seq);
  check_launch("gsproject");
  wait_result(s, w, seq);
  const unsigned got = *reinterpret_cast<volatile unsigned*>(w->host + kSlot + kMaxK);
  if (got != seq_dots) hipStreamSynchronize(s);
  for (int j = 0; j < nvec; j++) dots[j] = w->host[kSlot + j];
  *zz = w->host[0];
  return 0;
}
}  // namespace

extern "C" {

int NVHip_GSProject(N_Vector sv, N_Vector t, N_Vector z, int nvec, N_Vector* Y,
                    sunrealtype* dots, sunrealtype* zz) {
  if (!sv || !t || !z || nvec < 2 || nvec > kMaxK || !dots || !zz || Y[nvec - 1] != z ||
      LEN(sv) != LEN(z) || LEN(t) != LEN(z))
    return -1;
  return gs_project(sv, t, z, nvec, Y, -1, 1.0, dots, zz);
}

}  // extern "C"

namespace {

// ------------------------------------------------- device-resident SPGMR
// The left-preconditioned SPGMR solve of SUNLinSolSolve_SPGMR (zero initial
// guess, classical Gram-Schmidt, no restarts) with every scalar kept on the
// device: the Hessenberg column, its Givens rotation, the residual estimate
// and the convergence decision are made by one-thread kernels between the
// vector passes, and every kernel of an iteration after convergence exits
// at once (a device flag).  The host waits once, at the end.  Each scalar
// is computed with the same operations in the same order as the host
// integrator computes it (no contraction in this file; FP64 division and
// square root are correctly rounded on the device as on the host), so the
// solve is bit-identical to the host-driven one.
constexpr double kSpGsFactor = 1000.0;   // classical GS re-orthogonalisation test
enum SpStatus { SP_RUNNING = 0, SP_ZERO = 1, SP_CONVERGED = 2, SP_REORTH = 3, SP_QRFACT = 4,
                SP_QRSOL = 5, SP_REDUCED = 6, SP_FAILED = 7 };
struct SpDev {
  double H[kMaxK][kMaxK];   // Hessenberg matrix, rows 0..maxl, columns 0..maxl-1
  double q[2 * kMaxK];      // Givens rotations (c, s)
  double yg[kMaxK + 1];     // least-squares right-hand side / solution
  double zz, r_norm, beta, rho, rotprod, vscale, delta;
  int done, status, krydim, write;
};

__global__ void k_sp_init(SpDev* st, const double* rr, double delta) {
  if (threadIdx.x != 0) return;
  const double r_norm = sqrt(rr[0]);
  st->r_norm = r_norm;
  st->beta = r_norm;
  st->rho = r_norm;
  st->delta = delta;
  st->rotprod = 1.0;
  st->krydim = 0;
  st->write = 0;
  st->zz = 0.0;
  for (int i = 0; i < kMaxK; i++)
    for (int j = 0; j < kMaxK; j++) st->H[i][j] = 0.0;
  if (r_norm <= delta) {
    st->done = 1;
    st->status = SP_ZERO;
  } else {
    st->done = 0;
    st->status = SP_RUNNING;
    st->vscale = 1.0 / r_norm;
  }
}

__device__ inline void sp_rotation(double t1, double t2, double& c, double& s) {
  if (t2 == 0.0) {
    c = 1.0;
    s = 0.0;
  } else if (fabs(t2) >= fabs(t1)) {
    const double t3 = t1 / t2;
    s = -1.0 / sqrt(1.0 + t3 * t3);
    c = -s * t3;
  } else {
    const double t3 = t2 / t1;
    c = 1.0 / sqrt(1.0 + t3 * t3);
    s = -c * t3;
  }
}

// iteration l: dots[0..l+1] = <V[l+1], V[0..l+1]> before the projection,
// st->zz = |V[l+1]|^2 after it
__global__ void k_sp_step(SpDev* st, int l, const double* __restrict__ dots) {
  if (threadIdx.x != 0 || st->done) return;
  const int k = l + 1;
  const double vk_norm = sqrt(dots[k]);
  for (int i = k - 1; i >= 0; i--) st->H[i][k - 1] = dots[i];
  const double new_norm = sqrt(st->zz);
  if ((kSpGsFactor * new_norm) < vk_norm) {   // the host path re-orthogonalises
    st->done = 1;
    st->status = SP_REORTH;
    return;
  }
  st->H[l + 1][l] = new_norm;
  st->krydim = k;
  // new Hessenberg column: previous rotations, then its own
  for (int j = 0; j < l; j++) {
    const double t1 = st->H[j][l], t2 = st->H[j + 1][l], c = st->q[2 * j], s = st->q[2 * j + 1];
    st->H[j][l] = c * t1 - s * t2;
    st->H[j + 1][l] = s * t1 + c * t2;
  }
  double c, s;
  const double t1 = st->H[l][l], t2 = st->H[l + 1][l];
  sp_rotation(t1, t2, c, s);
  st->q[2 * l] = c;
  st->q[2 * l + 1] = s;
  if ((st->H[l][l] = c * t1 - s * t2) == 0.0) {
    st->done = 1;
    st->status = SP_QRFACT;
    return;
  }
  st->rotprod *= st->q[2 * l + 1];
  st->rho = fabs(st->rotprod * st->r_norm);
  if (st->rho <= st->delta) {
    st->done = 1;
    st->status = SP_CONVERGED;
    return;
  }
  st->vscale = 1.0 / st->H[l + 1][l];
}

// least-squares solution and the decision to write x (write 1: the basis
// combination, 2: zero, 0: untouched)
__global__ void k_sp_final(SpDev* st, int maxl) {
  if (threadIdx.x != 0) return;
  const int status = st->status;
  if (status == SP_ZERO) { st->write = 2; return; }
  if (status == SP_REORTH || status == SP_QRFACT) { st->write = 0; return; }
  const bool conv = status == SP_CONVERGED;
  const int n = conv ? st->krydim : maxl;
  if (!conv) st->krydim = maxl;
  double* b = st->yg;
  b[0] = st->r_norm;
  for (int i = 1; i <= n; i++) b[i] = 0.0;
  for (int k = 0; k < n; k++) {
    const double c = st->q[2 * k], s = st->q[2 * k + 1], t1 = b[k], t2 = b[k + 1];
    b[k] = c * t1 - s * t2;
    b[k + 1] = s * t1 + c * t2;
  }
  for (int k = n - 1; k >= 0; k--) {
    if (st->H[k][k] == 0.0) { st->status = SP_QRSOL; st->write = 0; return; }
    b[k] /= st->H[k][k];
    for (int i = 0; i < k; i++) b[i] -= b[k] * st->H[i][k];
  }
  if (conv || st->rho < st->beta) {
    st->write = 1;
    if (!conv) st->status = SP_REDUCED;
  } else {
    st->write = 0;
    st->status = SP_FAILED;
  }
}

// x = (0 + sum_k yg_k V_k) / s as the host's N_VLinearCombination into a
// zeroed vector and N_VDiv (k_lincombdiv) round it; or x = 0
__global__ void k_sp_update(const SpDev* __restrict__ st, PtrPack V,
                            const double* __restrict__ sv, double* x, sunindextype n) {
  const int w = st->write;
  if (w == 0) return;
  if (w == 2) {
    GRID_STRIDE(i, n) x[i] = 0.0;
    return;
  }
  const int nv = st->krydim;
  GRID_STRIDE(i, n) {
    double acc = 1.0 * 0.0;
    for (int j = 0; j < nv; j++) acc = fma(st->yg[j], V.p[j][i], acc);
    x[i] = acc / sv[i];
  }
}

__global__ void k_sp_publish(const SpDev* st, double* out, unsigned seq) {
  if (threadIdx.x != 0) return;
  out[0] = (double)st->status;
  out[1] = (double)st->krydim;
  out[2] = st->rho;
  out[3] = st->r_norm;
  __threadfence_system();
  __hip_atomic_store(reinterpret_cast<unsigned*>(out + kMaxK), seq, __ATOMIC_RELEASE,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace

extern "C" {

int NVHip_SpgmrSolve(NVHipDevOp op, void* op_data, N_Vector t0, N_Vector* V, int maxl,
                     N_Vector sv, N_Vector x, sunrealtype delta, int* status, int* iters,
                     sunrealtype* r_norm, sunrealtype* resnorm) {
  if (!op || !t0 || !V || !sv || !x || !status || !iters || maxl < 1 || maxl > kMaxK - 1 ||
      g_red_mode == 0 || V[0] != t0)
    return -1;
  const sunindextype n = LEN(x);
  for (int i = 0; i <= maxl; i++)
    if (!V[i] || LEN(V[i]) != n) return -1;
  if (LEN(sv) != n) return -1;
  hipStream_t s = S(x);
  RedWs* w = workspace(s);
  if (!w->sp && hipMalloc(&w->sp, sizeof(SpDev)) != hipSuccess) {
    w->sp = nullptr;
    return -1;
  }
  SpDev* st = static_cast<SpDev*>(w->sp);
  const int* skip = &st->done;
  const int g = red_grid(n);
  // V[0] = s .* t0 and its squared norm into dres[0] (the single-dot pass
  // of NVHip_ProdDotMulti, so the same summation order)
  {
    const unsigned seq0 = ++w->seq;
    const FProdDot f{D(sv), D(t0), D(V[0]), D(V[0]), true};
    hipLaunchKernelGGL(k_reduce1<FProdDot>, dim3(g), dim3(kBlock), 0, s, RED_SUM, f, n,
                       w->partials, (unsigned*)nullptr, w->host_dev, (double*)nullptr, seq0,
                       (const int*)nullptr);
    launch_finish<1>(s, RED_SUM, g, w, seq0, 0, nullptr, w->dres);
  }
  hipLaunchKernelGGL(k_sp_init, dim3(1), dim3(64), 0, s, st, (const double*)w->dres, delta);
  for (int l = 0; l < maxl; l++) {
    const int k = l + 1;
    if (op(&st->vscale, skip, V[l], sv, V[k], op_data) != 0) return -1;
    // dots of V[k] with V[0..k], V[l] normalised by vscale as it is read
    PtrPack yp{};
    for (int j = 0; j <= k; j++) yp.p[j] = D(V[j]);
    ProdIn pi;
    pi.self = 1u << k;
    pi.slot = 1;
    pi.sidx = l;
    pi.scd = &st->vscale;
    pi.skip = skip;
    ++w->seq;
    dotmulti_dispatch(k + 1, s, D(V[k]), yp, n, w, pi);
    // V[k] -= sum_{j<k} dots_j V[j] and its squared norm into st->zz
    FGsLincombSq f{};
    f.nv = k + 1;
    f.d = w->dres;
    f.z = D(V[k]);
    f.X.p[0] = D(V[k]);
    for (int j = 0; j < k; j++) f.X.p[j + 1] = D(V[j]);
    const unsigned seq = ++w->seq;
    hipLaunchKernelGGL(k_reduce1<FGsLincombSq>, dim3(g), dim3(kBlock), 0, s, RED_SUM, f, n,
                       w->partials, (unsigned*)nullptr, w->host_dev, (double*)nullptr, seq,
                       skip);
    launch_finish<1>(s, RED_SUM, g, w, seq, 0, skip, &st->zz);
    hipLaunchKernelGGL(k_sp_step, dim3(1), dim3(64), 0, s, st, l, (const double*)w->dres);
  }
  hipLaunchKernelGGL(k_sp_final, dim3(1), dim3(64), 0, s, st, maxl);
  {
    PtrPack vp{};
    for (int j = 0; j < maxl; j++) vp.p[j] = D(V[j]);
    hipLaunchKernelGGL(k_sp_update, dim3(ew_grid(n)), dim3(kBlock), 0, s, (const SpDev*)st, vp,
                       (const double*)D(sv), D(x), n);
  }
  const unsigned seq = ++w->seq;
  hipLaunchKernelGGL(k_sp_publish, dim3(1), dim3(64), 0, s, (const SpDev*)st, w->host_dev, seq);
  check_launch("spgmr");
  wait_result(s, w, seq);
  const int code = (int)w->host[0];
  *iters = (int)w->host[1];
  if (r_norm) *r_norm = w->host[3];
  if (resnorm) *resnorm = code == SP_ZERO ? w->host[3] : w->host[2];
  switch (code) {
    case SP_ZERO:
    case SP_CONVERGED: *status = NVHIP_SPGMR_SUCCESS; break;
    case SP_REDUCED: *status = NVHIP_SPGMR_RES_REDUCED; break;
    case SP_FAILED: *status = NVHIP_SPGMR_CONV_FAIL; break;
    case SP_QRFACT: *status = NVHIP_SPGMR_QRFACT_FAIL; break;
    case SP_QRSOL: *status = NVHIP_SPGMR_QRSOL_FAIL; break;
    default: *status = NVHIP_SPGMR_FALLBACK; break;   // re-orthogonalisation needed
  }
  if (code == SP_ZERO) *iters = 0;
  return 0;
}

}  // extern "C"

namespace {

// -------------------------------------- device-driven lagged SPGMR
// The scalar step of the integrator's SPGMR with lagged normalisation (its
// host loop in cvode_bdf.cpp, spgmr_lagged), on thread 0 of the kernel
// that sums a pass's partials, with the same operations in the same order
// (no contraction in this file; FP64 division and square root correctly
// rounded), so the Hessenberg matrix, rotations and decisions are the host
// loop's bit for bit.  After pass l the sums are r[0] = |u_l|^2,
// r[1 + j] = <w_l, u_j> (j <= l), r[2 + l] = |w_l|^2 (and for pass 0 the
// right-hand side's r[3] = |b s|^2).
constexpr double kLagGsFactor = 1000.0;   // classical GS re-orthogonalisation test
constexpr double kLagEstFloor = 1e-4;     // the Pythagorean estimate's floor
constexpr int kLagMax = 8;
enum LagRun { LAG_RUNNING = -1 };

struct LagDev {
  double H[kMaxK][kMaxK];
  double q[2 * kMaxK];
  double nu[kMaxK];
  double coef[kMaxK + 1];   // the next pass's a, b_0 .. b_l
  double r_norm, rotprod, rho, vk;
  double xc[kMaxK];         // the solution's coefficients c_k = y_k / |u_k|
  int done, pskip, status, krydim, iters, passes;
  int xk;                   // basis vectors in the solution
  int xready;               // the solution is formed (status and qr_sol as the host's exit)
  int xskip;                // the queued solution update found none (NVHip_LagUpdate)
  int xpad;
};
struct LagPub {
  NVHipLagResult r;
  unsigned seq;
};

__device__ inline double lag_givens_sine(double t1, double t2) {
  if (t2 == 0.0) return 0.0;
  if (fabs(t2) >= fabs(t1)) {
    const double t3 = t1 / t2;
    return -1.0 / sqrt(1.0 + t3 * t3);
  }
  const double t3 = t2 / t1;
  return -(1.0 / sqrt(1.0 + t3 * t3)) * t3;
}

__device__ inline void lag_end(LagDev* st, int status) {
  st->status = status;
  st->done = 1;
}

// The step functions take the column index as a template argument: every
// index into the state is then a constant (LDS accesses at immediate
// offsets, the column and the sums in registers), which is what keeps the
// serial step of one lane short.

// column L complete with h_{L+1,L} = nu_next: re-orthogonalisation test,
// its Givens rotation (qr_fact of the new column), residual estimate and
// convergence; true when the solve has ended
template <int L>
__device__ bool lag_finalize(LagDev* st, double nu_next, int maxl, double delta) {
  if ((kLagGsFactor * nu_next) < st->vk) {
    lag_end(st, NVHIP_LAG_FALLBACK);
    return true;
  }
  st->nu[L + 1] = nu_next;
  double h[L + 2];
#pragma unroll
  for (int k = 0; k <= L; k++) h[k] = st->H[k][L];
  h[L + 1] = nu_next;
#pragma unroll
  for (int k = 0; k < L; k++) {
    const double t1 = h[k], t2 = h[k + 1], c = st->q[2 * k], s = st->q[2 * k + 1];
    h[k] = c * t1 - s * t2;
    h[k + 1] = s * t1 + c * t2;
  }
  double c, s;
  const double t1 = h[L], t2 = h[L + 1];
  sp_rotation(t1, t2, c, s);
  st->q[2 * L] = c;
  st->q[2 * L + 1] = s;
  h[L] = c * t1 - s * t2;
#pragma unroll
  for (int k = 0; k <= L + 1; k++) st->H[k][L] = h[k];
  if (h[L] == 0.0) {
    lag_end(st, NVHIP_LAG_QRFACT);
    return true;
  }
  st->rotprod *= s;
  st->rho = fabs(st->rotprod * st->r_norm);
  if (st->rho <= delta) {
    lag_end(st, NVHIP_LAG_CONVERGED);
    return true;
  }
  if (L + 1 == maxl) {
    lag_end(st, NVHIP_LAG_MAXL);
    return true;
  }
  return false;
}

// column L from pass L's sums r: Hessenberg entries, the next pass's
// coefficients, |A V_L| and the Pythagorean estimate of h_{L+1,L}; a column
// the estimate shows converged (or the last one, when the estimate is well
// conditioned) is completed here; true when the solve has ended
template <int L>
__device__ bool lag_column(LagDev* st, const double* r, int maxl, double delta) {
  st->iters++;
  st->krydim = L + 1;
  const double inv = 1.0 / st->nu[L];
  double h[L + 1];
  double hh = 0.0;
#pragma unroll
  for (int j = 0; j <= L; j++) {
    h[j] = (r[1 + j] * inv) / st->nu[j];
    hh += h[j] * h[j];
    st->H[j][L] = h[j];
    st->coef[1 + j] = h[j] / st->nu[j];
  }
  st->coef[0] = inv;
  const double vk_norm = sqrt(r[2 + L]) * inv;
  st->vk = vk_norm;
  const double est = vk_norm * vk_norm - hh;
  bool use_est = false;
  if (est > kLagEstFloor * vk_norm * vk_norm) {
    if (L + 1 == maxl) {
      use_est = true;
    } else {
      // the trial rotation on a copy of the column
#pragma unroll
      for (int k = 0; k < L; k++) {
        const double c = st->q[2 * k], sn = st->q[2 * k + 1];
        const double t1 = h[k], t2 = h[k + 1];
        h[k] = c * t1 - sn * t2;
        h[k + 1] = sn * t1 + c * t2;
      }
      const double sn = lag_givens_sine(h[L], sqrt(est));
      use_est = fabs(st->rotprod * sn * st->r_norm) <= delta;
    }
  }
  if (use_est) {
    // the trial rotation is the one the completion forms, so an estimate
    // that shows convergence converges; anything else goes to the host
    if (!lag_finalize<L>(st, sqrt(est), maxl, delta)) lag_end(st, NVHIP_LAG_FALLBACK);
    return true;
  }
  if (L + 1 == maxl) st->pskip = 0;   // the last column's projection norm
  return false;
}

// a solve that has ended with n = krydim <= NMAX columns: the coefficients
// of its solution x = sum_k c_k u_k / s, formed exactly as the host's exit
// forms them (the same statuses, qr_sol -- the rotations applied to the
// right-hand side, then back substitution -- and residual test; c_k =
// y_k / |u_k|)
template <int NMAX>
__device__ void lag_solution(LagDev* st) {
  st->xready = 0;
  if (st->status != NVHIP_LAG_CONVERGED && st->status != NVHIP_LAG_MAXL) return;
  const int n = st->krydim;
  if (n > NMAX) {   // not reached: pass P ends a solve of at most P + 1 columns
    st->status = NVHIP_LAG_FALLBACK;
    return;
  }
  double yg[NMAX + 1];
  yg[0] = st->r_norm;
#pragma unroll
  for (int i = 1; i <= NMAX; i++) yg[i] = 0.0;
#pragma unroll
  for (int k = 0; k < NMAX; k++) {
    if (k < n) {
      const double c = st->q[2 * k], s = st->q[2 * k + 1], t1 = yg[k], t2 = yg[k + 1];
      yg[k] = c * t1 - s * t2;
      yg[k + 1] = s * t1 + c * t2;
    }
  }
#pragma unroll
  for (int k = NMAX - 1; k >= 0; k--) {
    if (k < n) {
      if (st->H[k][k] == 0.0) return;
      yg[k] /= st->H[k][k];
#pragma unroll
      for (int i = 0; i < k; i++) yg[i] -= yg[k] * st->H[i][k];
    }
  }
  if (!(st->status == NVHIP_LAG_CONVERGED || st->rho < st->r_norm)) return;
#pragma unroll
  for (int k = 0; k < NMAX; k++)
    if (k < n) st->xc[k] = yg[k] / st->nu[k];
  st->xk = n;
  st->xready = 1;
}

// the result for the host, written by every thread of the workgroup (one
// store per value), each thread's stores fenced at system scope, then the
// sequence word after the barrier
__device__ void lag_publish(const LagDev* st, LagPub* pub, unsigned seq, int maxl) {
  NVHipLagResult& o = pub->r;
  const int t = threadIdx.x;
  if (t == 0) {
    o.status = st->status;
    o.krydim = st->krydim;
    o.iters = st->iters;
    o.passes = st->passes;
    o.r_norm = st->r_norm;
    o.rho = st->rho;
  }
  if (t <= maxl) o.nu[t] = st->nu[t];
  if (t < 2 * maxl) o.q[t] = st->q[t];
  for (int e = t; e < (maxl + 1) * maxl; e += blockDim.x) {
    const int i = e / maxl, j = e - i * maxl;
    o.H[i][j] = st->H[i][j];
  }
  __threadfence_system();
  __syncthreads();
  if (t == 0) __hip_atomic_store(&pub->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// one step of the solve from pass P's sums r; true when the solve has ended
// (the state is then to be published)
template <int P>
__device__ bool lag_run(LagDev* st, const double* r, int maxl, double delta, double bfloor,
                        double nlen) {
  if constexpr (P == 0) {
    st->done = 0;
    st->pskip = 1;
    st->xready = 0;
    st->status = LAG_RUNNING;
    st->iters = 0;
    st->passes = 0;
    st->krydim = 0;
    st->rotprod = 1.0;
    st->r_norm = 0.0;
    st->rho = 0.0;
    if (bfloor >= 0.0 && sqrt(r[3] / nlen) <= bfloor) {
      lag_end(st, NVHIP_LAG_BNORM);
      return true;
    }
    const double r_norm = sqrt(r[0]);
    st->r_norm = r_norm;
    st->rho = r_norm;
    if (r_norm <= delta) {
      lag_end(st, NVHIP_LAG_ZERO);
      return true;
    }
    // H was zeroed when the state was staged
    st->nu[0] = r_norm;
    return lag_column<0>(st, r, maxl, delta);
  } else {
    st->passes++;
    return lag_finalize<P - 1>(st, sqrt(r[0]), maxl, delta) || lag_column<P>(st, r, maxl, delta);
  }
}


// PROJ false: the sums of pass P (P = 0 starts the solve); PROJ true: the
// last column's projection norm |u|^2 (one sum; P = maxl - 1).  last: no
// further step of the solve is queued yet (the state is published also
// when the solve goes on, with status LAG_RUNNING).  The state is staged in
// LDS by all threads while the partials are summed (its Hessenberg matrix
// zeroed for a new solve), stepped there by thread 0 (its dependent
// accesses then cost LDS latency, not memory round trips) and copied back.
// A solve that has ended (a pass queued past convergence) or a projection
// norm that is not needed leaves everything as it is: its flag, read with
// the state, is tested before the partials are summed.
template <int P, bool PROJ>
__global__ __launch_bounds__(kFinishBlock) void k_lag_step(int nblk, const double* partials,
                                                           int last, LagDev* gst, LagPub* pub,
                                                           unsigned seq, int maxl, double delta,
                                                           double bfloor, double nlen) {
  constexpr int K = PROJ ? 1 : (P == 0 ? 4 : P + 3);
  static_assert(sizeof(LagDev) % sizeof(double) == 0, "LagDev copied as doubles");
  static_assert(offsetof(LagDev, H) == 0, "the Hessenberg matrix leads the state");
  constexpr int NW = (int)(sizeof(LagDev) / sizeof(double));
  constexpr int NH = kMaxK * kMaxK;
  constexpr bool fresh = !PROJ && P == 0;
  __shared__ double lst[NW];
  __shared__ double sm[K][kFinishBlock / 64];
  __shared__ int publish_s;
  // the skip flag read with the state, tested before the partials' sum
  const int skip_now = PROJ ? gst->pskip : (P > 0 ? gst->done : 0);
  for (int i = threadIdx.x; i < NW; i += blockDim.x)
    lst[i] = (fresh && i < NH) ? 0.0 : reinterpret_cast<const double*>(gst)[i];
  if (skip_now) return;   // uniform
  double r[K];
  sum_partials<K, false, kFinishBlock / 64>(RED_SUM, nblk, partials, sm, r);   // barriers inside
  LagDev* st = reinterpret_cast<LagDev*>(lst);
  if (threadIdx.x == 0) {
    bool ended;
    if constexpr (PROJ) {
      st->pskip = 1;
      lag_finalize<P>(st, sqrt(r[0]), maxl, delta);
      ended = true;
    } else {
      ended = lag_run<P>(st, r, maxl, delta, bfloor, nlen);
    }
    if (ended) lag_solution<P + 1>(st);
    // a solve still running after the last step queued so far publishes too
    publish_s = ended || last;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NW; i += blockDim.x) reinterpret_cast<double*>(gst)[i] = lst[i];
  if (publish_s) lag_publish(st, pub, seq, maxl);   // uniform
}

// The solution of an ended solve and the first Newton update in one pass:
// x = (0 + sum_k c_k X_k) / s with the coefficients c_k of the solve's final
// step (as k_lincombdiv), then FNewtonUpd0 with x (ycor += x, y = zn0 +
// ycor, (x w)^2 summed) on the zero correction of a Newton iteration's
// start (ycor is written, not read).  NV, the number of basis vectors, is a template
// argument so that the coefficients sit in registers.
template <int NV>
struct FNewtonUpdLag0 {
  double c[NV]; PtrPack X; const double *s, *w, *zn0; double *x, *ycor, *y;
  __device__ double operator()(sunindextype i) const {
    double acc = 1.0 * 0.0;
#pragma unroll
    for (int j = 0; j < NV; j++) acc = fma(c[j], X.p[j][i], acc);
    const double d = acc / s[i];
    if (x) x[i] = d;   // uniform: NULL when the caller does not read the solution
    // ycor is zero before a Newton iteration's first update: +0.0 in place of
    // its load gives the bits of fma(1, ycor, d)
    const double cy = fma(1.0, 0.0, 1.0 * d);
    ycor[i] = cy;
    y[i] = fma(1.0, zn0[i], 1.0 * cy);
    const double p = d * w[i];
    return p * p;
  }
};

template <int NV>
__device__ void newton_upd_lag(const double* __restrict__ xc, const PtrPack& X, const double* s,
                               const double* w, const double* zn0, double* x, double* ycor,
                               double* y, sunindextype n, double* partials, double (*sm)[kWaves]) {
  FNewtonUpdLag0<NV> f{};
#pragma unroll
  for (int j = 0; j < NV; j++) f.c[j] = xc[j];
  f.X = X;
  f.s = s;
  f.w = w;
  f.zn0 = zn0;
  f.x = x;
  f.ycor = ycor;
  f.y = y;
  const double v = reduce1_block(RED_SUM, f, n, sm);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// xinfo = {xk, xready} and xc of the solve's state (separate restrict
// pointers, so the uniform reads are scalar loads).  When the solve left no
// solution to use, nothing, and the flag xskip stops the finish queued
// behind.  Partials as k_reduce1 writes them (the same grid and
// accumulation), so the update and its sum are those of k_lincombdiv and
// k_reduce1<FNewtonUpd0> run one after the other.  At most 128 VGPRs, so
// that four workgroups share a CU (the widest basis is register-heavy).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_newton_upd_lag(
    const int* __restrict__ xinfo, const double* __restrict__ xc, int* __restrict__ xskip,
    PtrPack X, const double* s, const double* w, const double* zn0, double* x, double* ycor,
    double* y, sunindextype n, double* partials) {
  __shared__ double sm[1][kWaves];
  const int nv = xinfo[0];
  const int go = xinfo[1] && nv >= 1 && nv <= kLagMax;
  if (blockIdx.x == 0 && threadIdx.x == 0) *xskip = go ? 0 : 1;
  if (!go) return;
  static_assert(kLagMax == 8, "one instantiation per basis size");
  switch (nv) {
    case 1: newton_upd_lag<1>(xc, X, s, w, zn0, x, ycor, y, n, partials, sm); break;
    case 2: newton_upd_lag<2>(xc, X, s, w, zn0, x, ycor, y, n, partials, sm); break;
    case 3: newton_upd_lag<3>(xc, X, s, w, zn0, x, ycor, y, n, partials, sm); break;
    case 4: newton_upd_lag<4>(xc, X, s, w, zn0, x, ycor, y, n, partials, sm); break;
    case 5: newton_upd_lag<5>(xc, X, s, w, zn0, x, ycor, y, n, partials, sm); break;
    case 6: newton_upd_lag<6>(xc, X, s, w, zn0, x, ycor, y, n, partials, sm); break;
    case 7: newton_upd_lag<7>(xc, X, s, w, zn0, x, ycor, y, n, partials, sm); break;
    default: newton_upd_lag<8>(xc, X, s, w, zn0, x, ycor, y, n, partials, sm); break;
  }
}

// u = a w - sum_j b_j U_j with (a, b) read from device memory, |u|^2
// (rounded as FProjSq rounds it)
struct FProjSqDev {
  int nb; const double* coef; const double* w; PtrPack U;
  __device__ double operator()(sunindextype i) const {
    double u = coef[0] * w[i];
    for (int j = 0; j < nb; j++) u = fma(-coef[1 + j], U.p[j][i], u);
    return u * u;
  }
};

template <int P>
void launch_lag_step(hipStream_t s, int nblk, const RedWs* w, bool proj, int last) {
  if (proj)
    hipLaunchKernelGGL((k_lag_step<P, true>), dim3(1), dim3(kFinishBlock), 0, s, nblk,
                       w->partials, last, w->lag, w->lagpub_dev, w->lagseq, w->lagmaxl,
                       w->lagdelta, w->lagbfloor, w->lagnlen);
  else
    hipLaunchKernelGGL((k_lag_step<P, false>), dim3(1), dim3(kFinishBlock), 0, s, nblk,
                       w->partials, last, w->lag, w->lagpub_dev, w->lagseq, w->lagmaxl,
                       w->lagdelta, w->lagbfloor, w->lagnlen);
}

// pass p's step (proj: the projection norm of column p = maxl - 1); 0 <= p < kLagMax
void lag_step_dispatch(hipStream_t s, int nblk, const RedWs* w, int p, bool proj, int last) {
  static_assert(kLagMax == 8, "one instantiation per column");
  switch (p) {
    case 0: launch_lag_step<0>(s, nblk, w, proj, last); break;
    case 1: launch_lag_step<1>(s, nblk, w, proj, last); break;
    case 2: launch_lag_step<2>(s, nblk, w, proj, last); break;
    case 3: launch_lag_step<3>(s, nblk, w, proj, last); break;
    case 4: launch_lag_step<4>(s, nblk, w, proj, last); break;
    case 5: launch_lag_step<5>(s, nblk, w, proj, last); break;
    case 6: launch_lag_step<6>(s, nblk, w, proj, last); break;
    default: launch_lag_step<7>(s, nblk, w, proj, last); break;
  }
}

}  // namespace

static bool is_hip(N_Vector v);

extern "C" {

int NVHip_LagBegin(N_Vector v, int maxl, sunrealtype delta, sunrealtype bnorm_floor,
                   sunrealtype nlen, const double** coef_dev, const int** skip_dev,
                   const int** pskip_dev) {
  if (!v || !coef_dev || !skip_dev || !pskip_dev || !is_hip(v) || maxl < 1 ||
      maxl > kLagMax || g_red_mode == 0 || !(nlen > 0.0))
    return -1;
  RedWs* w = workspace(S(v));
  if (!w->lag) {
    void* hb = nullptr;
    if (hipMalloc(&w->lag, sizeof(LagDev)) != hipSuccess) {
      w->lag = nullptr;
      return -1;
    }
    if (hipHostMalloc(&hb, sizeof(LagPub), hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void**)&w->lagpub_dev, hb, 0) != hipSuccess) {
      if (hb) hipHostFree(hb);
      hipFree(w->lag);
      w->lag = nullptr;
      return -1;
    }
    w->lagpub = static_cast<LagPub*>(hb);
    std::memset(hb, 0, sizeof(LagPub));
  }
  w->lagseq++;
  w->lagmaxl = maxl;
  w->lagdelta = delta;
  w->lagbfloor = bnorm_floor;
  w->lagnlen = nlen;
  *coef_dev = w->lag->coef;
  *skip_dev = &w->lag->done;
  *pskip_dev = &w->lag->pskip;
  return 0;
}

int NVHip_LagStep(N_Vector v, int K, int nblk, int p, int last) {
  if (!v || !is_hip(v) || K < 3 || K > 10 || nblk < 1 || nblk * K > kMaxRedBlocks * kMaxK ||
      p < 0)
    return -1;
  RedWs* w = workspace(S(v));
  if (!w->lag || p >= w->lagmaxl || K != (p == 0 ? 4 : p + 3)) return -1;
  lag_step_dispatch(S(v), nblk, w, p, false, last != 0);
  check_launch("lag step");
  return 0;
}

int NVHip_LagProjNorm(N_Vector wv, int nb, N_Vector* U) {
  if (!wv || !U || !is_hip(wv)) return -1;
  hipStream_t s = S(wv);
  RedWs* w = workspace(s);
  if (!w->lag || nb != w->lagmaxl) return -1;
  FProjSqDev f{};
  f.nb = nb;
  f.coef = w->lag->coef;
  f.w = D(wv);
  for (int j = 0; j < nb; j++) {
    if (!U[j] || LEN(U[j]) != LEN(wv)) return -1;
    f.U.p[j] = D(U[j]);
  }
  const sunindextype n = LEN(wv);
  const int g = red_grid(n);
  hipLaunchKernelGGL(k_reduce1<FProjSqDev>, dim3(g), dim3(kBlock), 0, s, RED_SUM, f, n,
                     w->partials, (unsigned*)nullptr, w->host_dev, (double*)nullptr, 0u,
                     (const int*)&w->lag->pskip);
  lag_step_dispatch(s, g, w, nb - 1, true, 0);
  check_launch("lag projection norm");
  return 0;
}

int NVHip_LagUpdate(N_Vector b, int nvec, N_Vector* V, N_Vector sv, N_Vector x, N_Vector ycor,
                    N_Vector zn0, N_Vector w, N_Vector y, unsigned* ticket) {
  if (!b || !V || !sv || !ycor || !zn0 || !w || !y || !ticket || !is_hip(b) || nvec < 1 ||
      nvec > kMaxK || g_red_mode == 0)
    return -1;
  const sunindextype n = LEN(b);
  hipStream_t s = S(b);
  if (LEN(sv) != n || LEN(ycor) != n || LEN(zn0) != n || LEN(w) != n || LEN(y) != n ||
      S(ycor) != s || (x && (LEN(x) != n || S(x) != s)))
    return -1;
  // the Newton update's aliasing rules (NVHip_NewtonUpdate), and x not a
  // basis vector or the scale (every element is read before x is written,
  // but x == X_j would change the values other elements read)
  double* const dx = x ? D(x) : nullptr;
  if (D(y) == D(ycor) || D(y) == D(zn0) ||
      (dx && (D(y) == dx || D(ycor) == dx || dx == D(sv))))
    return -1;
  RedWs* wk = workspace(s);
  if (!wk->lag) return -1;
  PtrPack X{};
  for (int j = 0; j < nvec; j++) {
    if (!V[j] || LEN(V[j]) != n || (dx && D(V[j]) == dx) || D(V[j]) == D(y)) return -1;
    X.p[j] = D(V[j]);
  }
  static_assert(offsetof(LagDev, xready) == offsetof(LagDev, xk) + sizeof(int),
                "xk, xready adjacent");
  const int g = red_grid(n);
  const unsigned seq = ++wk->seq;
  hipLaunchKernelGGL(k_newton_upd_lag, dim3(g), dim3(kBlock), 0, s, (const int*)&wk->lag->xk,
                     (const double*)wk->lag->xc, &wk->lag->xskip, X, D(sv), D(w), D(zn0), dx,
                     D(ycor), D(y), n, wk->partials);
  launch_finish<1>(s, RED_SUM, g, wk, seq, 0, &wk->lag->xskip);
  check_launch("lag update");
  *ticket = seq;
  return 0;
}

int NVHip_LagUpdateWait(N_Vector ycor, unsigned ticket, sunrealtype* dnorm) {
  if (!ycor || !dnorm || !is_hip(ycor)) return -1;
  hipStream_t s = S(ycor);
  RedWs* w = workspace(s);
  if (!wait_result_checked(s, w, ticket)) return -1;
  *dnorm = std::sqrt(w->host[0] / (double)LEN(ycor));
  return 0;
}

int NVHip_LagNext(N_Vector v) {
  if (!v || !is_hip(v)) return -1;
  RedWs* w = workspace(S(v));
  if (!w->lag) return -1;
  w->lagseq++;
  return 0;
}

int NVHip_LagWait(N_Vector v, NVHipLagResult* res) {
  if (!v || !res || !is_hip(v)) return -1;
  hipStream_t s = S(v);
  RedWs* w = workspace(s);
  if (!w->lag) return -1;
  g_syncs++;
  const unsigned* word = &w->lagpub->seq;
  bool got = false;
  if (g_sync_mode == 1 || g_sync_mode == 2) {
    auto t0 = std::chrono::steady_clock::now();
    for (unsigned it = 0;; it++) {
      if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == w->lagseq) { got = true; break; }
      if (g_sync_mode == 2 && (it & 15u) == 15u) std::this_thread::yield();
      if ((it & 4095u) == 4095u &&
          std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2))
        break;
    }
  }
  if (!got) {
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (__atomic_load_n(word, __ATOMIC_ACQUIRE) != w->lagseq) return -1;
  }
  std::memcpy(res, &w->lagpub->r, sizeof(NVHipLagResult));
  return 0;
}

int NVHip_GSProjectScaled(N_Vector z, int nvec, N_Vector* Y, int sidx, sunrealtype sc,
                          sunrealtype* dots, sunrealtype* zz) {
  if (!z || nvec < 2 || nvec > kMaxK || !dots || !zz || Y[nvec - 1] != z || sidx < 0 ||
      sidx >= nvec - 1)
    return -1;
  return gs_project(nullptr, nullptr, z, nvec, Y, sidx, sc, dots, zz);
}

static bool is_hip(N_Vector v) {
  return v->ops && v->ops->nvgetvectorid == N_VGetVectorID_Hip;
}

double* NVHip_ReducePartials(N_Vector v, int* capacity) {
  if (!v || !capacity || !is_hip(v)) return nullptr;
  RedWs* w = workspace(S(v));
  *capacity = kMaxRedBlocks * kMaxK;
  return w->partials;
}

int NVHip_ReduceFinish(N_Vector v, int K, int nblk, sunrealtype* out) {
  if (!v || !out || K < 1 || K > kMaxK || nblk < 1 || nblk > kMaxRedBlocks * kMaxK / K ||
      !is_hip(v))
    return -1;
  hipStream_t s = S(v);
  RedWs* w = workspace(s);
  const unsigned seq = ++w->seq;
  finish_dispatch(K, s, RED_SUM, nblk, w, seq);
  check_launch("finish");
  wait_result(s, w, seq);
  for (int k = 0; k < K; k++) out[k] = w->host[k];
  return 0;
}

int NVHip_ReduceFinishAsync(N_Vector v, int K, int nblk, const int* skip_dev,
                            unsigned* ticket) {
  if (!v || !ticket || K < 1 || K > kMaxK || nblk < 1 || nblk > kMaxRedBlocks * kMaxK / K ||
      !is_hip(v))
    return -1;
  hipStream_t s = S(v);
  RedWs* w = workspace(s);
  const unsigned seq = ++w->seq;
  switch (K) {
#define NVHIP_FIN_CASE(k) \
    case k: launch_finish<k>(s, RED_SUM, nblk, w, seq, 0, skip_dev); break;
    NVHIP_FIN_CASE(1) NVHIP_FIN_CASE(2) NVHIP_FIN_CASE(3) NVHIP_FIN_CASE(4)
    NVHIP_FIN_CASE(5) NVHIP_FIN_CASE(6) NVHIP_FIN_CASE(7) NVHIP_FIN_CASE(8)
    NVHIP_FIN_CASE(9) NVHIP_FIN_CASE(10) NVHIP_FIN_CASE(11) NVHIP_FIN_CASE(12)
    NVHIP_FIN_CASE(13) NVHIP_FIN_CASE(14) NVHIP_FIN_CASE(15)
    default: launch_finish<16>(s, RED_SUM, nblk, w, seq, 0, skip_dev); break;
#undef NVHIP_FIN_CASE
  }
  check_launch("finish");
  *ticket = seq;
  return 0;
}

int NVHip_ReduceWait(N_Vector v, unsigned ticket, int K, sunrealtype* out) {
  if (!v || !out || K < 1 || K > kMaxK || !is_hip(v)) return -1;
  hipStream_t s = S(v);
  RedWs* w = workspace(s);
  if (!wait_result_checked(s, w, ticket)) return -1;
  for (int k = 0; k < K; k++) out[k] = w->host[k];
  return 0;
}

int NVHip_GateOpen(N_Vector v, const double** coef_dev, const int** skip_dev) {
  if (!v || !coef_dev || !skip_dev || !is_hip(v) || g_sync_mode == 0) return -1;
  hipStream_t s = S(v);
  RedWs* w = workspace(s);
  if (w->gopen) return -1;
  if (!w->gbox) {
    void* hb = nullptr;
    if (hipHostMalloc(&hb, sizeof(GateBox), hipHostMallocMapped | hipHostMallocCoherent) !=
        hipSuccess)
      return -1;
    std::memset(hb, 0, sizeof(GateBox));
    if (hipMalloc(&w->gcoef, sizeof(double) * kMaxK) != hipSuccess ||
        hipMalloc(&w->gskip, sizeof(int)) != hipSuccess) {
      hipHostFree(hb);
      hipFree(w->gcoef);
      w->gcoef = nullptr;
      return -1;
    }
    if (hipHostGetDevicePointer((void**)&w->gbox_dev, hb, 0) != hipSuccess) {
      hipHostFree(hb);
      hipFree(w->gcoef);
      hipFree(w->gskip);
      w->gcoef = nullptr;
      w->gskip = nullptr;
      return -1;
    }
    w->gbox = static_cast<GateBox*>(hb);
  }
  const GateBox* dbox = w->gbox_dev;
  const unsigned expected = ++w->gseq;
  // a released gate is at most microseconds late; after a second the
  // kernels behind it are skipped and the host sees their results missing
  const unsigned long long limit = 100000000ull;   // 1 s at 100 MHz
  hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, s, dbox, expected, w->gcoef,
                     w->gskip, limit);
  check_launch("gate");
  w->gopen = true;
  *coef_dev = w->gcoef;
  *skip_dev = w->gskip;
  return 0;
}

int NVHip_GateRelease(N_Vector v, int stop, const sunrealtype* coef, int n) {
  if (!v || !is_hip(v) || n < 0 || n > kMaxK || (n > 0 && !coef)) return -1;
  RedWs* w = workspace(S(v));
  if (!w->gopen) return -1;
  GateBox* b = w->gbox;
  for (int k = 0; k < n; k++) b->coef[k] = coef[k];
  b->stop = stop ? 1 : 0;
  __atomic_store_n(&b->seq, w->gseq, __ATOMIC_RELEASE);
  w->gopen = false;
  return 0;
}

int NVHip_ProjNorm(int nb, sunrealtype a, const sunrealtype* b, N_Vector w, N_Vector* U,
                   N_Vector uout, sunrealtype* nn) {
  if (!w || !nn || nb < 0 || nb > kMaxK || (nb > 0 && (!U || !b))) return -1;
  FProjSq f{};
  f.nb = nb;
  f.a = a;
  f.w = D(w);
  for (int j = 0; j < nb; j++) {
    if (!U[j] || LEN(U[j]) != LEN(w) || U[j] == uout) return -1;
    f.b.c[j] = b[j];
    f.U.p[j] = D(U[j]);
  }
  if (uout && (uout == w || LEN(uout) != LEN(w))) return -1;
  f.out = uout ? D(uout) : nullptr;
  *nn = reduce1(w, RED_SUM, f);
  return 0;
}

int NVHip_LinCombDot(int nvec, sunrealtype* c, N_Vector* X, N_Vector z, sunrealtype* zz) {
  if (nvec < 1 || nvec > kMaxK || !z || !zz) return -1;
  FLincombSq f{};
  f.nv = nvec;
  f.z = D(z);
  for (int j = 0; j < nvec; j++) {
    if (!X[j] || LEN(X[j]) != LEN(z)) return -1;
    f.c.c[j] = c[j];
    f.X.p[j] = D(X[j]);
  }
  *zz = reduce1(z, RED_SUM, f);
  return 0;
}

void FusedNVec_Init(N_Vector v) {
  if (!v) return;
  N_VEnableFusedOps_Hip(v, SUNTRUE);
  workspace(S(v));
}

void FusedNVec_FreePool(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_ws) {
    hipFree(kv.second->partials);
    hipFree(kv.second->ticket);
    hipHostFree(kv.second->host);
    if (kv.second->gbox) {
      hipHostFree(kv.second->gbox);
      hipFree(kv.second->gcoef);
      hipFree(kv.second->gskip);
    }
    if (kv.second->lag) {
      hipFree(kv.second->lag);
      hipHostFree(kv.second->lagpub);
    }
    delete kv.second;
  }
  g_ws.clear();
}

}  // extern "C"

namespace {
void install_ops(N_Vector v, bool fused) {
  N_Vector_Ops o = v->ops;
  std::memset(o, 0, sizeof(*o));
  o->nvgetvectorid = N_VGetVectorID_Hip;
  o->nvclone = N_VClone_Hip;
  o->nvcloneempty = N_VCloneEmpty_Hip;
  o->nvdestroy = N_VDestroy_Hip;
  o->nvspace = N_VSpace_Hip;
  o->nvgetarraypointer = N_VGetHostArrayPointer_Hip;
  o->nvgetdevicearraypointer = N_VGetDeviceArrayPointer_Hip;
  o->nvsetarraypointer = nv_setarray;
  o->nvgetcommunicator = nv_comm;
  o->nvgetlength = N_VGetLength_Hip;
  o->nvgetlocallength = N_VGetLength_Hip;
  o->nvlinearsum = N_VLinearSum_Hip;
  o->nvconst = N_VConst_Hip;
  o->nvprod = N_VProd_Hip;
  o->nvdiv = N_VDiv_Hip;
  o->nvscale = N_VScale_Hip;
  o->nvabs = N_VAbs_Hip;
  o->nvinv = N_VInv_Hip;
  o->nvaddconst = N_VAddConst_Hip;
  o->nvdotprod = N_VDotProd_Hip;
  o->nvmaxnorm = N_VMaxNorm_Hip;
  o->nvwrmsnorm = N_VWrmsNorm_Hip;
  o->nvwrmsnormmask = N_VWrmsNormMask_Hip;
  o->nvmin = N_VMin_Hip;
  o->nvwl2norm = N_VWL2Norm_Hip;
  o->nvl1norm = N_VL1Norm_Hip;
  o->nvcompare = N_VCompare_Hip;
  o->nvinvtest = N_VInvTest_Hip;
  o->nvconstrmask = N_VConstrMask_Hip;
  o->nvminquotient = N_VMinQuotient_Hip;
  o->nvdotprodlocal = nv_dot_local;
  o->nvmaxnormlocal = N_VMaxNorm_Hip;
  o->nvminlocal = N_VMin_Hip;
  o->nvl1normlocal = N_VL1Norm_Hip;
  o->nvinvtestlocal = N_VInvTest_Hip;
  o->nvconstrmasklocal = N_VConstrMask_Hip;
  o->nvminquotientlocal = N_VMinQuotient_Hip;
  o->nvwsqrsumlocal = N_VWSqrSumLocal_Hip;
  o->nvwsqrsummasklocal = N_VWSqrSumMaskLocal_Hip;
  o->nvprint = nv_print;
  o->nvprintfile = nv_printfile;
  if (fused) N_VEnableFusedOps_Hip(v, SUNTRUE);
}
}  // namespace
