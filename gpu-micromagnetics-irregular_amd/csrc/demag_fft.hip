// demag_fft.hip -- FFT demagnetising field on gfx950 with a hand-written
// LDS Stockham FFT.  Replaces the reference's cuFFT pipeline
// gather -> D2Z(batch 3) -> multiply(9 complex planes) -> Z2D -> shift/scale
// (src/irregular/2d_fft/2d_i4/demag_fft.cu:94-161,184-363) with three
// fused passes per Demag_Apply:
//
//   1. row pass   : per (row, component) workgroup: gather y (x w for the
//                   windowed form) straight from the SoA state, real FFT of
//                   length Px as a complex FFT of length Px/2 in LDS plus
//                   the split post-process; writes Px/2+1 bins.
//   2. column pass: per spectral column kx: the three components' column
//                   FFTs (length Py) in LDS, the 3x3 tensor product, the
//                   inverse column FFTs; only the rows the output needs
//                   are written back.
//   3. row pass   : inverse real FFT per (row, component), writes the
//                   field directly into the SoA output (overwrite).
//
// The tensor is stored as real planes because the centred thin-film
// kernels are even (xx, yy, zz) or odd-odd (xy); only the periodic xy
// plane keeps an imaginary part (its -n/2 line breaks the odd symmetry).
// xz/yz/zx/zy vanish for the z=0 source (their spectra are rounding
// noise); the periodic shift and the strength/(Px*Py) scale are folded
// into the stored spectrum.  The Newell tensor itself (`ctt`/`calt`,
// 2d_i4/demag_fft.cu:27-91) is evaluated on the GPU over one quadrant of
// displacements and mirrored.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "llg/demag_fft.h"
#include "llg/nvector_hip.h"
#include "demag_reg.h"
#include "kprof.h"

// No implicit multiply-add contraction in this file: every fused
// multiply-add is written out (fmad below), so a kernel's rounding does
// not depend on how the compiler schedules it.  The general and the
// compile-time FFT plans therefore give identical bits.
#pragma clang fp contract(off)

namespace {

using rfft::cadd;
using rfft::cconj;
using rfft::cmul;
using rfft::cmuli;
using rfft::csub;
using rfft::fmad;
using rfft::mk;

thread_local std::string g_derr;
void derr(const std::string& e) { g_derr = e; }

constexpr int kMaxPasses = 16;
constexpr int kXcd = 8;
constexpr size_t kMaxLds = 160 * 1024;   // LDS of one gfx950 CU

// dynamic LDS above 64 KB has to be allowed per kernel
template <class K>
bool allow_lds(K kern, size_t bytes) {
  if (bytes <= 65536) return true;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess;
}

// complex types of the two pipeline precisions (a plain pair, same size
// and alignment as double2 / float2)
template <class T> using V2 = rfft::Cplx<T>;

template <class T>
struct Plan1D {
  int L;
  int npass;
  int radix[kMaxPasses];
  int bpt[kMaxPasses];      // butterflies per thread per pass (<= 8)
  int nsh[kMaxPasses];      // log2(Ns) of the pass when Ns is a power of two, else -1
  int nbsh[kMaxPasses];     // log2(L / radix) when a power of two, else -1
  int psh;                  // LDS padding: one spare slot per 2^psh values (31: none)
  const V2<T>* tw;          // exp(-2 pi i q / L), q in [0, L)
};

// a complex value in LDS (the FFT helpers are always inlined into the
// kernels, so these pointers resolve to the LDS address space)
template <class T>
using LdsV2 = V2<T>;

// in-register DFT of length R, kernel exp(sign * 2 pi i j k / R)
template <int R, class T>
struct Dft;

template <class T>
struct Dft<2, T> {
  __device__ static inline void run(V2<T> (&v)[2], T) {
    const V2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <class T>
struct Dft<4, T> {
  __device__ static inline void run(V2<T> (&v)[4], T sign) {
    const V2<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const V2<T> c = cadd(v[1], v[3]), d = cmuli(csub(v[1], v[3]), sign);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};
template <class T>
struct Dft<8, T> {
  __device__ static inline void run(V2<T> (&v)[8], T sign) {
    V2<T> e[4] = {v[0], v[2], v[4], v[6]};
    V2<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, T>::run(e, sign);
    Dft<4, T>::run(o, sign);
    const T r = T(0.70710678118654752440);
    const V2<T> w1 = mk<T>(r, sign * r);   // exp(sign 2 pi i / 8)
    const V2<T> w3 = mk<T>(-r, sign * r);
    const V2<T> t0 = o[0];
    const V2<T> t1 = cmul(o[1], w1);
    const V2<T> t2 = cmuli(o[2], sign);
    const V2<T> t3 = cmul(o[3], w3);
    v[0] = cadd(e[0], t0); v[4] = csub(e[0], t0);
    v[1] = cadd(e[1], t1); v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2); v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3); v[7] = csub(e[3], t3);
  }
};
template <class T>
struct Dft<3, T> {
  __device__ static inline void run(V2<T> (&v)[3], T sign) {
    const T c = T(-0.5), s = sign * T(0.86602540378443864676);
    const V2<T> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const V2<T> a = mk<T>(fmad(c, t1.x, v[0].x), fmad(c, t1.y, v[0].y));
    const V2<T> b = mk<T>(-s * t2.y, s * t2.x);   // i*s*t2
    v[0] = cadd(v[0], t1);
    v[1] = cadd(a, b);
    v[2] = csub(a, b);
  }
};
template <class T>
struct Dft<5, T> {
  __device__ static inline void run(V2<T> (&v)[5], T sign) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = sign * T(0.95105651629515357212), s2 = sign * T(0.58778525229247312917);
    const V2<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const V2<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const V2<T> a1 = mk<T>(fmad(c2, t2.x, fmad(c1, t1.x, v[0].x)), fmad(c2, t2.y, fmad(c1, t1.y, v[0].y)));
    const V2<T> a2 = mk<T>(fmad(c1, t2.x, fmad(c2, t1.x, v[0].x)), fmad(c1, t2.y, fmad(c2, t1.y, v[0].y)));
    const V2<T> b1r = mk<T>(fmad(s1, t3.x, s2 * t4.x), fmad(s1, t3.y, s2 * t4.y));
    const V2<T> b2r = mk<T>(fmad(s2, t3.x, -(s1 * t4.x)), fmad(s2, t3.y, -(s1 * t4.y)));
    const V2<T> b1 = mk<T>(-b1r.y, b1r.x);   // i * b1r
    const V2<T> b2 = mk<T>(-b2r.y, b2r.x);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
};

// LDS index padding: one spare slot per 2^sh complex values, so the
// stride-R accesses of the Stockham passes spread over the banks
// (sh = 31 disables it).  Padding costs LDS, hence occupancy.
__device__ __host__ inline int lpad(int i, int sh) { return i + (i >> sh); }
__device__ __host__ inline int lpad_len(int L, int sh) { return L + (L >> sh) + 1; }

// One Stockham pass of radix R over ncomp transforms of length L stored at
// buf + comp*cstride (padded index, in place: all reads, barrier, all
// writes, barrier).  Butterfly b of the pass is (comp, j), j = grp*Ns + k;
// the divisions become shifts when Ns and L/R are powers of two (nsh,
// nbsh >= 0), which they are for every power-of-two length.
template <int R, int B, int SIGN, class T>
__device__ __forceinline__ void pass(LdsV2<T>* buf, int L, int Ns, const V2<T>* __restrict__ tw, int ncomp,
                     int cstride, int sh, int nsh, int nbsh) {
  const T sign = T(SIGN);
  const int nb = L / R;
  const int total = nb * ncomp;
  const int tstep = L / (Ns * R);
  V2<T> v[B][R];
  int dbase[B];
#pragma unroll
  for (int r = 0; r < B; r++) {
    const int b = threadIdx.x + r * blockDim.x;
    dbase[r] = -1;
    if (b < total) {
      int comp, j;
      if (ncomp == 1) {
        comp = 0;
        j = b;
      } else if (nbsh >= 0) {
        comp = b >> nbsh;
        j = b & (nb - 1);
      } else {
        comp = b / nb;
        j = b - comp * nb;
      }
      int k, grp;
      if (nsh >= 0) {
        k = j & (Ns - 1);
        grp = j >> nsh;
      } else {
        grp = j / Ns;
        k = j - grp * Ns;
      }
      const LdsV2<T>* src = buf + comp * cstride;
      dbase[r] = comp * cstride + grp * Ns * R + k;
      const int tk = k * tstep;
#pragma unroll
      for (int t = 0; t < R; t++) {
        V2<T> x = src[lpad(j + t * nb, sh)];
        if (t > 0) {
          // branch-free: tw[0] = (1, 0) leaves x unchanged when k == 0
          V2<T> w = tw[t * tk];
          if (SIGN > 0) w.y = -w.y;
          x = cmul(x, w);
        }
        v[r][t] = x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < B; r++) {
    if (dbase[r] >= 0) {
      Dft<R, T>::run(v[r], sign);
      // dbase holds comp*cstride + the unpadded in-transform index: pad
      // the in-transform part only
      const int b = threadIdx.x + r * blockDim.x;
      int comp;
      if (ncomp == 1) comp = 0;
      else if (nbsh >= 0) comp = b >> nbsh;
      else comp = b / nb;
      LdsV2<T>* dst = buf + comp * cstride;
      const int base = dbase[r] - comp * cstride;
#pragma unroll
      for (int t = 0; t < R; t++) dst[lpad(base + t * Ns, sh)] = v[r][t];
    }
  }
  __syncthreads();
}

// register budget: R*B complex values per thread
template <int R>
struct MaxB { static constexpr int value = R == 8 ? 2 : (R == 5 ? 3 : (R == 2 ? 8 : 4)); };

template <int R, int B, int SIGN, class T>
__device__ __forceinline__ void pass_if(LdsV2<T>* buf, int L, int Ns, const V2<T>* tw, int ncomp,
                               int cstride, int sh, int nsh, int nbsh) {
  if constexpr (B <= MaxB<R>::value)
    pass<R, B, SIGN, T>(buf, L, Ns, tw, ncomp, cstride, sh, nsh, nbsh);
}

template <int R, int SIGN, class T>
__device__ __forceinline__ void pass_b(LdsV2<T>* buf, int L, int Ns, const V2<T>* tw, int ncomp, int cstride,
                       int B, int sh, int nsh, int nbsh) {
  switch (B) {
    case 1: pass_if<R, 1, SIGN, T>(buf, L, Ns, tw, ncomp, cstride, sh, nsh, nbsh); break;
    case 2: pass_if<R, 2, SIGN, T>(buf, L, Ns, tw, ncomp, cstride, sh, nsh, nbsh); break;
    case 3: pass_if<R, 3, SIGN, T>(buf, L, Ns, tw, ncomp, cstride, sh, nsh, nbsh); break;
    case 4: pass_if<R, 4, SIGN, T>(buf, L, Ns, tw, ncomp, cstride, sh, nsh, nbsh); break;
    default: pass_if<R, 8, SIGN, T>(buf, L, Ns, tw, ncomp, cstride, sh, nsh, nbsh); break;
  }
}

// in-place FFT of ncomp transforms in LDS; SIGN -1 forward, +1 inverse.
// Always inlined: a call would spill the plan (a kernel argument) to the
// stack and turn the twiddle loads into flat loads.
template <int SIGN, class T>
__device__ __forceinline__ void fft_lds(LdsV2<T>* buf, const Plan1D<T>& p, int ncomp, int cstride) {
  int Ns = 1;
  for (int i = 0; i < p.npass; i++) {
    const int R = p.radix[i], B = p.bpt[i];
    const int ns = p.nsh[i], nbs = p.nbsh[i];
    switch (R) {
      case 8: pass_b<8, SIGN, T>(buf, p.L, Ns, p.tw, ncomp, cstride, B, p.psh, ns, nbs); break;
      case 4: pass_b<4, SIGN, T>(buf, p.L, Ns, p.tw, ncomp, cstride, B, p.psh, ns, nbs); break;
      case 2: pass_b<2, SIGN, T>(buf, p.L, Ns, p.tw, ncomp, cstride, B, p.psh, ns, nbs); break;
      case 3: pass_b<3, SIGN, T>(buf, p.L, Ns, p.tw, ncomp, cstride, B, p.psh, ns, nbs); break;
      default: pass_b<5, SIGN, T>(buf, p.L, Ns, p.tw, ncomp, cstride, B, p.psh, ns, nbs); break;
    }
    Ns *= R;
  }
}

// ------------------------------------------------ compile-time plans
// The same Stockham passes with the length, radices, workgroup size and LDS
// padding as template constants: every index division folds to a shift or
// a multiply, and butterflies per thread are exact.  Same arithmetic as
// fft_lds, so the results are identical.
template <int L, int R, int Ns, int NCOMP, int BLOCK, int SH, int SIGN, class T>
__device__ __forceinline__ void pass_ct(LdsV2<T>* buf, const V2<T>* __restrict__ tw, int cs) {
  constexpr int nb = L / R, total = nb * NCOMP, B = (total + BLOCK - 1) / BLOCK;
  constexpr int tstep = L / (Ns * R);
  const T sign = T(SIGN);
  V2<T> v[B][R];
  int base[B], coff[B];
#pragma unroll
  for (int r = 0; r < B; r++) {
    const int b = threadIdx.x + r * BLOCK;
    coff[r] = -1;
    if (total % BLOCK == 0 || b < total) {
      const int comp = b / nb, j = b - comp * nb;
      const int grp = j / Ns, k = j - grp * Ns;
      coff[r] = comp * cs;
      base[r] = grp * Ns * R + k;
      const LdsV2<T>* src = buf + coff[r];
      const int tk = k * tstep;
#pragma unroll
      for (int t = 0; t < R; t++) {
        V2<T> x = src[lpad(j + t * nb, SH)];
        if (t > 0) {
          V2<T> w = tw[t * tk];
          if (SIGN > 0) w.y = -w.y;
          x = cmul(x, w);
        }
        v[r][t] = x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < B; r++) {
    if (coff[r] >= 0) {
      Dft<R, T>::run(v[r], sign);
      LdsV2<T>* dst = buf + coff[r];
#pragma unroll
      for (int t = 0; t < R; t++) dst[lpad(base[r] + t * Ns, SH)] = v[r][t];
    }
  }
  __syncthreads();
}

template <int SIGN, class T, int L, int NCOMP, int BLOCK, int SH, int Ns>
__device__ __forceinline__ void fft_ct_rec(LdsV2<T>*, const V2<T>*, int) {}

template <int SIGN, class T, int L, int NCOMP, int BLOCK, int SH, int Ns, int R, int... Rest>
__device__ __forceinline__ void fft_ct_rec(LdsV2<T>* buf, const V2<T>* tw, int cs) {
  pass_ct<L, R, Ns, NCOMP, BLOCK, SH, SIGN, T>(buf, tw, cs);
  fft_ct_rec<SIGN, T, L, NCOMP, BLOCK, SH, Ns * R, Rest...>(buf, tw, cs);
}

// FFT policies of the kernels: the runtime plan, or a compile-time one
template <class T>
struct GenFft {
  static constexpr bool kCompileTime = false;
  Plan1D<T> p;
  __device__ int len() const { return p.L; }
  __device__ int shift() const { return p.psh; }
  template <int SIGN>
  __device__ __forceinline__ void run(LdsV2<T>* buf, int ncomp, int cs) const {
    fft_lds<SIGN, T>(buf, p, ncomp, cs);
  }
};

template <class T, int L, int BLOCK, int SH, int... Rs>
struct CtFft {
  static constexpr bool kCompileTime = true;
  const V2<T>* tw;
  static constexpr int kL = L, kBlock = BLOCK, kSh = SH;
  static constexpr int kRad[] = {Rs...};
  static constexpr int kNpass = sizeof...(Rs);
  __device__ int len() const { return L; }
  __device__ int shift() const { return SH; }
  template <int SIGN>
  __device__ __forceinline__ void run(LdsV2<T>* buf, int ncomp, int cs) const {
    if (ncomp == 3) fft_ct_rec<SIGN, T, L, 3, BLOCK, SH, 1, Rs...>(buf, tw, cs);
    else fft_ct_rec<SIGN, T, L, 1, BLOCK, SH, 1, Rs...>(buf, tw, cs);
  }
};

// ----------------------------------------------------------- row forward
// grid (rows, ncomp); in = base + comp*comp_stride + row*nx_in (FP64 input)
template <class T, class F>
__global__ __launch_bounds__(512) void k_row_fwd(const double* __restrict__ in,
                                                  const double* __restrict__ w, long comp_stride,
                                                  int nx_in, int rows, F plan,
                                                  const V2<T>* __restrict__ post,
                                                  V2<T>* __restrict__ spec) {
  extern __shared__ char lds_raw[];
  LdsV2<T>* lds = reinterpret_cast<LdsV2<T>*>(lds_raw);
  const int row = blockIdx.x, comp = blockIdx.y;
  const int L = plan.len(), Hx = L + 1, sh = plan.shift();
  const double* src = in + comp * comp_stride + (long)row * nx_in;
  const double* wr = w ? w + (long)row * nx_in : nullptr;
  for (int m = threadIdx.x; m < L; m += blockDim.x) {
    const int n0 = 2 * m, n1 = 2 * m + 1;
    double a = 0.0, b = 0.0;
    if (n0 < nx_in) a = wr ? wr[n0] * src[n0] : src[n0];
    if (n1 < nx_in) b = wr ? wr[n1] * src[n1] : src[n1];
    lds[lpad(m, sh)] = mk<T>((T)a, (T)b);
  }
  __syncthreads();
  plan.template run<-1>(lds, 1, lpad_len(L, sh));
  V2<T>* out = spec + ((long)comp * rows + row) * Hx;
  for (int k = threadIdx.x; k <= L; k += blockDim.x) {
    const V2<T> zk = lds[lpad(k == L ? 0 : k, sh)];
    const V2<T> zc = cconj(lds[lpad(k == 0 ? 0 : L - k, sh)]);
    const V2<T> e = mk<T>(T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y));
    const V2<T> d = csub(zk, zc);
    const V2<T> o = mk<T>(T(0.5) * d.y, T(-0.5) * d.x);
    out[k] = cadd(e, cmul(post[k], o));
  }
}

// ---------------------------------------------------------- row inverse
template <class T, class F>
__global__ __launch_bounds__(512) void k_row_inv(const V2<T>* __restrict__ spec, int rows,
                                                  F plan, const V2<T>* __restrict__ post,
                                                  double* __restrict__ out, long comp_stride,
                                                  int nx_out) {
  extern __shared__ char lds_raw[];
  LdsV2<T>* lds = reinterpret_cast<LdsV2<T>*>(lds_raw);
  const int row = blockIdx.x, comp = blockIdx.y;
  const int L = plan.len(), Hx = L + 1, sh = plan.shift();
  const V2<T>* in = spec + ((long)comp * rows + row) * Hx;
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    const V2<T> xk = in[k];
    const V2<T> xc = cconj(in[L - k]);
    const V2<T> a = cadd(xk, xc);
    const V2<T> b = cmul(csub(xk, xc), cconj(post[k]));
    lds[lpad(k, sh)] = mk<T>(a.x - b.y, a.y + b.x);   // a + i b
  }
  __syncthreads();
  plan.template run<1>(lds, 1, lpad_len(L, sh));
  double* dst = out + comp * comp_stride + (long)row * nx_out;
  for (int m = threadIdx.x; m < L; m += blockDim.x) {
    const int n0 = 2 * m;
    const V2<T> z = lds[lpad(m, sh)];
    if (n0 < nx_out) dst[n0] = (double)z.x;
    if (n0 + 1 < nx_out) dst[n0 + 1] = (double)z.y;
  }
}

__device__ inline int xcd_remap(int bid, int G) {
  const int per = G / kXcd, rem = G % kXcd;
  const int x = bid % kXcd, idx = bid / kXcd;
  return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

// --------------------------------------------------------------- column
// mode 0: fwd FFT, 3x3 tensor product, inverse FFT (3 components)
// mode 1: fwd FFT only (1 component; tensor spectra at init)
template <class T, class F>
__global__ __launch_bounds__(512) void k_col(V2<T>* __restrict__ spec, int rows_in,
                                              int rows_out, int Hx, F plan,
                                              const T* __restrict__ K, int has_xy_im, int mode) {
  extern __shared__ char lds_raw[];
  LdsV2<T>* lds = reinterpret_cast<LdsV2<T>*>(lds_raw);
  const int kx = xcd_remap(blockIdx.x, gridDim.x);
  const int Py = plan.len();
  const int ncomp = mode == 0 ? 3 : 1;
  const int sh = plan.shift();
  const int cs = lpad_len(Py, sh);
  for (int c = 0; c < ncomp; c++)
    for (int y = threadIdx.x; y < Py; y += blockDim.x)
      lds[c * cs + lpad(y, sh)] = y < rows_in ? spec[((long)c * rows_in + y) * Hx + kx] : mk<T>(T(0), T(0));
  __syncthreads();
  plan.template run<-1>(lds, ncomp, cs);
  if (mode == 0) {
    const long P = (long)Hx * Py;
    const T* kxx = K + (long)kx * Py;
    const T* kyy = kxx + P;
    const T* kzz = kyy + P;
    const T* kxyr = kzz + P;
    const T* kxyi = kxyr + P;
    for (int ky = threadIdx.x; ky < Py; ky += blockDim.x) {
      const int q = lpad(ky, sh);
      const V2<T> mx = lds[q], my = lds[cs + q], mz = lds[2 * cs + q];
      const T a = kxx[ky], b = kyy[ky], cz = kzz[ky];
      const V2<T> xy = mk<T>(kxyr[ky], has_xy_im ? kxyi[ky] : T(0));
      const V2<T> cxy = cmul(xy, my), cyx = cmul(xy, mx);
      const V2<T> hx = mk<T>(fmad(a, mx.x, cxy.x), fmad(a, mx.y, cxy.y));
      const V2<T> hy = mk<T>(fmad(b, my.x, cyx.x), fmad(b, my.y, cyx.y));
      lds[q] = hx;
      lds[cs + q] = hy;
      lds[2 * cs + q] = mk<T>(cz * mz.x, cz * mz.y);
    }
    __syncthreads();
    plan.template run<1>(lds, 3, cs);
  }
  for (int c = 0; c < ncomp; c++)
    for (int y = threadIdx.x; y < rows_out; y += blockDim.x)
      spec[((long)c * rows_out + y) * Hx + kx] = lds[c * cs + lpad(y, sh)];
}

__global__ void k_to_float(const double* __restrict__ a, float* __restrict__ b, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// ------------------------------------------------------------ tensor
// Closed-form field of a uniformly magnetised cell of half-width a and
// half-thickness b at the in-plane offset (sx, sy), sz = 0
// (2d_i4/demag_fft.cu:27-61).  dm[1..9] = xx, xy, xz, yx, yy, yz, zx, zy, zz;
// ALL = false computes only the four the thin-film kernel uses (xx, xy,
// yy, zz; xz, yz, zx, zy vanish at sz = 0 and yx = xy).  One source for the
// device tensor builder and the host self-coupling / full-tensor export.
template <bool ALL>
__host__ __device__ void ctt(double b, double a, double sx, double sy, double dm[10]) {
  const double sz = 0.0;
  const double xn = sx - a, xp = sx + a, yn = sy - a, yp = sy + a, zn = sz - b, zp = sz + b;
  const double xn2 = xn * xn, xp2 = xp * xp, yn2 = yn * yn, yp2 = yp * yp, zn2 = zn * zn,
               zp2 = zp * zp;
  const double dnnn = sqrt(xn2 + yn2 + zn2), dpnn = sqrt(xp2 + yn2 + zn2);
  const double dnpn = sqrt(xn2 + yp2 + zn2), dnnp = sqrt(xn2 + yn2 + zp2);
  const double dppn = sqrt(xp2 + yp2 + zn2), dnpp = sqrt(xn2 + yp2 + zp2);
  const double dpnp = sqrt(xp2 + yn2 + zp2), dppp = sqrt(xp2 + yp2 + zp2);
  dm[1] = atan(zn * yn / (xn * dnnn)) - atan(zp * yn / (xn * dnnp)) -
          atan(zn * yp / (xn * dnpn)) + atan(zp * yp / (xn * dnpp)) -
          atan(zn * yn / (xp * dpnn)) + atan(zp * yn / (xp * dpnp)) +
          atan(zn * yp / (xp * dppn)) - atan(zp * yp / (xp * dppp));
  dm[2] = log((dnnn - zn) / (dnnp - zp)) - log((dpnn - zn) / (dpnp - zp)) -
          log((dnpn - zn) / (dnpp - zp)) + log((dppn - zn) / (dppp - zp));
  dm[5] = atan(zn * xn / (yn * dnnn)) - atan(zp * xn / (yn * dnnp)) -
          atan(zn * xp / (yn * dpnn)) + atan(zp * xp / (yn * dpnp)) -
          atan(zn * xn / (yp * dnpn)) + atan(zp * xn / (yp * dnpp)) +
          atan(zn * xp / (yp * dppn)) - atan(zp * xp / (yp * dppp));
  dm[9] = atan(xn * yn / (zn * dnnn)) - atan(xp * yn / (zn * dpnn)) -
          atan(xn * yp / (zn * dnpn)) + atan(xp * yp / (zn * dppn)) -
          atan(xn * yn / (zp * dnnp)) + atan(xp * yn / (zp * dpnp)) +
          atan(xn * yp / (zp * dnpp)) - atan(xp * yp / (zp * dppp));
  if constexpr (ALL) {
    dm[3] = log((dnnn - yn) / (dnpn - yp)) - log((dpnn - yn) / (dppn - yp)) -
            log((dnnp - yn) / (dnpp - yp)) + log((dpnp - yn) / (dppp - yp));
    dm[4] = log((dnnn - zn) / (dnnp - zp)) - log((dnpn - zn) / (dnpp - zp)) -
            log((dpnn - zn) / (dpnp - zp)) + log((dppn - zn) / (dppp - zp));
    dm[6] = log((dnnn - xn) / (dpnn - xp)) - log((dnpn - xn) / (dppn - xp)) -
            log((dnnp - xn) / (dpnp - xp)) + log((dnpp - xn) / (dppp - xp));
    dm[7] = log((dnnn - yn) / (dnpn - yp)) - log((dnnp - yn) / (dnpp - yp)) -
            log((dpnn - yn) / (dppn - yp)) + log((dpnp - yn) / (dppp - yp));
    dm[8] = log((dnnn - xn) / (dpnn - xp)) - log((dnnp - xn) / (dpnp - xp)) -
            log((dnpn - xn) / (dppn - xp)) + log((dnpp - xn) / (dppp - xp));
  }
}

// quadrant Q[c][dy][dx], dx in [0,qx), dy in [0,qy); c = xx, yy, zz, xy
__global__ void k_tensor_quadrant(double thick, int qx, int qy, double* __restrict__ Q) {
  const int dx = blockIdx.x * blockDim.x + threadIdx.x;
  const int dy = blockIdx.y;
  if (dx >= qx || dy >= qy) return;
  const double a = 0.49999, b = 0.5 * thick;
  double sxx = 0, syy = 0, szz = 0, sxy = 0;
  double dm[10];
  for (int jy = -4; jy <= 4; jy++) {
    const double sy = (double)dy + 0.1 * (double)jy;
    for (int ix = -4; ix <= 4; ix++) {
      const double sx = (double)dx + 0.1 * (double)ix;
      ctt<false>(b, a, sx, sy, dm);
      sxx += dm[1]; sxy += dm[2]; syy += dm[5]; szz += dm[9];
    }
  }
  const long P = (long)qx * qy, k = (long)dy * qx + dx;
  Q[k] = sxx / 81.;
  Q[P + k] = syy / 81.;
  Q[2 * P + k] = szz / 81.;
  Q[3 * P + k] = sxy / 81.;
}

// fill the four real kernel planes K[c][r][s] (Py x Px) from the quadrant.
// periodic: index s <-> displacement wrapped into [-Px/2, Px - Px/2)
// padded  : s < Px/2 -> d = s, s > Px/2 -> d = s - Px, s == Px/2 -> 0
__global__ void k_tensor_fill(const double* __restrict__ Q, int qx, int qy, int Px, int Py,
                              int padded, double* __restrict__ K) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (s >= Px) return;
  const long P = (long)Px * Py, k = (long)r * Px + s;
  int dx, dy;
  bool zero = false;
  if (padded) {
    const int hx = Px / 2, hy = Py / 2;
    if (s == hx || r == hy) zero = true;
    dx = s < hx ? s : s - Px;
    dy = r < hy ? r : r - Py;
  } else {
    const int hx = Px / 2, hy = Py / 2;
    dx = ((s + hx) % Px) - hx;
    dy = ((r + hy) % Py) - hy;
  }
  if (zero) {
    K[k] = K[P + k] = K[2 * P + k] = K[3 * P + k] = 0.0;
    return;
  }
  const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy;
  const long PQ = (long)qx * qy, q = (long)ay * qx + ax;
  const double sgn = ((dx < 0) != (dy < 0)) ? -1.0 : 1.0;
  K[k] = Q[q];
  K[P + k] = Q[PQ + q];
  K[2 * P + k] = Q[2 * PQ + q];
  K[3 * P + k] = sgn * Q[3 * PQ + q];
}

// spectra [ky][kx] (one plane) -> stored plane [kx][ky] (transposed), scaled
__global__ void k_extract(const double2* __restrict__ S, int Py, int Hx, double scale,
                          int take_imag, double* __restrict__ outp) {
  const int kx = blockIdx.x * blockDim.x + threadIdx.x;
  const int ky = blockIdx.y;
  if (kx >= Hx) return;
  const double2 v = S[(long)ky * Hx + kx];
  outp[(long)kx * Py + ky] = scale * (take_imag ? v.y : v.x);
}

// Zero-padded grids: the kernel planes are even in dy (xx, yy, zz) or odd
// (xy), and the row dy = -Py/2 is zero, so their column spectra satisfy
// K(Py - ky) = +-K(ky).  The FFT computes the two halves with different
// roundings; store the upper half as the exact mirror of the lower so a
// column kernel may read ky <= Py/2 only and lose no bit.
__global__ void k_mirror_ky(double* __restrict__ K, int Hx, int Py) {
  const int kx = blockIdx.x * blockDim.x + threadIdx.x;
  const int ky = Py / 2 + 1 + blockIdx.y;
  const int c = blockIdx.z;
  if (kx >= Hx || ky >= Py) return;
  double* plane = K + (long)c * Hx * Py + (long)kx * Py;
  plane[ky] = (c == 3 ? -1.0 : 1.0) * plane[Py - ky];
}

// ---------------------------------------------------------- host helpers
bool factorize(int L, std::vector<int>& rad) {
  rad.clear();
  int n = L;
  while (n % 8 == 0 && n > 1) { rad.push_back(8); n /= 8; }
  if (n % 2 == 0 && n % 4 != 0 && !rad.empty() && rad.back() == 8) {
    rad.back() = 4;   // 8*2 -> 4*4
    n *= 2;
  }
  while (n % 4 == 0 && n > 1) { rad.push_back(4); n /= 4; }
  while (n % 2 == 0 && n > 1) { rad.push_back(2); n /= 2; }
  while (n % 3 == 0 && n > 1) { rad.push_back(3); n /= 3; }
  while (n % 5 == 0 && n > 1) { rad.push_back(5); n /= 5; }
  return n == 1 && (int)rad.size() <= kMaxPasses;
}

template <class T>
struct HostPlan {
  Plan1D<T> p{};
  int ncomp = 1;
  int block = 64;
  V2<T>* d_tw = nullptr;
};

int max_b(int r) { return r == 8 ? 2 : (r == 5 ? 3 : (r == 2 ? 8 : 4)); }

template <class T>
bool make_plan(int L, int ncomp, HostPlan<T>& hp) {
  std::vector<int> rad;
  if (L < 1 || !factorize(L, rad)) {
    derr("unsupported FFT length " + std::to_string(L) + " (need 2^a 3^b 5^c)");
    return false;
  }
  hp.p.L = L;
  hp.p.npass = (int)rad.size();
  // row plans (one transform per workgroup, small LDS) pad every 8 values;
  // the 3-component column plan pads every 16 so three workgroups still
  // fit one CU's 160 KB of LDS at Py = 1024 in FP64
  hp.p.psh = ncomp == 3 ? 4 : 3;
  // default workgroup: the smallest multiple of 64 for which every pass
  // fits its per-thread register budget, at least min(widest pass, 256)
  int block = 64;
  for (int r : rad) {
    const int tot = ncomp * L / r;
    const int need = (tot + max_b(r) - 1) / max_b(r);
    block = std::max(block, ((need + 63) / 64) * 64);
  }
  int maxb = 1;
  for (int r : rad) maxb = std::max(maxb, ncomp * L / r);
  block = std::max(block, std::min(((maxb + 63) / 64) * 64, 256));
  // the Stockham kernels hold all ncomp transforms in one workgroup's LDS;
  // a length beyond one workgroup keeps only its twiddle table (the
  // register-resident kernels cover it) and block = 0 marks it
  if (block > 512 || (size_t)ncomp * lpad_len(L, hp.p.psh) * sizeof(V2<T>) > kMaxLds) block = 0;
  hp.ncomp = ncomp;
  hp.block = block;
  auto log2_if_pow2 = [](int x) {
    if (x < 1 || (x & (x - 1))) return -1;
    int l = 0;
    while ((1 << l) < x) l++;
    return l;
  };
  int ns_run = 1;
  for (int i = 0; i < hp.p.npass; i++) {
    hp.p.nsh[i] = log2_if_pow2(ns_run);
    hp.p.nbsh[i] = log2_if_pow2(L / rad[i]);
    ns_run *= rad[i];
    hp.p.radix[i] = rad[i];
    const int tot = ncomp * L / rad[i];
    int b = block ? (tot + block - 1) / block : 8;
    if (b > 4) b = 8;
    hp.p.bpt[i] = b;
  }
  std::vector<V2<T>> tw(L);
  for (int q = 0; q < L; q++) {
    const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)L;
    tw[q] = mk<T>((T)cosl(ang), (T)sinl(ang));
  }
  if (hipMalloc(&hp.d_tw, sizeof(V2<T>) * L) != hipSuccess) { derr("hipMalloc twiddles"); return false; }
  hipMemcpy(hp.d_tw, tw.data(), sizeof(V2<T>) * L, hipMemcpyHostToDevice);
  hp.p.tw = hp.d_tw;
  return true;
}

// workgroup size override (multiple of 64, <= 1024); every pass must fit
// its register budget at the resulting butterflies per thread
template <class T>
bool set_block(HostPlan<T>& hp, int block) {
  if (hp.block == 0) return block == 0;   // no Stockham form of this plan
  if (block < 64 || block > 512 || block % 64) return false;
  int bpt[kMaxPasses];
  for (int i = 0; i < hp.p.npass; i++) {
    const int r = hp.p.radix[i];
    int b = (hp.ncomp * hp.p.L / r + block - 1) / block;
    if (b > 4) b = 8;
    if (b > max_b(r)) return false;
    bpt[i] = b;
  }
  for (int i = 0; i < hp.p.npass; i++) hp.p.bpt[i] = bpt[i];
  hp.block = block;
  return true;
}

template <class T>
void free_plan(HostPlan<T>& hp) {
  if (hp.d_tw) hipFree(hp.d_tw);
  hp.d_tw = nullptr;
}

void tensor_at_host(double thick, int dx, int dy, double out9[9]) {
  const double a = 0.49999, b = 0.5 * thick;
  double dm[10];
  for (int c = 0; c < 9; c++) out9[c] = 0.0;
  for (int jy = -4; jy <= 4; jy++) {
    const double sy = (double)dy + 0.1 * (double)jy;
    for (int ix = -4; ix <= 4; ix++) {
      const double sx = (double)dx + 0.1 * (double)ix;
      ctt<true>(b, a, sx, sy, dm);
      for (int c = 0; c < 9; c++) out9[c] += dm[c + 1];
    }
  }
  for (int c = 0; c < 9; c++) out9[c] /= 81.;
}

}  // namespace

// ================================================================= state
struct DemagData {
  int nx, ny, px, py, hx, padded;
  double strength;
  double n0[3];
  hipStream_t stream;
  HostPlan<double> row, col, col1;
  HostPlan<float> row32, col32;
  double2* d_post;    // exp(-2 pi i k / Px), k in [0, Px/2]
  float2* d_post32;
  double2* d_spec;    // 3 * ny * hx complex (FP32 pipeline reuses it as float2)
  double* d_K;        // 5 planes [hx][py]
  float* d_K32;
  int has_xy_im;
  int fp32;
  int ct = 1;         // compile-time FFT plans where one matches the plan
  // register-resident FFT kernels (demag_reg_kernels.h): variant per pass,
  // -1 = the LDS Stockham kernels of this file
  int reg_row = 0;
  int reg_col = 0;
};

namespace {

bool launch_ok(const char* what) {
  const hipError_t e = hipPeekAtLastError();
  if (e != hipSuccess) {
    derr(std::string(what) + ": " + hipGetErrorString(e));
    return false;
  }
  return true;
}

void destroy(DemagData* d) {
  if (!d) return;
  free_plan(d->row);
  free_plan(d->col);
  free_plan(d->col1);
  free_plan(d->row32);
  free_plan(d->col32);
  hipFree(d->d_post);
  hipFree(d->d_post32);
  hipFree(d->d_spec);
  hipFree(d->d_K);
  hipFree(d->d_K32);
  delete d;
}

DemagData* init_common(int nx, int ny, double thick, double strength, int padded) {
  g_derr.clear();
  if (nx <= 0 || ny <= 0) { derr("Demag: bad grid"); return nullptr; }
  DemagData* d = new DemagData();
  d->nx = nx; d->ny = ny; d->padded = padded; d->strength = strength;
  d->px = padded ? 2 * nx : nx;
  d->py = padded ? 2 * ny : ny;
  if (d->px % 2) { derr("Demag: FFT row length must be even"); delete d; return nullptr; }
  d->hx = d->px / 2 + 1;
  d->stream = (hipStream_t)NVHip_DefaultStream();
  d->has_xy_im = padded ? 0 : 1;
  // self coupling N(0) on the host (2d_i4/demag_fft.cu:230-239)
  double t9[9];
  tensor_at_host(thick, 0, 0, t9);
  d->n0[0] = t9[0] * strength;
  d->n0[1] = t9[4] * strength;
  d->n0[2] = t9[8] * strength;
  if (!make_plan(d->px / 2, 1, d->row) || !make_plan(d->py, 3, d->col) ||
      !make_plan(d->py, 1, d->col1)) {
    destroy(d);
    return nullptr;
  }
  // the tensor spectra use the one-component Stockham transforms; the field
  // needs a row and a column kernel of either family
  if (d->row.block == 0 || d->col1.block == 0 ||
      (d->col.block == 0 && !demag_reg::has_col_plan<double>(d->py, 0))) {
    derr("Demag: FFT lengths " + std::to_string(d->px / 2) + " x " + std::to_string(d->py) +
         " exceed one workgroup");
    destroy(d);
    return nullptr;
  }
  const int Px = d->px, Py = d->py, Hx = d->hx;
  std::vector<double2> post(Hx);
  for (int k = 0; k < Hx; k++) {
    const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)Px;
    post[k] = make_double2((double)cosl(ang), (double)sinl(ang));
  }
  const size_t spec_elems = std::max((size_t)3 * ny * Hx, (size_t)Py * Hx);
  if (hipMalloc(&d->d_post, sizeof(double2) * Hx) != hipSuccess ||
      hipMalloc(&d->d_spec, sizeof(double2) * spec_elems) != hipSuccess ||
      hipMalloc(&d->d_K, sizeof(double) * 5 * (size_t)Hx * Py) != hipSuccess) {
    derr("Demag: device allocation failed");
    destroy(d);
    return nullptr;
  }
  hipMemcpy(d->d_post, post.data(), sizeof(double2) * Hx, hipMemcpyHostToDevice);
  hipMemset(d->d_K, 0, sizeof(double) * 5 * (size_t)Hx * Py);

  // tensor quadrant on the GPU
  const int qx = padded ? nx : Px / 2 + 1;
  const int qy = padded ? ny : Py / 2 + 1;
  double *dQ = nullptr, *dKr = nullptr;
  if (hipMalloc(&dQ, sizeof(double) * 4 * (size_t)qx * qy) != hipSuccess ||
      hipMalloc(&dKr, sizeof(double) * 4 * (size_t)Px * Py) != hipSuccess) {
    derr("Demag: tensor scratch allocation failed");
    hipFree(dQ);
    destroy(d);
    return nullptr;
  }
  hipLaunchKernelGGL(k_tensor_quadrant, dim3((qx + 63) / 64, qy), dim3(64), 0, d->stream, thick,
                     qx, qy, dQ);
  hipLaunchKernelGGL(k_tensor_fill, dim3((Px + 255) / 256, Py), dim3(256), 0, d->stream, dQ, qx,
                     qy, Px, Py, padded, dKr);
  // spectra of the 4 real planes: row R2C + column forward FFT
  const double scale = strength / ((double)Px * (double)Py);
  if (!allow_lds(k_row_fwd<double, GenFft<double>>,
                 sizeof(double2) * lpad_len(d->row.p.L, d->row.p.psh)) ||
      !allow_lds(k_col<double, GenFft<double>>, sizeof(double2) * lpad_len(Py, d->col1.p.psh))) {
    derr("Demag: LDS attribute");
    hipFree(dQ);
    hipFree(dKr);
    destroy(d);
    return nullptr;
  }
  for (int c = 0; c < 4; c++) {
    hipLaunchKernelGGL((k_row_fwd<double, GenFft<double>>), dim3(Py, 1), dim3(d->row.block),
                       sizeof(double2) * lpad_len(d->row.p.L, d->row.p.psh), d->stream,
                       (const double*)(dKr + (size_t)c * Px * Py), (const double*)nullptr,
                       (long)0, Px, Py, GenFft<double>{d->row.p}, (const V2<double>*)d->d_post,
                       (V2<double>*)d->d_spec);
    hipLaunchKernelGGL((k_col<double, GenFft<double>>), dim3(Hx), dim3(d->col1.block),
                       sizeof(double2) * lpad_len(Py, d->col1.p.psh), d->stream,
                       (V2<double>*)d->d_spec, Py, Py, Hx, GenFft<double>{d->col1.p},
                       (const double*)nullptr, 0, 1);
    const long P = (long)Hx * Py;
    hipLaunchKernelGGL(k_extract, dim3((Hx + 255) / 256, Py), dim3(256), 0, d->stream,
                       (const double2*)d->d_spec, Py, Hx, scale, 0, d->d_K + c * P);
    if (c == 3 && d->has_xy_im)
      hipLaunchKernelGGL(k_extract, dim3((Hx + 255) / 256, Py), dim3(256), 0, d->stream,
                         (const double2*)d->d_spec, Py, Hx, scale, 1, d->d_K + 4 * P);
  }
  if (padded && Py / 2 - 1 > 0)
    hipLaunchKernelGGL(k_mirror_ky, dim3((Hx + 255) / 256, Py / 2 - 1, 4), dim3(256), 0, d->stream,
                       d->d_K, Hx, Py);
  hipStreamSynchronize(d->stream);
  hipFree(dQ);
  hipFree(dKr);
  if (!launch_ok("Demag_Init")) { destroy(d); return nullptr; }
  // register-resident FFT kernels wherever a plan covers the lengths
  d->reg_row = demag_reg::has_row_plan<double>(d->px / 2, 0) ? 0 : -1;
  // columns: one transform per wave (variant 5) where it exists, else the
  // blocked form (4 columns per workgroup, variant 0's bits), else variant 0
  d->reg_col = demag_reg::has_col_plan<double>(d->py, 5) ? 5
             : demag_reg::has_col_plan<double>(d->py, 3) ? 3
             : demag_reg::has_col_plan<double>(d->py, 0) ? 0 : -1;
  return d;
}

// compile-time plans of the benchmarked shapes, each at the default
// workgroup size and LDS padding of its plan: the zero-padded 3072 x 1024
// grid (row transforms of 1536 points, columns of 1024) and the periodic
// 512 x 512 grid (rows of 256, columns of 512)
template <class T> using CtRow1536 = CtFft<T, 1536, 256, 3, 8, 8, 8, 3>;
template <class T> using CtRow256 = CtFft<T, 256, 64, 3, 8, 8, 4>;
template <class T> using CtCol1024 = CtFft<T, 1024, 256, 4, 8, 8, 4, 4>;
template <class T> using CtCol512 = CtFft<T, 512, 192, 4, 8, 8, 8>;

template <class F, class T>
bool ct_fits(const HostPlan<T>& hp) {
  if (hp.p.L != F::kL || hp.block != F::kBlock || hp.p.psh != F::kSh || hp.p.npass != F::kNpass)
    return false;
  for (int i = 0; i < F::kNpass; i++)
    if (hp.p.radix[i] != F::kRad[i]) return false;
  return true;
}

// calls fn with the FFT policy for a row plan (compile-time when one fits)
template <class T, class Fn>
void with_row_plan(const HostPlan<T>& hp, bool ct, Fn&& fn) {
  if (ct && ct_fits<CtRow1536<T>>(hp)) fn(CtRow1536<T>{hp.p.tw});
  else if (ct && ct_fits<CtRow256<T>>(hp)) fn(CtRow256<T>{hp.p.tw});
  else fn(GenFft<T>{hp.p});
}

template <class T, class Fn>
void with_col_plan(const HostPlan<T>& hp, bool ct, Fn&& fn) {
  if (ct && ct_fits<CtCol1024<T>>(hp)) fn(CtCol1024<T>{hp.p.tw});
  else if (ct && ct_fits<CtCol512<T>>(hp)) fn(CtCol512<T>{hp.p.tw});
  else fn(GenFft<T>{hp.p});
}

template <class T>
void apply_t(DemagData* d, const HostPlan<T>& row, const HostPlan<T>& col, const V2<T>* post,
             const T* K, const double* y, const double* w, double* h) {
  const int nx = d->nx, ny = d->ny, Hx = d->hx, Py = d->py;
  const long N = (long)nx * ny;
  hipStream_t s = d->stream;
  V2<T>* spec = reinterpret_cast<V2<T>*>(d->d_spec);
  const size_t row_lds = sizeof(V2<T>) * lpad_len(row.p.L, row.p.psh);
  const int half = d->padded;
  kprof::begin(kprof::DEMAG_ALL, s);
  kprof::begin(kprof::DEMAG_ROW_FWD, s);
  int rc = d->reg_row < 0 ? 0
                          : demag_reg::launch_row_fwd<T>(row.p.L, d->reg_row, half, y, w, N, nx,
                                                          ny, row.p.tw, post, spec, s);
  if (rc == 0)
    with_row_plan(row, d->ct, [&](auto f) {
      allow_lds(k_row_fwd<T, decltype(f)>, row_lds);
      hipLaunchKernelGGL((k_row_fwd<T, decltype(f)>), dim3(ny, 3), dim3(row.block), row_lds, s,
                         y, w, N, nx, ny, f, post, spec);
    });
  kprof::end(kprof::DEMAG_ROW_FWD, s);
  // the column kernel's timing events ride on its dispatch
  hipEvent_t ce0, ce1;
  kprof::ext_pair(kprof::DEMAG_COL, &ce0, &ce1);
  rc = d->reg_col < 0 ? 0
                      : demag_reg::launch_col<T>(Py, d->reg_col, half, spec, Hx, col.p.tw, K,
                                                  d->has_xy_im, s, ce0, ce1);
  if (rc == 0)
    with_col_plan(col, d->ct, [&](auto f) {
      const size_t col_lds = sizeof(V2<T>) * 3 * lpad_len(Py, col.p.psh);
      allow_lds(k_col<T, decltype(f)>, col_lds);
      hipExtLaunchKernelGGL((k_col<T, decltype(f)>), dim3(Hx), dim3(col.block),
                            (std::uint32_t)col_lds, s, ce0, ce1, 0u, spec, ny, ny, Hx, f, K,
                            d->has_xy_im, 0);
    });
  kprof::begin(kprof::DEMAG_ROW_INV, s);
  rc = d->reg_row < 0 ? 0
                      : demag_reg::launch_row_inv<T>(row.p.L, d->reg_row, half,
                                                      (const V2<T>*)spec, ny, row.p.tw, post, h,
                                                      N, nx, s);
  if (rc == 0)
    with_row_plan(row, d->ct, [&](auto f) {
      allow_lds(k_row_inv<T, decltype(f)>, row_lds);
      hipLaunchKernelGGL((k_row_inv<T, decltype(f)>), dim3(ny, 3), dim3(row.block), row_lds, s,
                         (const V2<T>*)spec, ny, f, post, h, N, nx);
    });
  kprof::end(kprof::DEMAG_ROW_INV, s);
  kprof::end(kprof::DEMAG_ALL, s);
  launch_ok("Demag_Apply");
}

void apply(DemagData* d, const double* y, const double* w, double* h) {
  if (d->fp32)
    apply_t<float>(d, d->row32, d->col32, (const V2<float>*)d->d_post32, d->d_K32, y, w, h);
  else
    apply_t<double>(d, d->row, d->col, (const V2<double>*)d->d_post, d->d_K, y, w, h);
}

}  // namespace

extern "C" {

const char* Demag_LastError(void) { return g_derr.c_str(); }

DemagData* Demag_Init(int nx, int ny, double thick, double demag_strength) {
  return init_common(nx, ny, thick, demag_strength, 0);
}

DemagData* Demag_InitPadded(int nx, int ny, double thick, double demag_strength) {
  return init_common(nx, ny, thick, demag_strength, 1);
}

void Demag_Apply(DemagData* d, const double* y_dev, double* h_out_dev) {
  if (!d) return;
  g_derr.clear();
  apply(d, y_dev, nullptr, h_out_dev);
}

void Demag_ApplyWindowed(DemagData* d, const double* y_dev, const double* w_dev, double* h_out_dev) {
  if (!d) return;
  g_derr.clear();
  apply(d, y_dev, w_dev, h_out_dev);
}

int Demag_ApplyBatch(int n, DemagData* const* d, const double* const* y,
                     const double* const* w, double* const* h, void* stream) {
  if (n < 0 || (n > 0 && (!d || !y || !h))) {
    derr("Demag_ApplyBatch: NULL argument");
    return -1;
  }
  g_derr.clear();
  hipStream_t s = (hipStream_t)stream;
  // instances share a launch when everything a pass kernel's shape and
  // arithmetic depend on is equal
  auto same_plan = [](const DemagData* a, const DemagData* b) {
    return a->nx == b->nx && a->ny == b->ny && a->px == b->px && a->py == b->py &&
           a->hx == b->hx && a->padded == b->padded && a->has_xy_im == b->has_xy_im &&
           a->reg_row == b->reg_row && a->reg_col == b->reg_col && !a->fp32 && !b->fp32;
  };
  std::vector<char> done(n, 0);
  for (int i = 0; i < n; i++) {
    if (done[i]) continue;
    DemagData* a = d[i];
    if (!a || !y[i] || !h[i]) {
      derr("Demag_ApplyBatch: NULL instance");
      return -1;
    }
    done[i] = 1;
    const bool reg = !a->fp32 && a->reg_row >= 0 && a->reg_col >= 0 &&
                     demag_reg::has_row_plan<double>(a->px / 2, a->reg_row) &&
                     demag_reg::has_col_plan<double>(a->py, a->reg_col);
    if (!reg) {   // its own passes, on the batch's stream (its own stream kept)
      hipStream_t own = a->stream;
      a->stream = s;
      apply(a, y[i], w ? w[i] : nullptr, h[i]);
      a->stream = own;
      if (!g_derr.empty()) return -1;
      continue;
    }
    std::vector<int> grp{i};
    for (int j = i + 1; j < n; j++)
      if (!done[j] && d[j] && same_plan(a, d[j])) {
        grp.push_back(j);
        done[j] = 1;
      }
    const long N = (long)a->nx * a->ny;
    const auto* tw_row = a->row.p.tw;
    const auto* tw_col = a->col.p.tw;
    const auto* post = (const V2<double>*)a->d_post;
    for (size_t c0 = 0; c0 < grp.size(); c0 += demag_reg::kBatchMax) {
      const int m = (int)std::min<size_t>(demag_reg::kBatchMax, grp.size() - c0);
      demag_reg::RegBatch<double> b{};
      for (int k = 0; k < m; k++) {
        const int j = grp[c0 + k];
        if (!y[j] || !h[j]) {
          derr("Demag_ApplyBatch: NULL instance");
          return -1;
        }
        b.in[k] = y[j];
        b.w[k] = w ? w[j] : nullptr;
        b.spec[k] = reinterpret_cast<V2<double>*>(d[j]->d_spec);
        b.K[k] = d[j]->d_K;
        b.out[k] = h[j];
      }
      if (demag_reg::launch_row_fwd_batch<double>(a->row.p.L, a->reg_row, a->padded, m, b, N,
                                                   a->nx, a->ny, tw_row, post, s) != 1 ||
          demag_reg::launch_col_batch<double>(a->py, a->reg_col, a->padded, m, b, a->hx, tw_col,
                                               a->has_xy_im, s) != 1 ||
          demag_reg::launch_row_inv_batch<double>(a->row.p.L, a->reg_row, a->padded, m, b,
                                                   a->ny, tw_row, post, N, a->nx, s) != 1) {
        if (g_derr.empty()) derr("Demag_ApplyBatch: a batched pass failed to launch");
        return -1;
      }
    }
  }
  return launch_ok("Demag_ApplyBatch") ? 0 : -1;
}

void Demag_SetStream(DemagData* d, void* stream) {
  if (!d) return;
  d->stream = stream ? (hipStream_t)stream : (hipStream_t)NVHip_DefaultStream();
}

void Demag_GetSelfCoupling(DemagData* d, double* nxx0, double* nyy0, double* nzz0) {
  if (!d) {
    if (nxx0) *nxx0 = 0.0;
    if (nyy0) *nyy0 = 0.0;
    if (nzz0) *nzz0 = 0.0;
    return;
  }
  if (nxx0) *nxx0 = d->n0[0];
  if (nyy0) *nyy0 = d->n0[1];
  if (nzz0) *nzz0 = d->n0[2];
}

void Demag_Destroy(DemagData* d) { destroy(d); }

int Demag_SetPrecision(DemagData* d, int fp32) {
  if (!d) return -1;
  g_derr.clear();
  if (!fp32) { d->fp32 = 0; return 0; }
  if (!d->d_K32) {
    const int Hx = d->hx, Py = d->py;
    const size_t P5 = 5 * (size_t)Hx * Py;
    if (!make_plan(d->px / 2, 1, d->row32) || !make_plan(d->py, 3, d->col32)) return -1;
    d->row32.p.psh = d->row.p.psh;
    d->col32.p.psh = d->col.p.psh;
    std::vector<float2> post(Hx);
    for (int k = 0; k < Hx; k++) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)d->px;
      post[k] = make_float2((float)cosl(ang), (float)sinl(ang));
    }
    if (hipMalloc(&d->d_post32, sizeof(float2) * Hx) != hipSuccess ||
        hipMalloc(&d->d_K32, sizeof(float) * P5) != hipSuccess) {
      derr("Demag_SetPrecision: allocation failed");
      return -1;
    }
    hipMemcpy(d->d_post32, post.data(), sizeof(float2) * Hx, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_to_float, dim3(1024), dim3(256), 0, d->stream, (const double*)d->d_K,
                       d->d_K32, (long)P5);
    hipStreamSynchronize(d->stream);
    if (!launch_ok("Demag_SetPrecision")) return -1;
  }
  d->fp32 = 1;
  return 0;
}

int Demag_SetWorkgroups(DemagData* d, int row_block, int col_block) {
  if (!d) return -1;
  if (row_block && (!set_block(d->row, row_block) ||
                    (d->d_K32 && !set_block(d->row32, row_block))))
    return -1;
  if (col_block && (!set_block(d->col, col_block) ||
                    (d->d_K32 && !set_block(d->col32, col_block))))
    return -1;
  return 0;
}

int Demag_SetSpecialized(DemagData* d, int on) {
  if (!d) return -1;
  d->ct = on ? 1 : 0;
  int n = 0;
  auto count = [&](auto f) { n += decltype(f)::kCompileTime ? 1 : 0; };
  if (d->fp32) {
    with_row_plan(d->row32, d->ct, count);
    with_col_plan(d->col32, d->ct, count);
  } else {
    with_row_plan(d->row, d->ct, count);
    with_col_plan(d->col, d->ct, count);
  }
  return n;
}

int Demag_SetLdsPadding(DemagData* d, int row_shift, int col_shift) {
  if (!d || row_shift < 1 || row_shift > 31 || col_shift < 1 || col_shift > 31) return -1;
  d->row.p.psh = d->row32.p.psh = row_shift;
  d->col.p.psh = d->col32.p.psh = col_shift;
  return 0;
}

int Demag_SetFftKernels(DemagData* d, int row_variant, int col_variant) {
  if (!d) return -1;
  const int Lr = d->px / 2, Lc = d->py;
  if (row_variant >= 0 && !demag_reg::has_row_plan<double>(Lr, row_variant)) return -1;
  if (col_variant >= 0 && !demag_reg::has_col_plan<double>(Lc, col_variant)) return -1;
  // the Stockham kernels exist only for lengths that fit one workgroup
  if (row_variant < 0 && (d->row.block == 0 || (d->d_K32 && d->row32.block == 0))) return -1;
  if (col_variant < 0 && (d->col.block == 0 || (d->d_K32 && d->col32.block == 0))) return -1;
  d->reg_row = row_variant < 0 ? -1 : row_variant;
  d->reg_col = col_variant < 0 ? -1 : col_variant;
  return 0;
}

int Demag_GetFftKernels(DemagData* d, int* row_variant, int* col_variant) {
  if (!d) return -1;
  if (row_variant) *row_variant = d->reg_row;
  if (col_variant) *col_variant = d->reg_col;
  return 0;
}

// read-modify-write of a buffer larger than the L2 and last-level caches
__global__ void k_evict(double* __restrict__ a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    a[i] = a[i] + 1.0;
}

int Demag_Autotune(DemagData* d, int reps, int* choice) {
  if (!d || reps < 1) return -1;
  const long n3 = 3L * d->nx * d->ny;
  double *y = nullptr, *h = nullptr, *ev = nullptr;
  if (hipMalloc(&y, sizeof(double) * n3) != hipSuccess) return -1;
  if (hipMalloc(&h, sizeof(double) * n3) != hipSuccess) { hipFree(y); return -1; }
  // each timed apply starts with cold caches, as inside an integration step
  // (where the vector passes between two demag applies evict the tensor and
  // the spectrum): warm back-to-back applies favour the kernels that gain
  // most from cache hits
  const long nev = 48L << 20;   // 384 MB
  if (hipMalloc(&ev, sizeof(double) * nev) != hipSuccess) {
    hipFree(y);
    hipFree(h);
    return -1;
  }
  hipMemsetAsync(ev, 0, sizeof(double) * nev, d->stream);
  hipMemsetAsync(y, 0, sizeof(double) * n3, d->stream);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const int rb0 = d->row.block, cb0 = d->col.block;
  const int rp0 = d->row.p.psh, cp0 = d->col.p.psh;
  const int rv0 = d->reg_row, cv0 = d->reg_col;
  // candidates: the current settings first, then the usual alternatives
  auto cands = [](int cur, std::initializer_list<int> alt) {
    std::vector<int> v{cur};
    for (int a : alt)
      if (a != cur) v.push_back(a);
    return v;
  };
  std::vector<int> row_vars, col_vars;
  for (int v : cands(rv0, {0, -1}))
    if (v < 0 ? d->row.block > 0 && (!d->d_K32 || d->row32.block > 0)
              : demag_reg::has_row_plan<double>(d->px / 2, v))
      row_vars.push_back(v);
  for (int v : cands(cv0, {0, 1, 2, 3, 4, 5, 6, -1}))
    if (v < 0 ? d->col.block > 0 && (!d->d_K32 || d->col32.block > 0)
              : demag_reg::has_col_plan<double>(d->py, v))
      col_vars.push_back(v);
  const std::vector<int> row_blocks = cands(rb0, {64, 128, 192, 256, 320});
  const std::vector<int> col_blocks = cands(cb0, {128, 192, 256, 320});
  const std::vector<int> row_pads = cands(rp0, {3, 31}), col_pads = cands(cp0, {4, 31});
  float best = 1e30f;
  int best_cfg[6] = {rb0, cb0, rp0, cp0, rv0, cv0};
  auto time_it = [&]() {
    apply(d, y, nullptr, h);   // warm
    float total = 0.f;
    for (int i = 0; i < reps; i++) {
      hipLaunchKernelGGL(k_evict, dim3(4096), dim3(256), 0, d->stream, ev, nev);
      hipEventRecord(e0, d->stream);
      apply(d, y, nullptr, h);
      hipEventRecord(e1, d->stream);
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      total += ms;
    }
    return total;
  };
  for (int rv : row_vars)
    for (int cv : col_vars) {
      d->reg_row = rv;
      d->reg_col = cv;
      // workgroup sizes and LDS paddings only matter to the Stockham kernels
      const std::vector<int> rbs = rv < 0 ? row_blocks : std::vector<int>{rb0};
      const std::vector<int> rps = rv < 0 ? row_pads : std::vector<int>{rp0};
      const std::vector<int> cbs = cv < 0 ? col_blocks : std::vector<int>{cb0};
      const std::vector<int> cps = cv < 0 ? col_pads : std::vector<int>{cp0};
      for (int rb : rbs)
        for (int cb : cbs)
          for (int rp : rps)
            for (int cp : cps) {
              if (!set_block(d->row, rb) || !set_block(d->col, cb)) continue;
              d->row.p.psh = rp;
              d->col.p.psh = cp;
              const float ms = time_it();
              if (ms < best) {
                best = ms;
                best_cfg[0] = rb; best_cfg[1] = cb; best_cfg[2] = rp; best_cfg[3] = cp;
                best_cfg[4] = rv; best_cfg[5] = cv;
              }
            }
    }
  set_block(d->row, best_cfg[0]);
  set_block(d->col, best_cfg[1]);
  d->row.p.psh = best_cfg[2];
  d->col.p.psh = best_cfg[3];
  d->reg_row = best_cfg[4];
  d->reg_col = best_cfg[5];
  if (d->d_K32) {
    set_block(d->row32, best_cfg[0]);
    set_block(d->col32, best_cfg[1]);
    d->row32.p.psh = best_cfg[2];
    d->col32.p.psh = best_cfg[3];
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(y);
  hipFree(h);
  hipFree(ev);
  if (choice)
    for (int i = 0; i < 6; i++) choice[i] = best_cfg[i];
  return launch_ok("Demag_Autotune") ? 0 : -1;
}

int Demag_GetPlanShape(DemagData* d, int* px, int* py, int* padded) {
  if (!d) return -1;
  if (px) *px = d->px;
  if (py) *py = d->py;
  if (padded) *padded = d->padded;
  return 0;
}

int Demag_GetSpectrum(DemagData* d, double* h_planes) {
  if (!d || !h_planes) return -1;
  const int Hx = d->hx, Py = d->py;
  const size_t P = (size_t)Hx * Py;
  std::vector<double> tmp(5 * P);
  hipStreamSynchronize(d->stream);
  if (hipMemcpy(tmp.data(), d->d_K, sizeof(double) * 5 * P, hipMemcpyDeviceToHost) != hipSuccess)
    return -1;
  for (int c = 0; c < 5; c++)
    for (int ky = 0; ky < Py; ky++)
      for (int kx = 0; kx < Hx; kx++)
        h_planes[c * P + (size_t)ky * Hx + kx] = tmp[c * P + (size_t)kx * Py + ky];
  return 0;
}

int Demag_NewellTensor(double thick, int mdx, int mdy, double* planes9, int nthreads) {
  if (mdx <= 0 || mdy <= 0 || !planes9) return -1;
  const int mdx2 = mdx / 2, mdy2 = mdy / 2;
  const size_t plane = (size_t)mdx * mdy;
  int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++) {
    pool.emplace_back([=]() {
      double t9[9];
      for (int j = t; j < mdy; j += nt)
        for (int i = 0; i < mdx; i++) {
          tensor_at_host(thick, i - mdx2, j - mdy2, t9);
          for (int c = 0; c < 9; c++) planes9[c * plane + (size_t)j * mdx + i] = t9[c];
        }
    });
  }
  for (auto& th : pool) th.join();
  return 0;
}

}  // extern "C"
