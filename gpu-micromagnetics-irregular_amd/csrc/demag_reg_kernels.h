// demag_reg_kernels.h -- the three demag passes on the register-resident
// FFT (fft_regs.h).  Included by one translation unit per precision.
//
// Pipeline (same spectrum layout and tensor planes as the LDS Stockham
// kernels of demag_fft.hip):
//   row forward : lane u of a (row, component) transform loads the complex
//                 pairs z[m] = (x[2m], x[2m+1]) of the gathered row for
//                 m = u + U b straight into registers (coalesced), runs the
//                 forward FFT of length L = Px/2, and the real-FFT split
//                 writes bins 0..L of the row.
//   column      : two workgroups per column.  In one, lane u loads rows
//                 u + U b of m_z, runs its forward FFT, multiplies by K_zz,
//                 runs the inverse and stores; the other does the same for
//                 m_x and m_y as a pair, with the xy block of the tensor
//                 applied in registers (each lane holds both components at
//                 the same ky).  Only the rows the field needs are stored.
//                 The blocked form (k_col_cb) moves several adjacent
//                 columns per workgroup through LDS; the wave form
//                 (k_col_wave, the default for 1024-point columns) runs
//                 each component's transform on one wave.
//   row inverse : the C2R pre-processing reads bins k and L-k straight from
//                 the spectrum into the registers of the last FFT stage,
//                 the inverse FFT ends in lane u holding z[u + U b], stored
//                 as pairs into the SoA field (overwrite).
// Zero-padded grids (half): rows and columns load only their non-zero first
// half and store only the first half of the inverse outputs.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <type_traits>

#include "demag_reg.h"
#include "fft_regs.h"

// no implicit multiply-add contraction: every fused multiply-add is an
// explicit fma, so the rounding of each pass is fixed by its source
#pragma clang fp contract(off)

namespace demag_reg {

using rfft::cadd;
using rfft::cconj;
using rfft::cmul;
using rfft::cmulc;
using rfft::csub;
using rfft::fmad;
using rfft::ldc;
using rfft::mk;
using rfft::stc;
using rfft::sfor;

// consecutive workgroup ids -> the same XCD (of 8) gets a contiguous range
__device__ inline int reg_xcd_remap(int bid, int G) {
  const int per = G / 8, rem = G % 8;
  const int x = bid % 8, idx = bid / 8;
  return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

struct Bar {
  __device__ __forceinline__ void operator()() const { __syncthreads(); }
};

// the batch argument of a pass kernel: the instances' buffers (RegBatch,
// blockIdx.y selects the instance) or nothing for a single launch
struct NoBatch {};
template <class T, bool B>
using BatchArg = std::conditional_t<B, RegBatch<T>, NoBatch>;

// ------------------------------------------------------------ row forward
template <class T, class F, bool HALF, bool BATCH = false>
__global__ __launch_bounds__(F::kThreads) void k_row_fwd_reg(
    const double* __restrict__ in, const double* __restrict__ w, long comp_stride, int nx_in,
    int rows, const C2<T>* __restrict__ tw, const C2<T>* __restrict__ post,
    C2<T>* __restrict__ spec, BatchArg<T, BATCH> bt) {
  if constexpr (BATCH) {
    in = bt.in[blockIdx.y];
    w = bt.w[blockIdx.y];
    spec = bt.spec[blockIdx.y];
  }
  using C = C2<T>;
  constexpr int U = F::U, V = F::V, L = F::L, G = F::kPerGroup;
  constexpr int NB = HALF ? V / 2 : V;
  extern __shared__ char lds_raw[];
  const int lane = threadIdx.x & (U - 1);
  const int slot = threadIdx.x / U;
  const int total = 3 * rows;
  int id = reg_xcd_remap(blockIdx.x, gridDim.x) * G + slot;
  const bool live = id < total;
  if (!live) id = total - 1;
  const int comp = id / rows, row = id - comp * rows;
  C* lds = reinterpret_cast<C*>(lds_raw) + slot * F::kSlots;
  const double* src = in + comp * comp_stride + (long)row * nx_in;
  const double* wr = w ? w + (long)row * nx_in : nullptr;
  C r[V];
  sfor<0, NB>([&](auto B) {
    const int m = lane + U * B;
    double2 v = *reinterpret_cast<const double2*>(src + 2 * m);
    if (wr) {
      const double2 ww = *reinterpret_cast<const double2*>(wr + 2 * m);
      v.x = ww.x * v.x;
      v.y = ww.y * v.y;
    }
    r[B] = mk<T>((T)v.x, (T)v.y);
  });
  F::template forward<HALF>(r, lds, lane, tw, Bar{});
  using Last = typename F::Last;
  __syncthreads();
  Last::put_out(lds, r, lane);   // natural order: slot(k) = k + k / V
  __syncthreads();
  if (!live) return;
  C* out = spec + ((long)comp * rows + row) * (L + 1);
  auto split = [&](C zk, C zc, C p) {
    const C e = mk<T>(T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y));
    const C d = csub(zk, zc);
    const C o = mk<T>(T(0.5) * d.y, T(-0.5) * d.x);
    return cadd(e, cmul(p, o));
  };
  sfor<0, V>([&](auto I) {
    const int k = lane + U * I;
    const int km = k == 0 ? 0 : L - k;
    const C zk = lds[k + k / V];
    const C zc = cconj(lds[km + km / V]);
    stc(out + k, split(zk, zc, ldc(post + k)));
  });
  if (lane == 0) {
    const C z0 = lds[0];
    stc(out + L, split(z0, cconj(z0), ldc(post + L)));
  }
}

// ------------------------------------------------------------ row inverse
template <class T, class F, bool HALF, bool BATCH = false>
__global__ __launch_bounds__(F::kThreads) void k_row_inv_reg(
    const C2<T>* __restrict__ spec, int rows, const C2<T>* __restrict__ tw,
    const C2<T>* __restrict__ post, double* __restrict__ out, long comp_stride, int nx_out,
    BatchArg<T, BATCH> bt) {
  if constexpr (BATCH) {
    spec = bt.spec[blockIdx.y];
    out = bt.out[blockIdx.y];
  }
  using C = C2<T>;
  constexpr int U = F::U, V = F::V, L = F::L, G = F::kPerGroup;
  constexpr int NB = HALF ? V / 2 : V;
  using Last = typename F::Last;
  extern __shared__ char lds_raw[];
  const int lane = threadIdx.x & (U - 1);
  const int slot = threadIdx.x / U;
  const int total = 3 * rows;
  int id = reg_xcd_remap(blockIdx.x, gridDim.x) * G + slot;
  const bool live = id < total;
  if (!live) id = total - 1;
  const int comp = id / rows, row = id - comp * rows;
  C* lds = reinterpret_cast<C*>(lds_raw) + slot * F::kSlots;
  const C* in = spec + ((long)comp * rows + row) * (L + 1);
  C r[V];
  sfor<0, Last::NG>([&](auto J) {
    sfor<0, Last::Q>([&](auto Tt) {
      const int k = Last::k_of(lane, J, Tt);
      const C xk = ldc(in + k);
      const C xc = cconj(ldc(in + (L - k)));
      const C a = cadd(xk, xc);
      const C b = cmulc(csub(xk, xc), ldc(post + k));
      r[J * Last::Q + Tt] = mk<T>(a.x - b.y, a.y + b.x);   // a + i b
    });
  });
  F::template inverse<HALF>(r, lds, lane, tw, Bar{});
  if (!live) return;
  double* dst = out + comp * comp_stride + (long)row * nx_out;
  sfor<0, NB>([&](auto B) {
    const int m = lane + U * B;
    double2 v;
    v.x = (double)r[B].x;
    v.y = (double)r[B].y;
    *reinterpret_cast<double2*>(dst + 2 * m) = v;
  });
}

// ------------------------------------------------------------------ column
// waves per SIMD the column kernel asks the register allocator for: the
// FP64 plans with 8 values per lane fit four (128 VGPRs), which lets a CU
// hold 8 columns of 1024 rows at once; otherwise no constraint
template <class T, class F>
struct ColWaves {
  static constexpr int value = (sizeof(T) == 8 && F::V == 8 && F::U >= 128) ? 4 : 1;
};
// XYI: the xy tensor plane has an imaginary part (periodic grids); the
// zero-padded grid's xy spectrum is real (odd-odd kernel)
template <class T, class F, bool HALF, bool XYI, bool BATCH = false>
__global__ __launch_bounds__(F::kThreads) __attribute__((amdgpu_waves_per_eu(ColWaves<T, F>::value))) void k_col_reg(C2<T>* __restrict__ spec, int Hx,
                                                         const C2<T>* __restrict__ tw,
                                                         const T* __restrict__ K,
                                                         BatchArg<T, BATCH> bt) {
  if constexpr (BATCH) {
    spec = bt.spec[blockIdx.y];
    K = bt.K[blockIdx.y];
  }
  using C = C2<T>;
  constexpr int U = F::U, V = F::V, L = F::L, G = F::kPerGroup;
  constexpr int NB = HALF ? V / 2 : V;
  constexpr int ROWS = HALF ? L / 2 : L;
  using Last = typename F::Last;
  constexpr int Q = Last::Q, CH = Q < 4 ? Q : 4;   // tensor values per chunk
  extern __shared__ char lds_raw[];
  const int lane = threadIdx.x & (U - 1);
  const int slot = threadIdx.x / U;
  // the grid's first half (a multiple of 8 workgroups, so that the XCD of a
  // workgroup is the same in both halves) runs the x/y pairs, the second
  // half the z components: twice the workgroups of one column each, and
  // the same arithmetic per component as one workgroup doing both
  const int half_grid = gridDim.x / 2;
  const bool zjob = (int)blockIdx.x >= half_grid;
  const int bid = zjob ? (int)blockIdx.x - half_grid : (int)blockIdx.x;
  int kx = reg_xcd_remap(bid, half_grid) * G + slot;
  const bool live = kx < Hx;
  if (!live) kx = Hx - 1;
  C* lds = reinterpret_cast<C*>(lds_raw) + slot * F::kSlots;
  // z: load, transform, multiply by K_zz (h_z = K_zz m_z), transform back,
  // store; x and y as a pair.  At most two components are live at once,
  // which keeps the FP64 plans with 8 values per lane within 128 VGPRs
  // without spilling (loading all three up front spilled).
  C r[3][V];
  auto load = [&](auto Cc) {
    const C* src = spec + ((long)Cc * ROWS + rfft::opaque(lane)) * Hx + kx;
    sfor<0, NB>([&](auto B) { r[Cc][B] = ldc(src + (long)U * B * Hx); });
  };
  auto store = [&](auto Cc) {
    if (live) {
      C* dst = spec + ((long)Cc * ROWS + rfft::opaque(lane)) * Hx + kx;
      sfor<0, NB>([&](auto B) { stc(dst + (long)U * B * Hx, r[Cc][B]); });
    }
  };
  // the tensor values this lane needs, a few at a time so the loads in
  // flight stay within the register budget (plane bases are uniform; the
  // per-lane offset is a 32-bit unsigned index)
  const long P = (long)Hx * L;
  const T* kxx = K + (long)kx * L;
  const T* kyy = kxx + P;
  const T* kzz = kyy + P;
  const T* kxyr = kzz + P;
  const T* kxyi = kxyr + P;
  auto each_k = [&](auto&& fn) {
    sfor<0, Last::NG>([&](auto J) {
      sfor<0, Q / CH>([&](auto Ch) {
        const int ln = rfft::opaque(lane);
        sfor<0, CH>([&](auto Tq) {
          constexpr int t = Ch * CH + Tq;
          fn(J * Q + t, (unsigned)Last::k_of(ln, J, t));
        });
      });
    });
  };
  constexpr auto c0 = std::integral_constant<int, 0>{};
  constexpr auto c1 = std::integral_constant<int, 1>{};
  constexpr auto c2 = std::integral_constant<int, 2>{};
  if (zjob) {
    load(c2);
    F::template forward<HALF>(r[2], lds, lane, tw, Bar{});
    each_k([&](int i, unsigned k) {
      const T cz = kzz[k];
      r[2][i] = mk<T>(cz * r[2][i].x, cz * r[2][i].y);
    });
    F::template inverse<HALF>(r[2], lds, lane, tw, Bar{});
    store(c2);
    return;
  }
  load(c0);
  load(c1);
  sfor<0, 2>([&](auto Cc) { F::template forward<HALF>(r[Cc], lds, lane, tw, Bar{}); });
  each_k([&](int i, unsigned k) {
    const T a = kxx[k], b = kyy[k], xr = kxyr[k];
    const C mx = r[0][i], my = r[1][i];
    C cxy, cyx;
    if constexpr (XYI) {
      const C xy = mk<T>(xr, kxyi[k]);
      cxy = cmul(xy, my);
      cyx = cmul(xy, mx);
    } else {
      cxy = mk<T>(xr * my.x, xr * my.y);
      cyx = mk<T>(xr * mx.x, xr * mx.y);
    }
    r[0][i] = mk<T>(fmad(a, mx.x, cxy.x), fmad(a, mx.y, cxy.y));
    r[1][i] = mk<T>(fmad(b, my.x, cyx.x), fmad(b, my.y, cyx.y));
  });
  sfor<0, 2>([&](auto Cc) {
    F::template inverse<HALF>(r[Cc], lds, lane, tw, Bar{});
    store(Cc);
  });
}

// ------------------------------------------------- column, blocked by CB
// CB adjacent columns per workgroup.  The one-column kernel above loads and
// stores one complex value per lane at a stride of a spectrum row (Hx * 16 B
// in FP64), so every wave instruction touches 64 cache lines and fills
// 16 B of each; partly written lines go back to memory before the
// neighbouring columns' workgroups complete them.  Here the spectrum rows
// move as runs of CB consecutive values (CB * 16 B): thread t reads row
// t / CB + U i of column t % CB and stages it in LDS, and each column's
// transforms then run on threads [c U, (c + 1) U) exactly as in the
// one-column kernel -- the same plan, the same values in the same lanes,
// hence the same bits.  The column stride in LDS is the plan's slot count
// padded so that the CB columns of one row land P / CB apart modulo the
// bank period of P complex values (P = 256 B / value size): each group of
// P lanes of a staging access hits distinct banks.
// HALF (zero-padded grids): the tensor planes are read for ky <= Py / 2
// only; the builder stores K(Py - ky) = K(ky) for xx, yy, zz and
// K(Py - ky) = -K(ky) for xy exactly, so the mirrored read changes no bit.
template <class T, class F, int CB>
struct ColCb {
  static constexpr int P = 256 / (int)sizeof(C2<T>);
  static constexpr int want = P / CB;
  static constexpr int CS = F::kSlots + (((want - F::kSlots % P) % P) + P) % P;
  static constexpr int kThreads = F::U * CB;
  static constexpr size_t kLds = sizeof(C2<T>) * (size_t)CB * CS;
  static constexpr bool ok = F::kPerGroup == 1 && kThreads <= 1024 && kLds <= 160 * 1024;
};

template <class T, class F, bool HALF, bool XYI, int CB, bool BATCH = false>
__global__ __launch_bounds__(F::U * CB) __attribute__((amdgpu_waves_per_eu(ColWaves<T, F>::value)))
void k_col_cb(C2<T>* __restrict__ spec, int Hx, const C2<T>* __restrict__ tw,
              const T* __restrict__ K, BatchArg<T, BATCH> bt) {
  if constexpr (BATCH) {
    spec = bt.spec[blockIdx.y];
    K = bt.K[blockIdx.y];
  }
  using C = C2<T>;
  constexpr int U = F::U, V = F::V, L = F::L;
  constexpr int NB = HALF ? V / 2 : V;
  constexpr int ROWS = HALF ? L / 2 : L;
  constexpr int CS = ColCb<T, F, CB>::CS;
  using Last = typename F::Last;
  constexpr int Q = Last::Q, CH = Q < 4 ? Q : 4;
  static_assert(ColCb<T, F, CB>::ok, "plan does not fit a blocked column workgroup");
  extern __shared__ char lds_raw[];
  C* const lds_all = reinterpret_cast<C*>(lds_raw);
  const int tid = threadIdx.x;
  const int c = tid / U, lane = tid - c * U;      // transform view
  const int cio = tid % CB, rio = tid / CB;       // staging view
  // first half of the grid: x/y pairs; second half: z (as the one-column
  // kernel; each half a multiple of 8 workgroups)
  const int half_grid = gridDim.x / 2;
  const bool zjob = (int)blockIdx.x >= half_grid;
  const int bid = zjob ? (int)blockIdx.x - half_grid : (int)blockIdx.x;
  const int kx0 = reg_xcd_remap(bid, half_grid) * CB;
  if (kx0 >= Hx) return;                          // whole workgroup
  int kx = kx0 + c;
  if (kx >= Hx) kx = Hx - 1;                      // computed, never stored
  const bool live_io = kx0 + cio < Hx;
  const int kx_io = live_io ? kx0 + cio : Hx - 1;
  C* const lds = lds_all + c * CS;
  C r[3][V];
  // global loads of one component: runs of CB consecutive columns
  auto fetch = [&](auto Cc, C* v) {
    const C* src = spec + (long)Cc * ROWS * Hx + kx_io;
    sfor<0, NB>([&](auto I) { v[I] = ldc(src + (long)(rio + U * I) * Hx); });
  };
  // staged values -> this lane's registers (natural order n = lane + U b)
  auto stage = [&](auto Cc, const C* v) {
    sfor<0, NB>([&](auto I) { lds_all[cio * CS + rio + U * I] = v[I]; });
    __syncthreads();
    sfor<0, NB>([&](auto B) { r[Cc][B] = lds[lane + U * B]; });
    __syncthreads();
  };
  auto store = [&](auto Cc) {
    __syncthreads();                              // the inverse's last LDS reads
    sfor<0, NB>([&](auto B) { lds[lane + U * B] = r[Cc][B]; });
    __syncthreads();
    if (live_io) {
      C* dst = spec + (long)Cc * ROWS * Hx + kx_io;
      sfor<0, NB>([&](auto I) { stc(dst + (long)(rio + U * I) * Hx, lds_all[cio * CS + rio + U * I]); });
    }
  };
  const long P = (long)Hx * L;
  const T* kxx = K + (long)kx * L;
  const T* kyy = kxx + P;
  const T* kzz = kyy + P;
  const T* kxyr = kzz + P;
  const T* kxyi = kxyr + P;
  // index of ky in the stored half and the sign of the odd (xy) plane there
  auto mirror = [](unsigned k) { return (HALF && k > (unsigned)(L / 2)) ? (unsigned)L - k : k; };
  auto odd_sign = [](unsigned k) { return (HALF && k > (unsigned)(L / 2)) ? T(-1) : T(1); };
  auto each_k = [&](auto&& fn) {
    sfor<0, Last::NG>([&](auto J) {
      sfor<0, Q / CH>([&](auto Ch) {
        const int ln = rfft::opaque(lane);
        sfor<0, CH>([&](auto Tq) {
          constexpr int t = Ch * CH + Tq;
          fn(J * Q + t, (unsigned)Last::k_of(ln, J, t));
        });
      });
    });
  };
  constexpr auto c0 = std::integral_constant<int, 0>{};
  constexpr auto c1 = std::integral_constant<int, 1>{};
  constexpr auto c2 = std::integral_constant<int, 2>{};
  if (zjob) {
    C v[NB];
    fetch(c2, v);
    stage(c2, v);
    F::template forward<HALF>(r[2], lds, lane, tw, Bar{});
    each_k([&](int i, unsigned k) {
      const T cz = kzz[mirror(k)];
      r[2][i] = mk<T>(cz * r[2][i].x, cz * r[2][i].y);
    });
    F::template inverse<HALF>(r[2], lds, lane, tw, Bar{});
    store(c2);
    return;
  }
  {
    C v0[NB], v1[NB];
    fetch(c0, v0);
    fetch(c1, v1);
    stage(c0, v0);
    stage(c1, v1);
  }
  sfor<0, 2>([&](auto Cc) { F::template forward<HALF>(r[Cc], lds, lane, tw, Bar{}); });
  each_k([&](int i, unsigned k) {
    const unsigned km = mirror(k);
    const T a = kxx[km], b = kyy[km], xr = odd_sign(k) * kxyr[km];
    const C mx = r[0][i], my = r[1][i];
    C cxy, cyx;
    if constexpr (XYI) {
      const C xy = mk<T>(xr, kxyi[km]);
      cxy = cmul(xy, my);
      cyx = cmul(xy, mx);
    } else {
      cxy = mk<T>(xr * my.x, xr * my.y);
      cyx = mk<T>(xr * mx.x, xr * mx.y);
    }
    r[0][i] = mk<T>(fmad(a, mx.x, cxy.x), fmad(a, mx.y, cxy.y));
    r[1][i] = mk<T>(fmad(b, my.x, cyx.x), fmad(b, my.y, cyx.y));
  });
  sfor<0, 2>([&](auto Cc) {
    F::template inverse<HALF>(r[Cc], lds, lane, tw, Bar{});
    store(Cc);
  });
}

// ---------------------------------------- column, one transform per wave
// The kernels above run a transform on 128 or more lanes, so every LDS
// exchange is a workgroup barrier and all waves of a workgroup step through
// load, FFT stages, tensor product and store together; with two workgroups
// per CU the CU then alternates between vector-memory, LDS and VALU phases
// instead of overlapping them.  Here a transform runs on ONE wave (64 lanes
// of V values: the 1024-point plan with 16 values per lane), so its
// exchanges only wait for the wave's own LDS accesses and the waves of a CU
// drift apart.  The exchanges move real and imaginary parts separately
// (RegFft::forward_split), so a wave's LDS buffer is kSlots values of T and
// the 12 waves a CU holds (register-bound, below) need 104 KB of LDS.
//
// A workgroup of W = 2 CB waves holds either CB adjacent columns' x and y
// components (wave 2c: x of column c, wave 2c + 1: its y) or W adjacent
// columns' z.  The spectrum moves in runs of CB (x/y) or W (z) consecutive
// values per row, staged through the waves' buffers (one barrier each way).
// The xy block of the tensor needs both components at each ky: the x and y
// waves of a column swap their forward spectra through LDS, half of the
// registers at a time (workgroup barriers), and each forms its own output
// with the one-column kernel's arithmetic -- same plan, same values, same
// operations: variant 1's bits for 1024-point columns, variant 0's for 512.
struct WaveBar {
  // a wave's own LDS writes complete before its later LDS reads (and the
  // compiler keeps the accesses on their side)
  __device__ __forceinline__ void operator()() const { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }
};

template <class T, class F, bool HALF, int CB>
struct ColWave {
  static constexpr int W = 2 * CB;                 // waves per workgroup
  static constexpr int kThreads = 64 * W;
  static constexpr int ROWS = HALF ? F::L / 2 : F::L;
  // buffer stride in T: the exchange slots or the staged column (ROWS
  // complex values), whichever is larger, padded to 32 B modulo 256 B so
  // that adjacent buffers' staging accesses fall on different banks
  static constexpr int PB = 256 / (int)sizeof(T), PAD = 32 / (int)sizeof(T);
  static constexpr int base = F::kSlots > 2 * ROWS ? F::kSlots : 2 * ROWS;
  static constexpr int kBuf = base + ((PAD - base % PB) % PB + PB) % PB;
  static constexpr size_t kLds = sizeof(T) * (size_t)W * kBuf;
  static constexpr bool ok = F::U == 64 && F::kPerGroup == 1 && kThreads <= 1024 &&
                             kLds <= 160 * 1024 && ROWS % (kThreads / CB) == 0;
};

// FP64: three waves per SIMD (168 VGPRs): the tensor values are fetched
// before the work that precedes their use (the z component's before its
// forward transform, the x/y pair's first half before the exchange of
// spectra, the second half before that half's exchange); at four waves per
// SIMD that spilled, and 12 waves per CU with the loads in flight beat 16
// waiting on them.  FP32 values take half the registers: four waves.
template <class T, class F, bool HALF, bool XYI, int CB, bool BATCH = false>
__global__ __launch_bounds__(128 * CB) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4)))
void k_col_wave(C2<T>* __restrict__ spec, int Hx, int nxy, const C2<T>* __restrict__ tw,
                const T* __restrict__ K, BatchArg<T, BATCH> bt) {
  if constexpr (BATCH) {
    spec = bt.spec[blockIdx.y];
    K = bt.K[blockIdx.y];
  }
  using C = C2<T>;
  using G = ColWave<T, F, HALF, CB>;
  constexpr int V = F::V, L = F::L, W = G::W, NT = G::kThreads, kBuf = G::kBuf;
  constexpr int NB = HALF ? V / 2 : V;
  constexpr int ROWS = G::ROWS;
  using Last = typename F::Last;
  constexpr int Q = Last::Q;
  static_assert(G::ok, "plan does not fit a wave-per-transform column workgroup");
  extern __shared__ char lds_raw[];
  T* const lds_all = reinterpret_cast<T*>(lds_raw);
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  T* const mybuf = lds_all + wave * kBuf;
  // grid: [0, nxy) x/y workgroups, [nxy, gridDim) z workgroups (nxy a
  // multiple of 8, so both parts keep the XCD mapping)
  const bool zjob = (int)blockIdx.x >= nxy;
  const int bid = zjob ? (int)blockIdx.x - nxy : (int)blockIdx.x;
  const int blk = reg_xcd_remap(bid, zjob ? (int)gridDim.x - nxy : nxy);
  const long P = (long)Hx * L;
  auto mirror = [](unsigned k) { return (HALF && k > (unsigned)(L / 2)) ? (unsigned)L - k : k; };
  auto odd_sign = [](unsigned k) { return (HALF && k > (unsigned)(L / 2)) ? T(-1) : T(1); };
  auto job = [&](auto Zc) {
    constexpr bool Z = decltype(Zc)::value;
    constexpr int NCOL = Z ? W : CB, NCOMP = Z ? 1 : 2;
    constexpr int RPR = NT / NCOL, RND = ROWS / RPR;   // staging rows per round, rounds
    const int kx0 = blk * NCOL;
    if (kx0 >= Hx) return;                           // whole workgroup
    const int wcol = Z ? wave : wave >> 1;
    int kx = kx0 + wcol;
    if (kx >= Hx) kx = Hx - 1;                       // computed, never stored
    const int cio = tid % NCOL, rio = tid / NCOL;    // staging view
    const bool live_io = kx0 + cio < Hx;
    const int kx_io = live_io ? kx0 + cio : Hx - 1;
    auto buf_of = [&](int col, int comp) { return lds_all + (Z ? col : 2 * col + comp) * kBuf; };
    // ---- load the workgroup's columns in runs of NCOL values per row
    {
      C v[NCOMP][RND];
      sfor<0, NCOMP>([&](auto Cc) {
        const C* src = spec + (long)(Z ? 2 : (int)Cc) * ROWS * Hx + kx_io;
        sfor<0, RND>([&](auto I) { v[Cc][I] = ldc(src + (long)(rio + RPR * I) * Hx); });
      });
      sfor<0, NCOMP>([&](auto Cc) {
        C* b = reinterpret_cast<C*>(buf_of(cio, Cc));
        sfor<0, RND>([&](auto I) { b[rio + RPR * I] = v[Cc][I]; });
      });
    }
    __syncthreads();
    C r[V];
    sfor<0, V>([&](auto B) {
      if constexpr (B < NB) r[B] = reinterpret_cast<const C*>(mybuf)[lane + 64 * B];
      else r[B] = mk<T>(T(0), T(0));                 // not read by the half-input stage
    });
    T kz[Z ? V : 1];
    if constexpr (Z) {
      const T* kzz = K + 2 * P + (long)kx * L;
      sfor<0, V>([&](auto I) {
        constexpr int i = I;
        const unsigned k = (unsigned)Last::k_of(rfft::opaque(lane), i / Q, i % Q);
        kz[i] = kzz[mirror(k)];
      });
    }
    F::template forward_split<HALF>(r, mybuf, lane, tw, WaveBar{});
    // ---- tensor product
    if constexpr (Z) {
      sfor<0, V>([&](auto I) {
        constexpr int i = I;
        r[i] = mk<T>(kz[i] * r[i].x, kz[i] * r[i].y);
      });
    } else {
      const int comp = wave & 1;
      const T* kd = K + (long)comp * P + (long)kx * L;        // K_xx or K_yy
      const T* kxyr = K + 3 * P + (long)kx * L;
      const T* kxyi = kxyr + P;
      const C* pbuf = reinterpret_cast<const C*>(lds_all + (wave ^ 1) * kBuf);
      // two rounds over halves of the registers: publish own values,
      // read the partner's at the same positions (same lane -> k map)
      T pd[V / 2], px[V / 2];
      auto fetch_half = [&](auto H) {
        constexpr int HV = V / 2;
        sfor<0, HV>([&](auto I) {
          constexpr int i = H * HV + I;
          const unsigned k = (unsigned)Last::k_of(rfft::opaque(lane), i / Q, i % Q);
          const unsigned km = mirror(k);
          pd[I] = kd[km];
          px[I] = odd_sign(k) * kxyr[km];
        });
      };
      fetch_half(std::integral_constant<int, 0>{});
      sfor<0, 2>([&](auto H) {
        constexpr int HV = V / 2;
        if constexpr (H == 1) fetch_half(std::integral_constant<int, 1>{});
        __syncthreads();                             // the partner read the previous round
        sfor<0, HV>([&](auto I) { reinterpret_cast<C*>(mybuf)[lane + 64 * I] = r[H * HV + I]; });
        __syncthreads();
        sfor<0, HV>([&](auto I) {
          constexpr int i = H * HV + I;
          const unsigned k = (unsigned)Last::k_of(rfft::opaque(lane), i / Q, i % Q);
          const unsigned km = mirror(k);
          const T d = pd[I], xr = px[I];
          const C o = pbuf[lane + 64 * I];          // the other component at this k
          C c;
          if constexpr (XYI) {
            c = cmul(mk<T>(xr, kxyi[km]), o);
          } else {
            c = mk<T>(xr * o.x, xr * o.y);
          }
          r[i] = mk<T>(fmad(d, r[i].x, c.x), fmad(d, r[i].y, c.y));
        });
      });
      __syncthreads();                               // the partner's last reads of this buffer
    }
    F::template inverse_split<HALF>(r, mybuf, lane, tw, WaveBar{});
    // ---- store in runs of NCOL values per row
    WaveBar{}();
    sfor<0, NB>([&](auto B) { reinterpret_cast<C*>(mybuf)[lane + 64 * B] = r[B]; });
    __syncthreads();
    if (live_io) {
      sfor<0, NCOMP>([&](auto Cc) {
        const C* b = reinterpret_cast<const C*>(buf_of(cio, Cc));
        C* dst = spec + (long)(Z ? 2 : (int)Cc) * ROWS * Hx + kx_io;
        sfor<0, RND>([&](auto I) { stc(dst + (long)(rio + RPR * I) * Hx, b[rio + RPR * I]); });
      });
    }
  };
  if (zjob) job(std::true_type{});
  else job(std::false_type{});
}

// ---------------------------------------------------------------- plans
// Row transforms (complex length L = Px/2, one component per transform).
template <class T> using RowP128 = rfft::RegFft<T, 128, 16, 4, 4>;
template <class T> using RowP256 = rfft::RegFft<T, 256, 32, 8, 4>;
template <class T> using RowP384 = rfft::RegFft<T, 384, 16, 4, 4>;
template <class T> using RowP512 = rfft::RegFft<T, 512, 64, 8, 8>;
template <class T> using RowP1024 = rfft::RegFft<T, 1024, 64, 8, 8>;
template <class T> using RowP1536 = rfft::RegFft<T, 1536, 64, 8, 8>;
template <class T> using RowP2048 = rfft::RegFft<T, 2048, 64, 8, 8>;
template <class T> using RowP4096 = rfft::RegFft<T, 4096, 256, 16, 16>;
// Column transforms (length Py, three components per lane).  Variant 0
// holds 8 values per component and lane (about 150 VGPRs in FP64, three
// waves per SIMD); variant 1 holds 16 (one more radix per stage, one fewer
// LDS exchange, but over 256 registers in FP64).
template <class T> using ColP128 = rfft::RegFft<T, 128, 16, 4, 4>;
template <class T> using ColP256 = rfft::RegFft<T, 256, 32, 8, 4>;
template <class T> using ColP512 = rfft::RegFft<T, 512, 64, 8, 8>;
template <class T> using ColP1024 = rfft::RegFft<T, 1024, 128, 8, 4, 4>;
template <class T> using ColP1024w = rfft::RegFft<T, 1024, 64, 8, 8>;
template <class T> using ColP2048 = rfft::RegFft<T, 2048, 256, 8, 8, 4>;
template <class T> using ColP2048w = rfft::RegFft<T, 2048, 128, 16, 8>;
template <class T> using ColP4096 = rfft::RegFft<T, 4096, 512, 8, 8, 8>;
template <class T> using ColP4096w = rfft::RegFft<T, 4096, 256, 16, 16>;
template <class T> using ColP8192 = rfft::RegFft<T, 8192, 1024, 8, 8, 4, 4>;
template <class T> using ColP8192w = rfft::RegFft<T, 8192, 512, 8, 8, 8>;

template <class T, class Fn>
bool with_row(int L, int variant, Fn&& fn) {
  if (variant != 0) return false;
  switch (L) {
    case 128: fn(RowP128<T>{}); return true;
    case 256: fn(RowP256<T>{}); return true;
    case 384: fn(RowP384<T>{}); return true;
    case 512: fn(RowP512<T>{}); return true;
    case 1024: fn(RowP1024<T>{}); return true;
    case 1536: fn(RowP1536<T>{}); return true;
    case 2048: fn(RowP2048<T>{}); return true;
    case 4096: fn(RowP4096<T>{}); return true;
    default: return false;
  }
}

template <class T, class Fn>
bool with_col(int L, int variant, Fn&& fn) {
  if (variant == 0) {
    switch (L) {
      case 128: fn(ColP128<T>{}); return true;
      case 256: fn(ColP256<T>{}); return true;
      case 512: fn(ColP512<T>{}); return true;
      case 1024: fn(ColP1024<T>{}); return true;
      case 2048: fn(ColP2048<T>{}); return true;
      case 4096: fn(ColP4096<T>{}); return true;
      case 8192: fn(ColP8192<T>{}); return true;
      default: return false;
    }
  }
  if (variant == 1) {
    switch (L) {
      case 1024: fn(ColP1024w<T>{}); return true;
      case 2048: fn(ColP2048w<T>{}); return true;
      case 4096: fn(ColP4096w<T>{}); return true;
      case 8192: fn(ColP8192w<T>{}); return true;
      default: return false;
    }
  }
  return false;
}

template <class K>
inline int prep_lds(K kern, size_t bytes) {
  if (bytes > 65536 &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    return -1;
  return 0;
}

template <class T>
bool has_row_plan(int L, int variant) {
  return with_row<T>(L, variant, [](auto) {});
}
// column variants 2, 3, 4: the blocked kernel with CB = 2, 4, 8 columns per
// workgroup over the variant-0 plan, for the column lengths of the
// benchmarked grids, where the plan fits a workgroup (threads, LDS)
constexpr int col_cb_of(int variant) { return variant == 2 ? 2 : variant == 3 ? 4 : variant == 4 ? 8 : 0; }

template <class T, class Fn>
bool with_col_cb(int L, int variant, Fn&& fn) {
  const int cb = col_cb_of(variant);
  bool done = false;
  auto with_plan = [&](auto f) {
    using F = decltype(f);
    auto go = [&](auto CBc) {
      constexpr int CBv = decltype(CBc)::value;
      if constexpr (ColCb<T, F, CBv>::ok) {
        fn(f, CBc);
        done = true;
      }
    };
    if (cb == 2) go(std::integral_constant<int, 2>{});
    else if (cb == 4) go(std::integral_constant<int, 4>{});
    else if (cb == 8) go(std::integral_constant<int, 8>{});
  };
  switch (L) {
    case 512: with_plan(ColP512<T>{}); break;
    case 1024: with_plan(ColP1024<T>{}); break;
    case 2048: with_plan(ColP2048<T>{}); break;
    default: break;
  }
  return done;
}

// column variants 5, 6: the wave-per-transform kernel with CB = 2, 4
// (x/y workgroups of CB columns, z workgroups of 2 CB columns) on the
// 1024-point plan with 16 values per lane, where both the zero-padded and
// the periodic form fit a workgroup
constexpr int col_wave_of(int variant) { return variant == 5 ? 2 : variant == 6 ? 4 : 0; }

template <class T, class Fn>
bool with_col_wave(int L, int variant, Fn&& fn) {
  const int cb = col_wave_of(variant);
  bool done = false;
  auto with_plan = [&](auto f) {
    using F = decltype(f);
    auto go = [&](auto CBc) {
      constexpr int CBv = decltype(CBc)::value;
      if constexpr (ColWave<T, F, true, CBv>::ok && ColWave<T, F, false, CBv>::ok) {
        fn(f, CBc);
        done = true;
      }
    };
    if (cb == 2) go(std::integral_constant<int, 2>{});
    else if (cb == 4) go(std::integral_constant<int, 4>{});
  };
  if (L == 1024) with_plan(ColP1024w<T>{});
  return done;
}

template <class T>
bool has_col_plan(int L, int variant) {
  if (col_cb_of(variant)) return with_col_cb<T>(L, variant, [](auto, auto) {});
  if (col_wave_of(variant)) return with_col_wave<T>(L, variant, [](auto, auto) {});
  return with_col<T>(L, variant, [](auto) {});
}

// The launchers: BATCH launches n instances (gridDim.y = n) with buffers
// from bt; otherwise one instance (n = 1) with the buffers passed directly.
template <class T, bool BATCH>
int launch_row_fwd_t(int L, int variant, int half, int n, const BatchArg<T, BATCH>& bt,
                     const double* in, const double* w, long comp_stride, int nx_in, int rows,
                     const C2<T>* tw, const C2<T>* post, C2<T>* spec, hipStream_t s) {
  int rc = 0;
  with_row<T>(L, variant, [&](auto f) {
    using F = decltype(f);
    const size_t lds = sizeof(C2<T>) * F::kSlots * F::kPerGroup;
    const int grid = (3 * rows + F::kPerGroup - 1) / F::kPerGroup;
    auto kern = half ? k_row_fwd_reg<T, F, true, BATCH> : k_row_fwd_reg<T, F, false, BATCH>;
    if (prep_lds(kern, lds)) { rc = -1; return; }
    hipLaunchKernelGGL(kern, dim3(grid, n), dim3(F::kThreads), lds, s, in, w, comp_stride, nx_in,
                       rows, tw, post, spec, bt);
    rc = hipPeekAtLastError() == hipSuccess ? 1 : -1;
  });
  return rc;
}

template <class T, bool BATCH>
int launch_col_t(int L, int variant, int half, int n, const BatchArg<T, BATCH>& bt, C2<T>* spec,
                 int Hx, const C2<T>* tw, const T* K, int has_xy_im, hipStream_t s,
                 hipEvent_t e0, hipEvent_t e1) {
  int rc = 0;
  if (col_cb_of(variant)) {
    // zero-padded grids (real xy spectrum) or periodic grids (complex xy)
    if ((half != 0) == (has_xy_im != 0)) return 0;
    with_col_cb<T>(L, variant, [&](auto f, auto CBc) {
      using F = decltype(f);
      constexpr int CBv = decltype(CBc)::value;
      using G = ColCb<T, F, CBv>;
      const int nblk = (Hx + CBv - 1) / CBv;
      const int grid = 2 * ((nblk + 7) / 8 * 8);
      auto kern = half ? k_col_cb<T, F, true, false, CBv, BATCH>
                       : k_col_cb<T, F, false, true, CBv, BATCH>;
      if (prep_lds(kern, G::kLds)) { rc = -1; return; }
      hipExtLaunchKernelGGL(kern, dim3(grid, n), dim3(G::kThreads), (std::uint32_t)G::kLds, s,
                            e0, e1, 0u, spec, Hx, tw, K, bt);
      rc = hipPeekAtLastError() == hipSuccess ? 1 : -1;
    });
    return rc;
  }
  if (col_wave_of(variant)) {
    if ((half != 0) == (has_xy_im != 0)) return 0;
    with_col_wave<T>(L, variant, [&](auto f, auto CBc) {
      using F = decltype(f);
      constexpr int CBv = decltype(CBc)::value;
      // x/y workgroups, then z workgroups, each part a multiple of 8
      const int nxy = ((Hx + CBv - 1) / CBv + 7) / 8 * 8;
      const int nz = ((Hx + 2 * CBv - 1) / (2 * CBv) + 7) / 8 * 8;
      auto kern = half ? k_col_wave<T, F, true, false, CBv, BATCH>
                       : k_col_wave<T, F, false, true, CBv, BATCH>;
      const size_t lds = half ? ColWave<T, F, true, CBv>::kLds : ColWave<T, F, false, CBv>::kLds;
      if (prep_lds(kern, lds)) { rc = -1; return; }
      hipExtLaunchKernelGGL(kern, dim3(nxy + nz, n), dim3(ColWave<T, F, true, CBv>::kThreads),
                            (std::uint32_t)lds, s, e0, e1, 0u, spec, Hx, nxy, tw, K, bt);
      rc = hipPeekAtLastError() == hipSuccess ? 1 : -1;
    });
    return rc;
  }
  with_col<T>(L, variant, [&](auto f) {
    using F = decltype(f);
    const size_t lds = sizeof(C2<T>) * F::kSlots * F::kPerGroup;
    // x/y pairs, then z components: each half rounded up to a multiple of 8
    const int grid = 2 * (((Hx + F::kPerGroup - 1) / F::kPerGroup + 7) / 8 * 8);
    auto kern = half ? (has_xy_im ? k_col_reg<T, F, true, true, BATCH>
                                  : k_col_reg<T, F, true, false, BATCH>)
                     : (has_xy_im ? k_col_reg<T, F, false, true, BATCH>
                                  : k_col_reg<T, F, false, false, BATCH>);
    if (prep_lds(kern, lds)) { rc = -1; return; }
    hipExtLaunchKernelGGL(kern, dim3(grid, n), dim3(F::kThreads), (std::uint32_t)lds, s, e0, e1,
                          0u, spec, Hx, tw, K, bt);
    rc = hipPeekAtLastError() == hipSuccess ? 1 : -1;
  });
  return rc;
}

template <class T, bool BATCH>
int launch_row_inv_t(int L, int variant, int half, int n, const BatchArg<T, BATCH>& bt,
                     const C2<T>* spec, int rows, const C2<T>* tw, const C2<T>* post,
                     double* out, long comp_stride, int nx_out, hipStream_t s) {
  int rc = 0;
  with_row<T>(L, variant, [&](auto f) {
    using F = decltype(f);
    const size_t lds = sizeof(C2<T>) * F::kSlots * F::kPerGroup;
    const int grid = (3 * rows + F::kPerGroup - 1) / F::kPerGroup;
    auto kern = half ? k_row_inv_reg<T, F, true, BATCH> : k_row_inv_reg<T, F, false, BATCH>;
    if (prep_lds(kern, lds)) { rc = -1; return; }
    hipLaunchKernelGGL(kern, dim3(grid, n), dim3(F::kThreads), lds, s, spec, rows, tw, post, out,
                       comp_stride, nx_out, bt);
    rc = hipPeekAtLastError() == hipSuccess ? 1 : -1;
  });
  return rc;
}

template <class T>
int launch_row_fwd(int L, int variant, int half, const double* in, const double* w, long comp_stride, int nx_in,
                   int rows, const C2<T>* tw, const C2<T>* post, C2<T>* spec, hipStream_t s) {
  return launch_row_fwd_t<T, false>(L, variant, half, 1, NoBatch{}, in, w, comp_stride, nx_in,
                                    rows, tw, post, spec, s);
}

template <class T>
int launch_col(int L, int variant, int half, C2<T>* spec, int Hx, const C2<T>* tw, const T* K, int has_xy_im,
               hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  return launch_col_t<T, false>(L, variant, half, 1, NoBatch{}, spec, Hx, tw, K, has_xy_im, s,
                                e0, e1);
}

template <class T>
int launch_row_inv(int L, int variant, int half, const C2<T>* spec, int rows, const C2<T>* tw,
                   const C2<T>* post, double* out, long comp_stride, int nx_out, hipStream_t s) {
  return launch_row_inv_t<T, false>(L, variant, half, 1, NoBatch{}, spec, rows, tw, post, out,
                                    comp_stride, nx_out, s);
}

template <class T>
int launch_row_fwd_batch(int L, int variant, int half, int n, const RegBatch<T>& b,
                         long comp_stride, int nx_in, int rows, const C2<T>* tw,
                         const C2<T>* post, hipStream_t s) {
  if (n < 1 || n > kBatchMax) return -1;
  return launch_row_fwd_t<T, true>(L, variant, half, n, b, nullptr, nullptr, comp_stride, nx_in,
                                   rows, tw, post, nullptr, s);
}

template <class T>
int launch_col_batch(int L, int variant, int half, int n, const RegBatch<T>& b, int Hx,
                     const C2<T>* tw, int has_xy_im, hipStream_t s) {
  if (n < 1 || n > kBatchMax) return -1;
  return launch_col_t<T, true>(L, variant, half, n, b, nullptr, Hx, tw, nullptr, has_xy_im, s,
                               nullptr, nullptr);
}

template <class T>
int launch_row_inv_batch(int L, int variant, int half, int n, const RegBatch<T>& b, int rows,
                         const C2<T>* tw, const C2<T>* post, long comp_stride, int nx_out,
                         hipStream_t s) {
  if (n < 1 || n > kBatchMax) return -1;
  return launch_row_inv_t<T, true>(L, variant, half, n, b, nullptr, rows, tw, post, nullptr,
                                   comp_stride, nx_out, s);
}

}  // namespace demag_reg
