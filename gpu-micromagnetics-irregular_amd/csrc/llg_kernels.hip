// llg_kernels.hip -- the LLG CVODE plug-ins on gfx950: RHS, analytic Jv,
// 3x3 block-Jacobi setup/solve, unit-sphere normalisation, and the
// context that replaces the reference's byte-mirrored UserData structs.
//
// Reference behaviour restated here (paths under src/irregular/2d_fft/):
//   RHS    2d_i4/2d_fft.cu:262-319 (ymsk), 2d_i6/2d_i6.cu:250-343 (compact),
//          2d_i1/2d_fft.cu:296-360 (linear anisotropy + DMI),
//          2d_i3/2d_fft.cu:451-627 (LDS-tiled windowed polycrystal);
//          wrapper order 2d_i4/2d_fft.cu:328-366
//   Jv     2d_i4/jtv.cu:47-147, 2d_i6/jtv.cu:47-183, 2d_i1/jtv.cu:43-119
//   Psetup 2d_i4/precond.cu:120-252,324-358; 2d_i6/precond.cu:118-249;
//          two-phase jok protocol src/2d/2d_heter/precond.cu:154-192,265-314
//   Psolve 2d_i4/precond.cu:255-286,361-385; 2d_i6/precond.cu:252-285,361-396
//   normalize 2d_i1/2d_fft.cu:253-275 (exact), 2d_i2/2d_fft.cu:262-283 (+1e-3)
//
// MI355X layout: SoA FP64 state [mx|my|mz]; geometry as ONE byte per cell
// (the reference replicates an FP64 0/1 mask in three planes: 24 B/cell
// of extra traffic); compact mode launches one thread per active cell and
// zeroes the hole cells in the same launch.  Stencil kernels use 64x4
// workgroups so one wavefront covers 64 consecutive cells of a row
// (512 B coalesced per plane); the neighbour rows come through L1/L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "llg/llg_plugin.h"
#include "llg/llg_batch.h"
#include "llg/learned_pc.h"
#include "llg_hook.h"
#include "llg/nvector_hip.h"
#include "kprof.h"

// No implicit multiply-add contraction: every product and sum is rounded
// on its own, so a value does not depend on which kernel forms it (the
// fused Krylov operator below reproduces the separate launches bit for bit).
#pragma clang fp contract(off)

namespace {

thread_local std::string g_err;
void set_err(const std::string& e) { g_err = e; }

constexpr int kBX = 64, kBY = 4;   // 2D stencil workgroup (one wave per row)
constexpr int k1D = 256;

struct TermsD {
  double che, chk, cha, chb;
  double msk[3], nsk[3], happ[3];
  int aniso, cell_axes, window, gate;
};

struct GeoD {
  int ng, ny, ncell;
  int boundary, exch_self;
  int compact;
  const unsigned char* active;  // ncell
  const double* w;              // ncell or null
  const double* cmsk;           // 3*ncell or null
  const double* cnsk;           // 3*ncell or null
  const int* ids;               // active ids (compact)
  const int* hole_ids;          // inactive ids (compact)
  int n_active, n_hole;
  double alpha, chg;
};

TermsD to_dev(const llg_terms& t) {
  TermsD d;
  d.che = t.che; d.chk = t.chk; d.cha = t.cha; d.chb = t.chb;
  for (int a = 0; a < 3; a++) { d.msk[a] = t.msk[a]; d.nsk[a] = t.nsk[a]; d.happ[a] = t.happ[a]; }
  d.aniso = t.aniso; d.cell_axes = t.cell_axes; d.window = t.window; d.gate = t.gate;
  return d;
}

__device__ inline void nbrs(const GeoD& g, int gx, int gy, int& lc, int& rc, int& uc, int& dc) {
  const int self = gy * g.ng + gx;
  if (g.boundary == LLG_BC_NEUMANN) {
    lc = gx > 0 ? self - 1 : self;
    rc = gx < g.ng - 1 ? self + 1 : self;
    uc = gy > 0 ? self - g.ng : self;
    dc = gy < g.ny - 1 ? self + g.ng : self;
  } else {
    const int xl = gx == 0 ? g.ng - 1 : gx - 1;
    const int xr = gx == g.ng - 1 ? 0 : gx + 1;
    const int yu = gy == 0 ? g.ny - 1 : gy - 1;
    const int yd = gy == g.ny - 1 ? 0 : gy + 1;
    lc = gy * g.ng + xl; rc = gy * g.ng + xr; uc = yu * g.ng + gx; dc = yd * g.ng + gx;
  }
}

__device__ inline void axes(const GeoD& g, const TermsD& T, int cell, double msk[3], double nsk[3]) {
  const int N = g.ncell;
  if (T.cell_axes && g.cmsk) {
    msk[0] = g.cmsk[cell]; msk[1] = g.cmsk[N + cell]; msk[2] = g.cmsk[2 * N + cell];
  } else {
    msk[0] = T.msk[0]; msk[1] = T.msk[1]; msk[2] = T.msk[2];
  }
  if (T.cell_axes && g.cnsk) {
    nsk[0] = g.cnsk[cell]; nsk[1] = g.cnsk[N + cell]; nsk[2] = g.cnsk[2 * N + cell];
  } else {
    nsk[0] = T.nsk[0]; nsk[1] = T.nsk[1]; nsk[2] = T.nsk[2];
  }
}

// Read access to the direction vector v of a Jacobian product: stored
// (VPlain), or formed on the fly as (c x) / s, the SPGMR basis vector
// normalised and unscaled exactly as N_VScale then N_VDiv round it.
// at(a, N, idx, nb): component a at cell idx, which is the centre (nb 0)
// or its left / right / up / down neighbour (nb 1..4) of the cell being
// evaluated.
struct VPlain {
  const double* __restrict__ v;
  __device__ double at(int a, int N, int idx, int) const { return v[a * N + idx]; }
};
struct VScaled {
  const double* __restrict__ x;
  const double* __restrict__ s;
  double c;
  __device__ double at(int a, int N, int idx, int) const {
    const int i = a * N + idx;
    const double t = c * x[i];
    return t / s[i];
  }
};
// the same (c x) / s values, formed once per cell of an LDS tile (with a
// one-cell halo) and read from there: tile[a][ty][tx] at the cell's tile
// position (tx, ty), the neighbours at tx -+ 1 / ty -+ 1
template <int SY, int SX>
struct VTile {
  const double (*t)[SY][SX];
  int tx, ty;
  __device__ double at(int a, int, int, int nb) const {
    const int dx = nb == 1 ? -1 : (nb == 2 ? 1 : 0);
    const int dy = nb == 3 ? -1 : (nb == 4 ? 1 : 0);
    return t[a][ty + dy][tx + dx];
  }
};

// the state y read by field / jtv_cell: stored, or staged in an LDS tile
// with VTile's one-cell halo (the lagged Krylov pass stages it beside v)
__device__ inline double yat(const double* __restrict__ y, int a, int N, int idx, int) {
  return y[a * N + idx];
}
template <int SY, int SX>
__device__ inline double yat(const VTile<SY, SX>& y, int a, int N, int idx, int nb) {
  return y.at(a, N, idx, nb);
}
// h (without demag) from terms T; optional linearised partner dh for v
template <bool LIN, class VA = VPlain, class YP = const double*>
__device__ inline void field(const GeoD& g, const TermsD& T, const YP& y,
                             const VA& v, int cell, int gx, int gy,
                             double h[3], double dh[3]) {
  const int N = g.ncell;
  int lc, rc, uc, dc;
  nbrs(g, gx, gy, lc, rc, uc, dc);
  double msk[3], nsk[3];
  axes(g, T, cell, msk, nsk);
  const double m[3] = {yat(y, 0, N, cell, 0), yat(y, 1, N, cell, 0), yat(y, 2, N, cell, 0)};
  const bool win = T.window && g.w;
  const double wl = win ? g.w[lc] : 1.0, wr = win ? g.w[rc] : 1.0;
  const double wu = win ? g.w[uc] : 1.0, wd = win ? g.w[dc] : 1.0;
  const double mdotmsk = m[0] * msk[0] + m[1] * msk[1] + m[2] * msk[2];
  double vv[3] = {0, 0, 0}, vdotmsk = 0.0;
  if (LIN) {
    vv[0] = v.at(0, N, cell, 0); vv[1] = v.at(1, N, cell, 0); vv[2] = v.at(2, N, cell, 0);
    vdotmsk = vv[0] * msk[0] + vv[1] * msk[1] + vv[2] * msk[2];
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double L = wl * yat(y, a, N, lc, 1), R = wr * yat(y, a, N, rc, 2);
    const double U = wu * yat(y, a, N, uc, 3), D = wd * yat(y, a, N, dc, 4);
    double ex = L + R + U + D;
    if (g.exch_self) ex = ex - 4.0 * m[a];
    double an;
    if (T.aniso == LLG_ANISO_CUBIC) an = msk[a] * T.chk * m[a] * (m[a] * m[a] - 1.0);
    else an = msk[a] * (T.chk * mdotmsk + T.cha);
    h[a] = T.happ[a] + T.che * ex + an + T.chb * nsk[a] * (L + R);
    if (LIN) {
      const double vL = wl * v.at(a, N, lc, 1), vR = wr * v.at(a, N, rc, 2);
      const double vU = wu * v.at(a, N, uc, 3), vD = wd * v.at(a, N, dc, 4);
      double vex = vL + vR + vU + vD;
      if (g.exch_self) vex = vex - 4.0 * vv[a];
      double dan;
      if (T.aniso == LLG_ANISO_CUBIC) dan = (msk[a] * T.chk * (3.0 * m[a] * m[a] - 1.0)) * vv[a];
      else dan = msk[a] * T.chk * vdotmsk;
      dh[a] = T.che * vex + dan + T.chb * nsk[a] * (vL + vR);
    }
  }
}

__device__ inline double gate_of(const GeoD& g, int gate, int cell) {
  if (gate == LLG_GATE_MASK) return g.active[cell] ? 1.0 : 0.0;
  if (gate == LLG_GATE_WINDOW) return g.w ? g.w[cell] : 1.0;
  return 1.0;
}

// Tile of a 2D stencil launch, XCD-aware: the hardware deals consecutive
// workgroups round-robin over the 8 XCDs (each with its own L2), so in
// launch order neighbouring tiles, which share halo rows, would sit in
// different L2s and each fetch the halos from memory.  Renumbering gives
// each XCD a contiguous band of tiles in row-major order; only the band
// edges are fetched twice.
__device__ inline void tile_of(int& tbx, int& tby) {
  const int G = gridDim.x * gridDim.y;
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  const int per = G / 8, rem = G % 8;
  const int x = bid % 8, idx = bid / 8;
  const int t = x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
  tbx = t % gridDim.x;
  tby = t / gridDim.x;
}

// thread -> cell mapping; returns false when the thread has no cell.
// compact: threads [0, n_active) take active ids, [n_active, n_active+n_hole)
// zero a hole cell of `out` (3 planes) and return false.
__device__ inline bool map_cell(const GeoD& g, double* out, int& cell, int& gx, int& gy) {
  if (g.compact) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < g.n_active) {
      cell = g.ids[t];
    } else {
      const int hidx = t - g.n_active;
      if (hidx < g.n_hole) {
        const int hc = g.hole_ids[hidx];
        out[hc] = 0.0; out[g.ncell + hc] = 0.0; out[2 * g.ncell + hc] = 0.0;
      }
      return false;
    }
    gx = cell % g.ng;
    gy = cell / g.ng;
    return true;
  }
  int tbx, tby;
  tile_of(tbx, tby);
  gx = tbx * blockDim.x + threadIdx.x;
  gy = tby * blockDim.y + threadIdx.y;
  if (gx >= g.ng || gy >= g.ny) return false;
  cell = gy * g.ng + gx;
  return true;
}

// ------------------------------------------------------------------ RHS
// The Newton residual formed in the RHS pass: out = c0 zn1 + c1 ycor + c2 f,
// combined in the order N_VLinearCombination uses (c0 x0, then fma with
// each further term), with f the value the plain RHS pass would write.
struct ResIn {
  const double* zn1;
  const double* ycor;
  double c0, c1, c2;
};
__device__ inline double res_of(const ResIn& r, int i, double f) {
  double acc = r.c0 * r.zn1[i];
  acc = fma(r.c1, r.ycor[i], acc);
  return fma(r.c2, f, acc);
}

template <bool RES>
__global__ __launch_bounds__(256) void k_rhs(GeoD g, TermsD T, const double* __restrict__ y,
                                             const double* __restrict__ hd, double* __restrict__ yd,
                                             ResIn rin) {
  int cell, gx, gy;
  if (!map_cell(g, yd, cell, gx, gy)) return;
  const int N = g.ncell;
  double h[3], dummy[3];
  field<false>(g, T, y, VPlain{nullptr}, cell, gx, gy, h, dummy);
  if (hd) { h[0] += hd[cell]; h[1] += hd[N + cell]; h[2] += hd[2 * N + cell]; }
  const double m1 = y[cell], m2 = y[N + cell], m3 = y[2 * N + cell];
  const double mh = m1 * h[0] + m2 * h[1] + m3 * h[2];
  const double f1 = g.chg * (m3 * h[1] - m2 * h[2]) + g.alpha * (h[0] - mh * m1);
  const double f2 = g.chg * (m1 * h[2] - m3 * h[0]) + g.alpha * (h[1] - mh * m2);
  const double f3 = g.chg * (m2 * h[0] - m1 * h[1]) + g.alpha * (h[2] - mh * m3);
  if (g.compact) {
    yd[cell] = f1; yd[N + cell] = f2; yd[2 * N + cell] = f3;
  } else {
    const double gt = gate_of(g, T.gate, cell);
    if constexpr (RES) {
      yd[cell] = res_of(rin, cell, gt * f1);
      yd[N + cell] = res_of(rin, N + cell, gt * f2);
      yd[2 * N + cell] = res_of(rin, 2 * N + cell, gt * f3);
    } else {
      yd[cell] = gt * f1; yd[N + cell] = gt * f2; yd[2 * N + cell] = gt * f3;
    }
  }
}

// LDS-tiled RHS on the dense grid (the windowed polycrystal form is
// 2d_i3/2d_fft.cu:451-627; without a window w = 1): a (BX+2)x(BY+2) halo of
// w*m per component is staged in LDS once and the four neighbours are read
// from it, the same values field() reads (periodic wrap, or the edge cell
// itself at a Neumann edge), so the result equals k_rhs bit for bit.
// One tile (tbx, tby) of the RHS; `tile` is the workgroup's LDS halo.
template <int BX, int BY, bool RES>
__device__ __forceinline__ void rhs_tile(double (&tile)[3][BY + 2][BX + 2], const GeoD& g,
                                         const TermsD& T, const double* __restrict__ y,
                                         const double* __restrict__ hd,
                                         double* __restrict__ yd, const ResIn& rin, int tbx,
                                         int tby) {
  constexpr int SX = BX + 2, SY = BY + 2;
  const int N = g.ncell;
  const int x0 = tbx * BX, y0 = tby * BY;
  const bool win = T.window && g.w;
  for (int s = threadIdx.y * BX + threadIdx.x; s < SX * SY; s += BX * BY) {
    const int sx = s % SX, sy = s / SX;
    int gx = x0 + sx - 1, gy = y0 + sy - 1;
    if (g.boundary == LLG_BC_NEUMANN) {
      // a missing neighbour is the cell itself (nbrs): the halo slot holds
      // the edge cell's value
      gx = gx < 0 ? 0 : gx;
      gy = gy < 0 ? 0 : gy;
    } else {
      gx = gx < 0 ? gx + g.ng : (gx >= g.ng ? gx - g.ng : gx);
      gy = gy < 0 ? gy + g.ny : (gy >= g.ny ? gy - g.ny : gy);
    }
    gx = gx > g.ng - 1 ? g.ng - 1 : gx;   // Neumann edge, or never read (see k_krylov_tiled)
    gy = gy > g.ny - 1 ? g.ny - 1 : gy;
    const int c = gy * g.ng + gx;
    const double ww = win ? g.w[c] : 1.0;
    tile[0][sy][sx] = ww * y[c];
    tile[1][sy][sx] = ww * y[N + c];
    tile[2][sy][sx] = ww * y[2 * N + c];
  }
  __syncthreads();
  const int gx = x0 + threadIdx.x, gy = y0 + threadIdx.y;
  if (gx >= g.ng || gy >= g.ny) return;
  const int cell = gy * g.ng + gx;
  const int sx = threadIdx.x + 1, sy = threadIdx.y + 1;
  double msk[3], nsk[3];
  axes(g, T, cell, msk, nsk);
  const double m[3] = {y[cell], y[N + cell], y[2 * N + cell]};
  const double mdotmsk = m[0] * msk[0] + m[1] * msk[1] + m[2] * msk[2];
  double h[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double L = tile[a][sy][sx - 1], R = tile[a][sy][sx + 1];
    const double U = tile[a][sy - 1][sx], D = tile[a][sy + 1][sx];
    double ex = L + R + U + D;
    if (g.exch_self) ex = ex - 4.0 * m[a];
    double an;
    if (T.aniso == LLG_ANISO_CUBIC) an = msk[a] * T.chk * m[a] * (m[a] * m[a] - 1.0);
    else an = msk[a] * (T.chk * mdotmsk + T.cha);
    h[a] = T.happ[a] + T.che * ex + an + T.chb * nsk[a] * (L + R);
  }
  if (hd) { h[0] += hd[cell]; h[1] += hd[N + cell]; h[2] += hd[2 * N + cell]; }
  const double m1 = m[0], m2 = m[1], m3 = m[2];
  const double mh = m1 * h[0] + m2 * h[1] + m3 * h[2];
  const double gt = gate_of(g, T.gate, cell);
  const double f1 = gt * (g.chg * (m3 * h[1] - m2 * h[2]) + g.alpha * (h[0] - mh * m1));
  const double f2 = gt * (g.chg * (m1 * h[2] - m3 * h[0]) + g.alpha * (h[1] - mh * m2));
  const double f3 = gt * (g.chg * (m2 * h[0] - m1 * h[1]) + g.alpha * (h[2] - mh * m3));
  if constexpr (RES) {
    yd[cell] = res_of(rin, cell, f1);
    yd[N + cell] = res_of(rin, N + cell, f2);
    yd[2 * N + cell] = res_of(rin, 2 * N + cell, f3);
  } else {
    yd[cell] = f1; yd[N + cell] = f2; yd[2 * N + cell] = f3;
  }
}

template <int BX, int BY, bool RES>
__global__ __launch_bounds__(BX* BY) void k_rhs_tiled(GeoD g, TermsD T,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ hd,
                                                       double* __restrict__ yd, ResIn rin) {
  __shared__ double tile[3][BY + 2][BX + 2];
  int tbx, tby;
  tile_of(tbx, tby);
  rhs_tile<BX, BY, RES>(tile, g, T, y, hd, yd, rin, tbx, tby);
}

// ------------------------------------------------------------------- Jv
// Jv = J(y) v; with AT the SPGMR operator z = v + (-gamma) Jv is written
// instead, rounded exactly as N_VLinearSum(1, v, -gamma, Jv) rounds it
template <bool AT, class VA, class YP = const double*>
__device__ inline void jtv_cell(const GeoD& g, const TermsD& T, const YP& y,
                                const VA& v, int cell, int gx, int gy, double mgamma,
                                double o[3]) {
  const int N = g.ncell;
  double h[3], dh[3];
  field<true>(g, T, y, v, cell, gx, gy, h, dh);
  const double m1 = yat(y, 0, N, cell, 0), m2 = yat(y, 1, N, cell, 0), m3 = yat(y, 2, N, cell, 0);
  const double v1 = v.at(0, N, cell, 0), v2 = v.at(1, N, cell, 0), v3 = v.at(2, N, cell, 0);
  const double mh = m1 * h[0] + m2 * h[1] + m3 * h[2];
  const double dmh = (v1 * h[0] + v2 * h[1] + v3 * h[2]) + (m1 * dh[0] + m2 * dh[1] + m3 * dh[2]);
  const double j1 = g.chg * (v3 * h[1] + m3 * dh[1] - v2 * h[2] - m2 * dh[2]) +
                    g.alpha * (dh[0] - dmh * m1 - mh * v1);
  const double j2 = g.chg * (v1 * h[2] + m1 * dh[2] - v3 * h[0] - m3 * dh[0]) +
                    g.alpha * (dh[1] - dmh * m2 - mh * v2);
  const double j3 = g.chg * (v2 * h[0] + m2 * dh[0] - v1 * h[1] - m1 * dh[1]) +
                    g.alpha * (dh[2] - dmh * m3 - mh * v3);
  double o1 = j1, o2 = j2, o3 = j3;
  if (!g.compact) {
    const double gt = gate_of(g, T.gate, cell);
    o1 = gt * j1; o2 = gt * j2; o3 = gt * j3;
  }
  if (AT) {
    const double p1 = mgamma * o1, p2 = mgamma * o2, p3 = mgamma * o3;
    o1 = v1 + p1; o2 = v2 + p2; o3 = v3 + p3;
  }
  o[0] = o1; o[1] = o2; o[2] = o3;
}

template <bool AT>
__global__ __launch_bounds__(256) void k_jtv(GeoD g, TermsD T, const double* __restrict__ y,
                                             const double* __restrict__ v, double* __restrict__ Jv,
                                             double mgamma) {
  int cell, gx, gy;
  if (AT && g.compact) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int hidx = t - g.n_active;
    if (hidx >= 0 && hidx < g.n_hole) {
      const int hc = g.hole_ids[hidx];
      for (int c = 0; c < 3; c++) {
        const double p = mgamma * 0.0;
        Jv[c * g.ncell + hc] = v[c * g.ncell + hc] + p;
      }
      return;
    }
  }
  if (!map_cell(g, Jv, cell, gx, gy)) return;
  const int N = g.ncell;
  double o[3];
  jtv_cell<AT>(g, T, y, VPlain{v}, cell, gx, gy, mgamma, o);
  Jv[cell] = o[0]; Jv[N + cell] = o[1]; Jv[2 * N + cell] = o[2];
}

// ---------------------------------------------------------- block-Jacobi
__device__ inline void invert_block(const double J[9], double gamma, double P[9]) {
  const double A00 = 1.0 - gamma * J[0], A01 = -gamma * J[1], A02 = -gamma * J[2];
  const double A10 = -gamma * J[3], A11 = 1.0 - gamma * J[4], A12 = -gamma * J[5];
  const double A20 = -gamma * J[6], A21 = -gamma * J[7], A22 = 1.0 - gamma * J[8];
  const double c00 = A11 * A22 - A12 * A21;
  const double c10 = A12 * A20 - A10 * A22;
  const double c20 = A10 * A21 - A11 * A20;
  const double det = A00 * c00 + A01 * c10 + A02 * c20;
  if (!(fabs(det) > 1e-30)) {
#pragma unroll
    for (int i = 0; i < 9; i++) P[i] = (i % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  const double inv = 1.0 / det;
  P[0] = (A11 * A22 - A12 * A21) * inv;
  P[1] = (A02 * A21 - A01 * A22) * inv;
  P[2] = (A01 * A12 - A02 * A11) * inv;
  P[3] = (A12 * A20 - A10 * A22) * inv;
  P[4] = (A00 * A22 - A02 * A20) * inv;
  P[5] = (A02 * A10 - A00 * A12) * inv;
  P[6] = (A10 * A21 - A11 * A20) * inv;
  P[7] = (A01 * A20 - A00 * A21) * inv;
  P[8] = (A00 * A11 - A01 * A10) * inv;
}

__global__ __launch_bounds__(256) void k_build_P(GeoD g, TermsD T, const double* __restrict__ y,
                                                 const double* __restrict__ hd, double n0x,
                                                 double n0y, double n0z, double gamma,
                                                 int rebuild_J, double* __restrict__ Jst,
                                                 double* __restrict__ P) {
  int cell, gx, gy;
  GeoD gg = g;
  gg.n_hole = 0;   // hole blocks of P are never read: skip them
  if (!map_cell(gg, nullptr, cell, gx, gy)) return;
  const int N = g.ncell;
  double J[9];
  if (rebuild_J) {
    double h[3], dummy[3];
    field<false>(g, T, y, VPlain{nullptr}, cell, gx, gy, h, dummy);
    if (hd) { h[0] += hd[cell]; h[1] += hd[N + cell]; h[2] += hd[2 * N + cell]; }
    double msk[3], nsk[3];
    axes(g, T, cell, msk, nsk);
    const double m1 = y[cell], m2 = y[N + cell], m3 = y[2 * N + cell];
    const double m[3] = {m1, m2, m3};
    double self_ex = 0.0;
    if (g.boundary == LLG_BC_NEUMANN || g.exch_self) {
      int ghosts = 0;
      if (g.boundary == LLG_BC_NEUMANN)
        ghosts = (gx == 0) + (gx == g.ng - 1) + (gy == 0) + (gy == g.ny - 1);
      self_ex = T.che * ((double)ghosts - (g.exch_self ? 4.0 : 0.0));
    }
    const double n0[3] = {n0x, n0y, n0z};
    double k[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      double dan;
      if (T.aniso == LLG_ANISO_CUBIC) dan = T.chk * (3.0 * m[a] * m[a] - 1.0) * msk[a];
      else dan = msk[a] * msk[a] * T.chk;
      k[a] = dan + n0[a] + self_ex;
    }
    const double e1 = h[0] + m1 * k[0], e2 = h[1] + m2 * k[1], e3 = h[2] + m3 * k[2];
    const double mh = m1 * h[0] + m2 * h[1] + m3 * h[2];
    const double al = g.alpha, cg = g.chg;
    J[0] = al * (k[0] - e1 * m1 - mh);
    J[1] = cg * (m3 * k[1] - h[2]) - al * e2 * m1;
    J[2] = cg * (h[1] - m2 * k[2]) - al * e3 * m1;
    J[3] = cg * (h[2] - m3 * k[0]) - al * e1 * m2;
    J[4] = al * (k[1] - e2 * m2 - mh);
    J[5] = cg * (m1 * k[2] - h[0]) - al * e3 * m2;
    J[6] = cg * (m2 * k[0] - h[1]) - al * e1 * m3;
    J[7] = cg * (h[0] - m1 * k[1]) - al * e2 * m3;
    J[8] = al * (k[2] - e3 * m3 - mh);
    if (Jst) {
#pragma unroll
      for (int s = 0; s < 9; s++) Jst[s * N + cell] = J[s];
    }
  } else {
#pragma unroll
    for (int s = 0; s < 9; s++) J[s] = Jst[s * N + cell];
  }
  double Pi[9];
  invert_block(J, gamma, Pi);
#pragma unroll
  for (int s = 0; s < 9; s++) P[s * N + cell] = Pi[s];
}

// z = gate * P_cell r for one cell (P stored SoA, 9 planes)
// z = gate * P_cell r for one cell, from the cell's 9 block entries
__device__ inline void apply_P_vals(const GeoD& g, int gate, const double Pv[9], int cell,
                                    const double r[3], double z[3]) {
  const double z0 = Pv[0] * r[0] + Pv[1] * r[1] + Pv[2] * r[2];
  const double z1 = Pv[3] * r[0] + Pv[4] * r[1] + Pv[5] * r[2];
  const double z2 = Pv[6] * r[0] + Pv[7] * r[1] + Pv[8] * r[2];
  if (g.compact) {
    z[0] = z0; z[1] = z1; z[2] = z2;
  } else {
    const double gt = gate_of(g, gate, cell);
    z[0] = gt * z0; z[1] = gt * z1; z[2] = gt * z2;
  }
}

// z = gate * P_cell r for one cell (P stored SoA, 9 planes)
__device__ inline void apply_P_cell(const GeoD& g, int gate, const double* __restrict__ P, int cell,
                                    const double r[3], double z[3]) {
  const int N = g.ncell;
  double Pv[9];
#pragma unroll
  for (int q = 0; q < 9; q++) Pv[q] = P[q * N + cell];
  apply_P_vals(g, gate, Pv, cell, r, z);
}

// One SPGMR iteration's operator chain with left block-Jacobi
// preconditioning, in one pass:
//   u = (c vl) / s          (formed where read; vl itself is not written)
//   z = u + (-gamma) J u,   vk = s .* (gate P z)
// Every value is rounded as the separate N_VScale + N_VDiv, atimes,
// PrecondSolve and N_VProd launches round it.  Compact hole cells get
// vk = s * 0, as PrecondSolve's zeroed hole followed by the product.
__global__ __launch_bounds__(256) void k_krylov(GeoD g, TermsD T, int pgate,
                                                const double* __restrict__ P,
                                                const double* __restrict__ y,
                                                const double* __restrict__ vl,
                                                const double* __restrict__ s, double c,
                                                double mgamma, double* __restrict__ vk,
                                                const double* cdev, const int* skip) {
  if (skip && *skip) return;   // a device-resident solve that has finished
  if (cdev) c = *cdev;         // the basis scale from device memory
  int cell, gx, gy;
  const int N = g.ncell;
  if (g.compact) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int hidx = t - g.n_active;
    if (hidx >= 0 && hidx < g.n_hole) {
      const int hc = g.hole_ids[hidx];
      for (int a = 0; a < 3; a++) vk[a * N + hc] = s[a * N + hc] * 0.0;
      return;
    }
  }
  if (!map_cell(g, vk, cell, gx, gy)) return;
  double o[3], z[3];
  jtv_cell<true>(g, T, y, VScaled{vl, s, c}, cell, gx, gy, mgamma, o);
  apply_P_cell(g, pgate, P, cell, o, z);
  vk[cell] = s[cell] * z[0];
  vk[N + cell] = s[N + cell] * z[1];
  vk[2 * N + cell] = s[2 * N + cell] * z[2];
}

// k_krylov on the dense grid with the basis vector (c vl) / s formed once
// per cell in an LDS tile with a one-cell halo (the halo cell of an edge is
// the neighbour nbrs() names: wrapped when periodic, the cell itself when
// Neumann), instead of once per neighbour read.  Same values, same rounding.
template <int BX, int BY>
__global__ __launch_bounds__(BX* BY) void k_krylov_tiled(GeoD g, TermsD T, int pgate,
                                                        const double* __restrict__ P,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ vl,
                                                        const double* __restrict__ s, double c,
                                                        double mgamma, double* __restrict__ vk,
                                                        const double* cdev, const int* skip) {
  if (skip && *skip) return;   // a device-resident solve that has finished
  if (cdev) c = *cdev;         // the basis scale from device memory
  constexpr int SX = BX + 2, SY = BY + 2;
  __shared__ double tile[3][SY][SX];
  const int N = g.ncell;
  int tbx, tby;
  tile_of(tbx, tby);
  const int x0 = tbx * BX, y0 = tby * BY;
  for (int i = threadIdx.y * BX + threadIdx.x; i < SX * SY; i += BX * BY) {
    const int sy = i / SX, sx = i - sy * SX;
    if ((sx == 0 || sx == SX - 1) && (sy == 0 || sy == SY - 1)) continue;   // corners unused
    int gx = x0 + sx - 1, gy = y0 + sy - 1;
    if (g.boundary == LLG_BC_NEUMANN) {
      gx = gx < 0 ? 0 : (gx > g.ng - 1 ? g.ng - 1 : gx);
      gy = gy < 0 ? 0 : (gy > g.ny - 1 ? g.ny - 1 : gy);
    } else {
      gx = gx < 0 ? gx + g.ng : (gx >= g.ng ? gx - g.ng : gx);
      gy = gy < 0 ? gy + g.ny : (gy >= g.ny ? gy - g.ny : gy);
      // staged cells beyond the first halo column/row of a partial tile are
      // never read; keep them in bounds on grids narrower than the tile
      gx = gx > g.ng - 1 ? g.ng - 1 : gx;
      gy = gy > g.ny - 1 ? g.ny - 1 : gy;
    }
    const int cidx = gy * g.ng + gx;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double t = c * vl[a * N + cidx];
      tile[a][sy][sx] = t / s[a * N + cidx];
    }
  }
  __syncthreads();
  const int gx = x0 + threadIdx.x, gy = y0 + threadIdx.y;
  if (gx >= g.ng || gy >= g.ny) return;
  const int cell = gy * g.ng + gx;
  double o[3], z[3];
  jtv_cell<true>(g, T, y, VTile<SY, SX>{tile, (int)threadIdx.x + 1, (int)threadIdx.y + 1}, cell,
                 gx, gy, mgamma, o);
  apply_P_cell(g, pgate, P, cell, o, z);
  vk[cell] = s[cell] * z[0];
  vk[N + cell] = s[N + cell] * z[1];
  vk[2 * N + cell] = s[2 * N + cell] * z[2];
}

// Inputs of one Arnoldi step with lagged normalisation (k_arnoldi).
constexpr int kArnMax = 8;   // basis vectors before the new one (SPGMR maxl <= 8)
struct ArnIn {
  const double* win;          // w of the previous step; b for the first step with PRE
  const double* U[kArnMax];   // unnormalised basis vectors u_0 .. u_{L-1}
  double* ul;                 // u_L, written for the tile's own cells (L > 0 or PRE)
  double a;                   // u = a win - sum_j b_j U_j
  double b[kArnMax];
  // PRE with nb > 0: b = sum_{k<nb} bc_k bx_k formed where read (the Newton
  // residual), rounded as N_VLinearCombination rounds it; win unused
  const double* bx[3];
  double bc[3];
  int nb;
};

// One left-preconditioned SPGMR iteration with the normalisation of the
// basis vector left to the caller (the operator is linear, so A(u/|u|) =
// (A u)/|u|), on the dense grid:
//   u = a win - sum_{j<L} b_j U_j     (L > 0: the classical Gram-Schmidt
//                                      projection of the previous step)
//     = s .* gate P win                (L == 0 with PRE: the first
//                                      preconditioner solve and scaling)
//     = U_0                            (L == 0 without PRE)
//   w = s .* gate P (v + (-gamma) J v),  v = u / s
// u is formed once per cell in an LDS tile with a one-cell halo (the halo
// cell of an edge is the neighbour nbrs() names), staged there with the
// state y that the operator reads at the same cells (and, for the own
// cells, the scale s that w = s .* P(...) needs again), and written for the
// tile's own cells; w is written; each workgroup leaves its partial sums
//   [|u|^2, <w,U_0>, .., <w,U_{L-1}>, <w,u>, |w|^2] (and with PRE also
//   sum (win s)^2, the square of the WRMS norm of the right-hand side)
// at partials[k * gridDim.x + blockIdx.x], summed in workgroup order by the
// caller's finish pass.  Workgroup b takes the tiles of the band of XCD
// (b mod 8) in grid-stride order, so any grid size covers the grid.
// (Keeping the own cell's U_j and P in LDS instead of re-reading them
// removed ~25 % of the fetched bytes but no time: the kernel is bound by
// the latency of its two dependent load phases, not by bandwidth.)
template <int L, bool PRE, int BX, int BY>
__global__ __launch_bounds__(BX* BY) void k_arnoldi(GeoD g, TermsD T, int pgate,
                                                   const double* __restrict__ P,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ s, double mgamma,
                                                   ArnIn in, double* __restrict__ wout,
                                                   double* __restrict__ partials, int ntx,
                                                   int nty, const double* coef,
                                                   const int* skip) {
  if (skip && *skip) return;   // queued behind a gate that stopped
  constexpr int SX = BX + 2, SY = BY + 2, NT = BX * BY, K = L + 3 + (PRE ? 1 : 0);
  // projection coefficients: by value, or from device memory (a gate)
  double ca = in.a, cb[L > 0 ? L : 1];
#pragma unroll
  for (int j = 0; j < L; j++) cb[j] = in.b[j];
  if (coef) {
    ca = coef[0];
#pragma unroll
    for (int j = 0; j < L; j++) cb[j] = coef[1 + j];
  }
  __shared__ double tile[3][SY][SX];   // u / s with a one-cell halo
  __shared__ double ytile[3][SY][SX];  // the state y, same halo
  __shared__ double own[3][BY][BX];    // u of the tile's own cells
  __shared__ double sown[3][BY][BX];   // s of the tile's own cells
  __shared__ double sm[K][NT / 64];
  const int N = g.ncell;
  const int tid = threadIdx.y * BX + threadIdx.x;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;

  const int G = gridDim.x, xcd = blockIdx.x & 7, idx = blockIdx.x >> 3;
  const int nwg = G / 8 + (xcd < G % 8 ? 1 : 0);
  const int ntile = ntx * nty, per = ntile / 8, rem = ntile % 8;
  const int t_begin = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
  const int t_end = t_begin + per + (xcd < rem ? 1 : 0);
  for (int t = t_begin + idx; t < t_end; t += nwg) {
    const int x0 = (t % ntx) * BX, y0 = (t / ntx) * BY;
    for (int i = tid; i < SX * SY; i += NT) {
      const int sy = i / SX, sx = i - sy * SX;
      if ((sx == 0 || sx == SX - 1) && (sy == 0 || sy == SY - 1)) continue;   // corners unused
      const int ux = x0 + sx - 1, uy = y0 + sy - 1;
      int gx = ux, gy = uy;
      if (g.boundary == LLG_BC_NEUMANN) {
        gx = gx < 0 ? 0 : (gx > g.ng - 1 ? g.ng - 1 : gx);
        gy = gy < 0 ? 0 : (gy > g.ny - 1 ? g.ny - 1 : gy);
      } else {
        gx = gx < 0 ? gx + g.ng : (gx >= g.ng ? gx - g.ng : gx);
        gy = gy < 0 ? gy + g.ny : (gy >= g.ny ? gy - g.ny : gy);
        gx = gx > g.ng - 1 ? g.ng - 1 : gx;   // never read (see k_krylov_tiled)
        gy = gy > g.ny - 1 ? g.ny - 1 : gy;
      }
      const int cidx = gy * g.ng + gx;
      const bool mine = sx >= 1 && sx <= BX && sy >= 1 && sy <= BY && ux < g.ng && uy < g.ny;
      double u[3];
      if (L == 0 && PRE) {
        double r[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const int e = a * N + cidx;
          if (in.nb > 0) {
            double v = in.bc[0] * in.bx[0][e];
            for (int k = 1; k < in.nb; k++) v = fma(in.bc[k], in.bx[k][e], v);
            r[a] = v;
          } else {
            r[a] = in.win[e];
          }
        }
        double z[3];
        apply_P_cell(g, pgate, P, cidx, r, z);
#pragma unroll
        for (int a = 0; a < 3; a++) {
          const double sa = s[a * N + cidx];
          u[a] = sa * z[a];
          if (mine) {
            const double p = r[a] * sa;
            acc[K - 1] = fma(p, p, acc[K - 1]);
          }
        }
      } else if (L == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) u[a] = in.U[0][a * N + cidx];
      } else {
#pragma unroll
        for (int a = 0; a < 3; a++) {
          double v = ca * in.win[a * N + cidx];
#pragma unroll
          for (int j = 0; j < L; j++) v = fma(-cb[j], in.U[j][a * N + cidx], v);
          u[a] = v;
        }
      }
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const double sv = s[a * N + cidx];
        tile[a][sy][sx] = u[a] / sv;
        ytile[a][sy][sx] = y[a * N + cidx];
        if (mine) {
          own[a][sy - 1][sx - 1] = u[a];
          sown[a][sy - 1][sx - 1] = sv;
          if (L > 0 || PRE) in.ul[a * N + cidx] = u[a];
        }
      }
    }
    __syncthreads();
    const int gx = x0 + threadIdx.x, gy = y0 + threadIdx.y;
    if (gx < g.ng && gy < g.ny) {
      const int cell = gy * g.ng + gx;
      double o[3], z[3];
      jtv_cell<true>(g, T, VTile<SY, SX>{ytile, (int)threadIdx.x + 1, (int)threadIdx.y + 1},
                     VTile<SY, SX>{tile, (int)threadIdx.x + 1, (int)threadIdx.y + 1},
                     cell, gx, gy, mgamma, o);
      apply_P_cell(g, pgate, P, cell, o, z);
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const double w = sown[a][threadIdx.y][threadIdx.x] * z[a];
        wout[a * N + cell] = w;
        const double u = own[a][threadIdx.y][threadIdx.x];
        acc[0] = fma(u, u, acc[0]);
#pragma unroll
        for (int j = 0; j < L; j++) acc[1 + j] = fma(w, in.U[j][a * N + cell], acc[1 + j]);
        acc[1 + L] = fma(w, u, acc[1 + L]);
        acc[2 + L] = fma(w, w, acc[2 + L]);
      }
    }
    __syncthreads();   // the tiles are restaged by the next iteration
  }
  const int lane = tid & 63, wave = tid >> 6;
  // the K cross-lane chains interleaved so their shuffle latencies overlap
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] += __shfl_xor(acc[k], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) sm[k][wave] = acc[k];
  }
  __syncthreads();
  if (tid < K) {
    double a = sm[tid][0];
    for (int w = 1; w < NT / 64; w++) a += sm[tid][w];
    partials[tid * gridDim.x + blockIdx.x] = a;
  }
}

__global__ __launch_bounds__(256) void k_apply_P(GeoD g, int gate, const double* __restrict__ P,
                                                 const double* __restrict__ r,
                                                 double* __restrict__ z) {
  const int N = g.ncell;
  int cell;
  if (g.compact) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n_active) {
      const int hidx = t - g.n_active;
      if (hidx < g.n_hole) {
        const int hc = g.hole_ids[hidx];
        z[hc] = 0.0; z[N + hc] = 0.0; z[2 * N + hc] = 0.0;
      }
      return;
    }
    cell = g.ids[t];
  } else {
    cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= N) return;
  }
  const double rr[3] = {r[cell], r[N + cell], r[2 * N + cell]};
  double zz[3];
  apply_P_cell(g, gate, P, cell, rr, zz);
  z[cell] = zz[0]; z[N + cell] = zz[1]; z[2 * N + cell] = zz[2];
}

// --------------------------------------------------------------- normalise
__device__ __forceinline__ void normalize_cell(double* y, int ncell, int mode, int cell) {
  const double m1 = y[cell], m2 = y[ncell + cell], m3 = y[2 * ncell + cell];
  const double ymp = sqrt(m1 * m1 + m2 * m2 + m3 * m3);
  if (mode == LLG_NORM_EXACT) {
    if (ymp > 1.0e-30) {
      const double inv = 1.0 / ymp;
      y[cell] = m1 * inv; y[ncell + cell] = m2 * inv; y[2 * ncell + cell] = m3 * inv;
    }
  } else {
    const double inv = 1.0 / (ymp + 1.0e-3);
    y[cell] = m1 * inv; y[ncell + cell] = m2 * inv; y[2 * ncell + cell] = m3 * inv;
  }
}

__global__ void k_normalize(double* y, int ncell, int mode) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= ncell) return;
  normalize_cell(y, ncell, mode, cell);
}

// ------------------------------------------------------------- batches
// One launch per pass over many instances (ensembles): a workgroup finds
// its instance and its tile there, then does exactly the work of that
// instance's own launch on that tile -- the same per-cell arithmetic --
// so each instance gets the bits of its single-instance call.  What stays
// fixed per instance (geometry, term sets, demag field buffer) sits in a
// device table; the per-call vectors come by value in the arguments.
constexpr int kBatchMax = 32;

struct BatchStatic {
  GeoD g;
  TermsD rhs, jv;
  const double* hd;   // demag field of the instance's context, or null
  int ntx, nty;       // tiles of kBatchBX x kBatchBY cells
  int norm;           // normalisation mode of f()
  int pad;
};
constexpr int kBatchBX = 32, kBatchBY = 8;

struct StencilBatch {
  int n;                     // instances in this launch
  int t0[kBatchMax + 1];     // tile prefix: instance k owns tiles [t0[k], t0[k+1])
  int idx[kBatchMax];        // row of instance k in the BatchStatic table
  const double* y[kBatchMax];
  const double* v[kBatchMax];   // Jv: the direction
  double* out[kBatchMax];
};
struct ResBatch {
  ResIn r[kBatchMax];
};

// XCD-aware tile over the whole batch (as tile_of), then the instance that
// owns it (binary search over the prefix) and the tile inside it
__device__ inline void batch_tile(const StencilBatch& b, int& k, int& lt) {
  const int G = gridDim.x;
  const int per = G / 8, rem = G % 8;
  const int x = blockIdx.x % 8, i = blockIdx.x / 8;
  const int t = x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
  int lo = 0, hi = b.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t >= b.t0[mid]) lo = mid;
    else hi = mid - 1;
  }
  k = lo;
  lt = t - b.t0[lo];
}

template <bool RES>
__global__ __launch_bounds__(kBatchBX* kBatchBY) void k_rhs_batch(
    const BatchStatic* __restrict__ st, StencilBatch b, ResBatch rb) {
  __shared__ double tile[3][kBatchBY + 2][kBatchBX + 2];
  int k, lt;
  batch_tile(b, k, lt);
  const BatchStatic& S = st[b.idx[k]];
  const GeoD g = S.g;
  const TermsD T = S.rhs;
  ResIn rin{nullptr, nullptr, 0.0, 0.0, 0.0};
  if constexpr (RES) rin = rb.r[k];
  rhs_tile<kBatchBX, kBatchBY, RES>(tile, g, T, b.y[k], S.hd, b.out[k], rin, lt % S.ntx,
                                    lt / S.ntx);
}

__global__ __launch_bounds__(kBatchBX* kBatchBY) void k_jtv_batch(
    const BatchStatic* __restrict__ st, StencilBatch b) {
  int k, lt;
  batch_tile(b, k, lt);
  const BatchStatic& S = st[b.idx[k]];
  const GeoD g = S.g;
  const TermsD T = S.jv;
  const int gx = (lt % S.ntx) * kBatchBX + threadIdx.x;
  const int gy = (lt / S.ntx) * kBatchBY + threadIdx.y;
  if (gx >= g.ng || gy >= g.ny) return;
  const int cell = gy * g.ng + gx, N = g.ncell;
  double o[3];
  jtv_cell<false>(g, T, b.y[k], VPlain{b.v[k]}, cell, gx, gy, 0.0, o);
  double* Jv = b.out[k];
  Jv[cell] = o[0]; Jv[N + cell] = o[1]; Jv[2 * N + cell] = o[2];
}

// blockIdx.y = instance k; cells beyond its grid return
__global__ void k_normalize_batch(const BatchStatic* __restrict__ st, StencilBatch b) {
  const int k = blockIdx.y;
  const BatchStatic& S = st[b.idx[k]];
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (S.norm == LLG_NORM_NONE || cell >= S.g.ncell) return;
  normalize_cell(const_cast<double*>(b.y[k]), S.g.ncell, S.norm, cell);
}


// Mean magnetisation of the active cells: kMeanBlocks workgroups sum
// three components over a fixed cell partition (grid-stride), a second
// launch sums the partials in slot order -- the same bits on every call.
constexpr int kMeanBlocks = 512;
__global__ __launch_bounds__(256) void k_mean_partial(const double* __restrict__ y,
                                                      const unsigned char* __restrict__ act,
                                                      int ncell, double* __restrict__ part) {
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < ncell; i += kMeanBlocks * 256)
    if (act[i]) {
      s[0] += y[i];
      s[1] += y[ncell + i];
      s[2] += y[2 * (long)ncell + i];
    }
  __shared__ double red[3][256];
  for (int a = 0; a < 3; a++) red[a][threadIdx.x] = s[a];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; a++) red[a][threadIdx.x] += red[a][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[threadIdx.x * kMeanBlocks + blockIdx.x] = red[threadIdx.x][0];
}
__global__ __launch_bounds__(256) void k_mean_final(double* __restrict__ part) {
  __shared__ double red[3][256];
  for (int a = 0; a < 3; a++)
    red[a][threadIdx.x] = part[a * kMeanBlocks + threadIdx.x] +
                          part[a * kMeanBlocks + 256 + threadIdx.x];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; a++) red[a][threadIdx.x] += red[a][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * kMeanBlocks + threadIdx.x] = red[threadIdx.x][0];
}
static_assert(kMeanBlocks == 512, "k_mean_final folds exactly 2 x 256 partials");

}  // namespace

// ============================================================ context
struct PrecondData {
  int ng, ny, ncell;
  double* d_P;   // 9*ncell SoA: P[s*ncell + cell]
  double* d_J;   // 9*ncell, two-phase protocol only
};

struct llg_ctx {
  llg_config cfg;
  int ncell, neq;
  int n_active, n_hole;
  hipStream_t stream;
  unsigned char* d_active;
  double* d_w;
  double* d_cmsk;
  double* d_cnsk;
  int* d_ids;
  int* d_hole_ids;
  double* d_hd;        // field of the most recent f() (read by PrecondSetup)
  bool hd_valid;
  PrecondData* pd;
  DemagData* demag;
  double n0[3];
  std::vector<unsigned char> h_active;
  std::vector<int> h_ids, h_hole_ids;
  long nf, njv, npset, npsol;
  llg_lpc* lpc = nullptr;   // learned correction applied by PrecondSolve
  int bx = kBX, by = kBY;   // 2D stencil workgroup shape (LLG_SetStencilBlock)
  double* d_mean = nullptr; // partial sums of LLG_MeanMagnetization
  // f() handed to the ensemble layer instead of launched (LLG_SetRhsHook)
  llg_rhs_hook rhs_hook = nullptr;
  void* rhs_hook_data = nullptr;
  long terms_ver = 0;       // bumped when a term set changes (batch tables re-read it)
};

namespace {

GeoD geo_of(const llg_ctx* c) {
  GeoD g;
  g.ng = c->cfg.ng; g.ny = c->cfg.ny; g.ncell = c->ncell;
  g.boundary = c->cfg.boundary; g.exch_self = c->cfg.exch_self;
  g.compact = c->cfg.exec == LLG_EXEC_COMPACT;
  g.active = c->d_active; g.w = c->d_w; g.cmsk = c->d_cmsk; g.cnsk = c->d_cnsk;
  g.ids = c->d_ids; g.hole_ids = c->d_hole_ids;
  g.n_active = c->n_active; g.n_hole = c->n_hole;
  g.alpha = c->cfg.alpha; g.chg = c->cfg.chg;
  return g;
}

void grid_for(const llg_ctx* c, dim3& grid, dim3& block) {
  if (c->cfg.exec == LLG_EXEC_COMPACT) {
    const int n = c->n_active + c->n_hole;
    block = dim3(k1D);
    grid = dim3((n + k1D - 1) / k1D);
  } else {
    block = dim3(c->bx, c->by);
    grid = dim3((c->cfg.ng + c->bx - 1) / c->bx, (c->cfg.ny + c->by - 1) / c->by);
  }
}

hipStream_t stream_of(N_Vector v, const llg_ctx* c) {
  if (v && v->ops && v->ops->nvgetvectorid && v->ops->nvgetvectorid(v) == SUNDIALS_NVEC_HIP)
    return (hipStream_t)N_VGetStream_Hip(v);
  return c->stream;
}

int launch_status(const char* what) {
  hipError_t e = hipPeekAtLastError();
  if (e != hipSuccess) {
    fprintf(stderr, ">>> ERROR in %s: kernel launch failed: %s\n", what, hipGetErrorString(e));
    set_err(std::string(what) + ": " + hipGetErrorString(e));
    return -1;
  }
  return 0;
}

template <class T>
bool upload(T** dst, const T* src, size_t n, hipStream_t s) {
  *dst = nullptr;
  if (n == 0) return true;
  if (hipMalloc(dst, sizeof(T) * n) != hipSuccess) return false;
  if (src) {
    if (hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s) != hipSuccess) return false;
  } else {
    if (hipMemsetAsync(*dst, 0, sizeof(T) * n, s) != hipSuccess) return false;
  }
  return true;
}

}  // namespace

namespace kprof {
State& state() {
  static State st;
  return st;
}
}  // namespace kprof

extern "C" {

const char* LLG_LastError(void) { return g_err.c_str(); }

int LLG_ProfileStart(int kernel_id, int capacity) {
  kprof::State& st = kprof::state();
  for (hipEvent_t e : st.ev) hipEventDestroy(e);
  st.ev.clear();
  st.used = 0;
  st.dropped = 0;
  st.every = 1;
  st.seen = 0;
  st.open = false;
  st.target = kprof::NONE;
  if (kernel_id == kprof::NONE) return 0;
  if (capacity < 1) capacity = 1;
  st.ev.resize(2 * (size_t)capacity);
  for (auto& e : st.ev)
    if (hipEventCreate(&e) != hipSuccess) return -1;
  st.target = kernel_id;
  return 0;
}

int LLG_ProfileSetEvery(int every) {
  if (every < 1) return -1;
  kprof::State& st = kprof::state();
  st.every = every;
  st.seen = 0;
  return 0;
}

void* LLG_EventCreate(void) {
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? (void*)e : nullptr;
}
int LLG_EventRecord(void* ev, void* stream) {
  return hipEventRecord((hipEvent_t)ev, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}
double LLG_EventElapsedMs(void* e0, void* e1) {
  float ms = -1.f;
  if (hipEventSynchronize((hipEvent_t)e1) != hipSuccess) return -1.0;
  hipEventElapsedTime(&ms, (hipEvent_t)e0, (hipEvent_t)e1);
  return ms;
}
void LLG_EventDestroy(void* ev) { hipEventDestroy((hipEvent_t)ev); }
int LLG_DeviceSynchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }

int LLG_ProfileStop(double* total_ms, long* launches, long* dropped) {
  kprof::State& st = kprof::state();
  st.target = kprof::NONE;
  double tot = 0.0;
  if (st.used) hipEventSynchronize(st.ev[st.used - 1]);
  for (size_t i = 0; i + 1 < st.used; i += 2) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, st.ev[i], st.ev[i + 1]);
    tot += ms;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = (long)(st.used / 2);
  if (dropped) *dropped = st.dropped;
  return 0;
}

PrecondData* Precond_Create(int ng, int ny, int ncell) {
  if (ng <= 0 || ny <= 0 || ncell != ng * ny) { set_err("Precond_Create: bad shape"); return nullptr; }
  PrecondData* pd = (PrecondData*)calloc(1, sizeof(PrecondData));
  if (!pd) return nullptr;
  pd->ng = ng; pd->ny = ny; pd->ncell = ncell;
  const size_t bytes = sizeof(double) * 9 * (size_t)ncell;
  if (hipMalloc(&pd->d_P, bytes) != hipSuccess) { free(pd); set_err("Precond_Create: hipMalloc"); return nullptr; }
  hipMemset(pd->d_P, 0, bytes);
  return pd;
}

void Precond_Destroy(PrecondData* pd) {
  if (!pd) return;
  if (pd->d_P) hipFree(pd->d_P);
  if (pd->d_J) hipFree(pd->d_J);
  free(pd);
}

llg_ctx* LLG_Create(const llg_config* cfg, const unsigned char* active, const double* w,
                    const double* cell_msk, const double* cell_nsk, void* stream) {
  if (!cfg || cfg->ng <= 0 || cfg->ny <= 0) { set_err("LLG_Create: bad grid"); return nullptr; }
  if (cfg->exec != LLG_EXEC_YMSK && cfg->exec != LLG_EXEC_COMPACT) { set_err("LLG_Create: bad exec"); return nullptr; }
  llg_ctx* c = new llg_ctx();
  c->cfg = *cfg;
  c->ncell = cfg->ng * cfg->ny;
  c->neq = 3 * c->ncell;
  c->stream = stream ? (hipStream_t)stream : (hipStream_t)NVHip_DefaultStream();
  const int N = c->ncell;
  c->h_active.resize(N);
  for (int i = 0; i < N; i++) c->h_active[i] = active ? (active[i] != 0) : 1;
  // ascending id lists (2d_i6/2d_i6.cu:514-518)
  for (int i = 0; i < N; i++) (c->h_active[i] ? c->h_ids : c->h_hole_ids).push_back(i);
  c->n_active = (int)c->h_ids.size();
  c->n_hole = (int)c->h_hole_ids.size();
  bool ok = upload(&c->d_active, c->h_active.data(), N, c->stream) &&
            upload(&c->d_w, w, w ? (size_t)N : 0, c->stream) &&
            upload(&c->d_cmsk, cell_msk, cell_msk ? (size_t)3 * N : 0, c->stream) &&
            upload(&c->d_cnsk, cell_nsk, cell_nsk ? (size_t)3 * N : 0, c->stream) &&
            upload(&c->d_ids, c->h_ids.data(), c->h_ids.size(), c->stream) &&
            upload(&c->d_hole_ids, c->h_hole_ids.data(), c->h_hole_ids.size(), c->stream) &&
            upload<double>(&c->d_hd, nullptr, (size_t)3 * N, c->stream);
  if (!ok) { set_err("LLG_Create: device allocation failed"); LLG_Destroy(c); return nullptr; }
  hipStreamSynchronize(c->stream);
  c->pd = Precond_Create(cfg->ng, cfg->ny, N);
  if (!c->pd) { LLG_Destroy(c); return nullptr; }
  if (cfg->pc_mode == LLG_PC_JOK) {
    if (hipMalloc(&c->pd->d_J, sizeof(double) * 9 * (size_t)N) != hipSuccess) {
      set_err("LLG_Create: J storage"); LLG_Destroy(c); return nullptr;
    }
    hipMemset(c->pd->d_J, 0, sizeof(double) * 9 * (size_t)N);
  }
  if (cfg->demag != LLG_DEMAG_OFF && cfg->demag_strength > 0.0) {
    c->demag = (cfg->demag == LLG_DEMAG_PADDED)
                   ? Demag_InitPadded(cfg->ng, cfg->ny, cfg->demag_thick, cfg->demag_strength)
                   : Demag_Init(cfg->ng, cfg->ny, cfg->demag_thick, cfg->demag_strength);
    if (!c->demag) { set_err(std::string("LLG_Create: demag: ") + Demag_LastError()); LLG_Destroy(c); return nullptr; }
    Demag_SetStream(c->demag, c->stream);
    if (cfg->demag_fp32) Demag_SetPrecision(c->demag, 1);
    Demag_GetSelfCoupling(c->demag, &c->n0[0], &c->n0[1], &c->n0[2]);
  }
  hipDeviceSynchronize();
  return c;
}

void LLG_Destroy(llg_ctx* c) {
  if (c) hipFree(c->d_mean);
  if (!c) return;
  hipFree(c->d_active); hipFree(c->d_w); hipFree(c->d_cmsk); hipFree(c->d_cnsk);
  hipFree(c->d_ids); hipFree(c->d_hole_ids); hipFree(c->d_hd);
  Precond_Destroy(c->pd);
  Demag_Destroy(c->demag);
  delete c;
}

int LLG_GetCounts(llg_ctx* c, int* ncell, int* n_active, int* n_inactive) {
  if (!c) return -1;
  if (ncell) *ncell = c->ncell;
  if (n_active) *n_active = c->n_active;
  if (n_inactive) *n_inactive = c->n_hole;
  return 0;
}

int LLG_GetMask(llg_ctx* c, double* h) {
  if (!c || !h) return -1;
  const int N = c->ncell;
  for (int i = 0; i < N; i++) {
    const double m = c->h_active[i] ? 1.0 : 0.0;
    h[i] = h[N + i] = h[2 * N + i] = m;
  }
  return 0;
}

int LLG_GetActiveIds(llg_ctx* c, int* ids, int* hole_ids) {
  if (!c) return -1;
  if (ids && c->n_active)
    hipMemcpy(ids, c->d_ids, sizeof(int) * c->n_active, hipMemcpyDeviceToHost);
  if (hole_ids && c->n_hole)
    hipMemcpy(hole_ids, c->d_hole_ids, sizeof(int) * c->n_hole, hipMemcpyDeviceToHost);
  return 0;
}

int LLG_GetHdemag(llg_ctx* c, double* h) {
  if (!c || !h) return -1;
  hipStreamSynchronize(c->stream);
  return hipMemcpy(h, c->d_hd, sizeof(double) * 3 * (size_t)c->ncell, hipMemcpyDeviceToHost) ==
                 hipSuccess ? 0 : -1;
}

int LLG_GetPinv(llg_ctx* c, double* h) {
  if (!c || !h) return -1;
  hipDeviceSynchronize();
  return hipMemcpy(h, c->pd->d_P, sizeof(double) * 9 * (size_t)c->ncell, hipMemcpyDeviceToHost) ==
                 hipSuccess ? 0 : -1;
}

DemagData* LLG_GetDemag(llg_ctx* c) { return c ? c->demag : nullptr; }
void* LLG_GetStream(llg_ctx* c) { return c ? (void*)c->stream : nullptr; }

int LLG_GetCallCounts(llg_ctx* c, long* nf, long* njv, long* nps, long* npv) {
  if (!c) return -1;
  if (nf) *nf = c->nf;
  if (njv) *njv = c->njv;
  if (nps) *nps = c->npset;
  if (npv) *npv = c->npsol;
  return 0;
}

// f(): [normalize] -> demag (overwrite hd) -> RHS (compact zeroes holes in-launch);
// with rin the RHS pass writes the Newton residual built from f instead of f
static int rhs_launch(llg_ctx* c, N_Vector y, N_Vector out, const ResIn* rin) {
  double* yd = N_VGetDeviceArrayPointer(y);
  double* fd = N_VGetDeviceArrayPointer(out);
  if (c->rhs_hook) {   // the ensemble batches this call with other instances'
    llg_res_in r{};
    if (rin) r = llg_res_in{rin->zn1, rin->ycor, rin->c0, rin->c1, rin->c2};
    return c->rhs_hook(c->rhs_hook_data, c, yd, fd, rin ? &r : nullptr);
  }
  c->nf++;
  hipStream_t s = stream_of(out, c);
  if (c->cfg.normalize != LLG_NORM_NONE) {
    hipLaunchKernelGGL(k_normalize, dim3((c->ncell + k1D - 1) / k1D), dim3(k1D), 0, s, yd,
                       c->ncell, c->cfg.normalize);
  }
  const double* hd = nullptr;
  if (c->demag) {
    Demag_SetStream(c->demag, s);
    if (c->cfg.rhs.window && c->d_w) Demag_ApplyWindowed(c->demag, yd, c->d_w, c->d_hd);
    else Demag_Apply(c->demag, yd, c->d_hd);
    if (Demag_LastError()[0]) { set_err(Demag_LastError()); return -1; }
    hd = c->d_hd;
  }
  const GeoD g = geo_of(c);
  const TermsD T = to_dev(c->cfg.rhs);
  const ResIn ri = rin ? *rin : ResIn{nullptr, nullptr, 0.0, 0.0, 0.0};
  if (c->cfg.exec == LLG_EXEC_YMSK) {
    dim3 block(32, 8), grid((c->cfg.ng + 31) / 32, (c->cfg.ny + 7) / 8);
    kprof::begin(kprof::RHS, s);
    if (rin) hipLaunchKernelGGL((k_rhs_tiled<32, 8, true>), grid, block, 0, s, g, T, yd, hd, fd, ri);
    else hipLaunchKernelGGL((k_rhs_tiled<32, 8, false>), grid, block, 0, s, g, T, yd, hd, fd, ri);
    kprof::end(kprof::RHS, s);
  } else {
    dim3 grid, block;
    grid_for(c, grid, block);
    kprof::begin(kprof::RHS, s);
    if (rin) hipLaunchKernelGGL(k_rhs<true>, grid, block, 0, s, g, T, yd, hd, fd, ri);
    else hipLaunchKernelGGL(k_rhs<false>, grid, block, 0, s, g, T, yd, hd, fd, ri);
    kprof::end(kprof::RHS, s);
  }
  return launch_status("f");
}

int LLG_Rhs(sunrealtype t, N_Vector y, N_Vector ydot, void* user_data) {
  (void)t;
  llg_ctx* c = (llg_ctx*)user_data;
  if (!c) return -1;
  return rhs_launch(c, y, ydot, nullptr);
}

int LLG_RhsResidual(sunrealtype t, N_Vector y, const sunrealtype* cf, N_Vector zn1,
                    N_Vector ycor, N_Vector res, void* user_data) {
  (void)t;
  llg_ctx* c = (llg_ctx*)user_data;
  if (!c || !cf || !y || !zn1 || !ycor || !res) return -1;
  if (c->cfg.exec == LLG_EXEC_COMPACT) return 2;   // holes: the plain RHS and combination
  const double* rd = N_VGetDeviceArrayPointer(res);
  if (rd == N_VGetDeviceArrayPointer(y) || rd == N_VGetDeviceArrayPointer(zn1) ||
      rd == N_VGetDeviceArrayPointer(ycor)) {
    set_err("LLG_RhsResidual: res aliases an input");
    return -1;
  }
  const ResIn ri{N_VGetDeviceArrayPointer(zn1), N_VGetDeviceArrayPointer(ycor), cf[0], cf[1],
                 cf[2]};
  return rhs_launch(c, y, res, &ri);
}

int JtvProduct(N_Vector v, N_Vector Jv, sunrealtype t, N_Vector y, N_Vector fy, void* user_data,
               N_Vector tmp) {
  (void)t; (void)fy; (void)tmp;
  llg_ctx* c = (llg_ctx*)user_data;
  if (!c) return -1;
  c->njv++;
  hipStream_t s = stream_of(Jv, c);
  dim3 grid, block;
  grid_for(c, grid, block);
  kprof::begin(kprof::JTV, s);
  hipLaunchKernelGGL(k_jtv<false>, grid, block, 0, s, geo_of(c), to_dev(c->cfg.jv),
                     (const double*)N_VGetDeviceArrayPointer(y),
                     (const double*)N_VGetDeviceArrayPointer(v), N_VGetDeviceArrayPointer(Jv),
                     0.0);
  kprof::end(kprof::JTV, s);
  return launch_status("jtv");
}

int LLG_Atimes(N_Vector v, N_Vector z, sunrealtype gamma, sunrealtype t, N_Vector y,
               N_Vector fy, void* user_data) {
  (void)t; (void)fy;
  llg_ctx* c = (llg_ctx*)user_data;
  if (!c || !v || !z || !y || N_VGetDeviceArrayPointer(v) == N_VGetDeviceArrayPointer(z))
    return -1;
  c->njv++;
  hipStream_t s = stream_of(z, c);
  dim3 grid, block;
  grid_for(c, grid, block);
  kprof::begin(kprof::JTV, s);
  hipLaunchKernelGGL(k_jtv<true>, grid, block, 0, s, geo_of(c), to_dev(c->cfg.jv),
                     (const double*)N_VGetDeviceArrayPointer(y),
                     (const double*)N_VGetDeviceArrayPointer(v), N_VGetDeviceArrayPointer(z),
                     -gamma);
  kprof::end(kprof::JTV, s);
  return launch_status("atimes");
}

int PrecondSetup(sunrealtype t, N_Vector y, N_Vector fy, sunbooleantype jok,
                 sunbooleantype* jcurPtr, sunrealtype gamma, void* user_data) {
  (void)t; (void)fy;
  llg_ctx* c = (llg_ctx*)user_data;
  if (!c) return -1;
  c->npset++;
  const bool two_phase = c->cfg.pc_mode == LLG_PC_JOK;
  const int rebuild = (!two_phase || !jok) ? 1 : 0;
  *jcurPtr = rebuild ? SUNTRUE : SUNFALSE;
  hipStream_t s = stream_of(y, c);
  dim3 grid, block;
  GeoD g = geo_of(c);
  if (c->cfg.exec == LLG_EXEC_COMPACT) {
    block = dim3(k1D);
    grid = dim3((c->n_active + k1D - 1) / k1D > 0 ? (c->n_active + k1D - 1) / k1D : 1);
  } else {
    grid_for(c, grid, block);
  }
  kprof::begin(kprof::BUILD_P, s);
  hipLaunchKernelGGL(k_build_P, grid, block, 0, s, g, to_dev(c->cfg.pc),
                     (const double*)N_VGetDeviceArrayPointer(y),
                     (const double*)(c->demag ? c->d_hd : nullptr), c->n0[0], c->n0[1],
                     c->n0[2], gamma, rebuild, c->pd->d_J, c->pd->d_P);
  kprof::end(kprof::BUILD_P, s);
  return launch_status("PrecondSetup");
}

int PrecondSolve(sunrealtype t, N_Vector y, N_Vector fy, N_Vector r, N_Vector z,
                 sunrealtype gamma, sunrealtype delta, int lr, void* user_data) {
  (void)t; (void)y; (void)fy; (void)gamma; (void)delta; (void)lr;
  llg_ctx* c = (llg_ctx*)user_data;
  if (!c) return -1;
  c->npsol++;
  hipStream_t s = stream_of(z, c);
  const int n = (c->cfg.exec == LLG_EXEC_COMPACT) ? c->n_active + c->n_hole : c->ncell;
  // with a learned correction the block-Jacobi output goes to the
  // network's input buffer and z = u + active N(u) is formed from it
  double* u = c->lpc ? LLG_LpcInputBuffer(c->lpc) : N_VGetDeviceArrayPointer(z);
  kprof::begin(kprof::APPLY_P, s);
  hipLaunchKernelGGL(k_apply_P, dim3((n + k1D - 1) / k1D), dim3(k1D), 0, s, geo_of(c),
                     c->cfg.pc.gate, (const double*)c->pd->d_P,
                     (const double*)N_VGetDeviceArrayPointer(r), u);
  kprof::end(kprof::APPLY_P, s);
  if (c->lpc &&
      LLG_LpcApply(c->lpc, N_VGetDeviceArrayPointer(y), gamma, u, N_VGetDeviceArrayPointer(z)) != 0)
    return -1;
  return launch_status("PrecondSolve");
}

}  // extern "C"

namespace {
// queue the operator chain vk = s .* gate P ((v + (-gamma) J v), v = (c vl) / s,
// with c by value or (cdev) from device memory; skip: device exit flag
int launch_krylov(llg_ctx* ctx, sunrealtype c, const double* cdev, const int* skip, N_Vector vl,
                  N_Vector s, N_Vector vk, sunrealtype gamma, N_Vector y) {
  // the in-kernel preconditioner is the plain block-Jacobi one
  if (!ctx || ctx->lpc || !ctx->pd || !vl || !s || !vk || !y) return -1;
  const double* dvl = N_VGetDeviceArrayPointer(vl);
  const double* ds = N_VGetDeviceArrayPointer(s);
  double* dvk = N_VGetDeviceArrayPointer(vk);
  if (dvk == dvl || dvk == ds || dvk == N_VGetDeviceArrayPointer(y)) return -1;
  hipStream_t st = stream_of(vk, ctx);
  kprof::begin(kprof::JTV, st);
  if (ctx->cfg.exec == LLG_EXEC_YMSK) {
    constexpr int BX = 32, BY = 8;
    const dim3 grid((ctx->cfg.ng + BX - 1) / BX, (ctx->cfg.ny + BY - 1) / BY);
    hipLaunchKernelGGL((k_krylov_tiled<BX, BY>), grid, dim3(BX, BY), 0, st, geo_of(ctx),
                       to_dev(ctx->cfg.jv), ctx->cfg.pc.gate, (const double*)ctx->pd->d_P,
                       (const double*)N_VGetDeviceArrayPointer(y), dvl, ds, c, -gamma, dvk,
                       cdev, skip);
  } else {
    dim3 grid, block;
    grid_for(ctx, grid, block);
    hipLaunchKernelGGL(k_krylov, grid, block, 0, st, geo_of(ctx), to_dev(ctx->cfg.jv),
                       ctx->cfg.pc.gate, (const double*)ctx->pd->d_P,
                       (const double*)N_VGetDeviceArrayPointer(y), dvl, ds, c, -gamma, dvk,
                       cdev, skip);
  }
  kprof::end(kprof::JTV, st);
  return launch_status("krylov");
}

template <int L, bool PRE>
void launch_arnoldi_t(llg_ctx* ctx, hipStream_t st, int nwg, int ntx, int nty, const double* y,
                      const double* s, double mgamma, const ArnIn& in, double* wout,
                      double* partials, const double* coef, const int* skip) {
  hipLaunchKernelGGL((k_arnoldi<L, PRE, 32, 8>), dim3(nwg), dim3(32, 8), 0, st, geo_of(ctx),
                     to_dev(ctx->cfg.jv), ctx->cfg.pc.gate, (const double*)ctx->pd->d_P, y, s,
                     mgamma, in, wout, partials, ntx, nty, coef, skip);
}

struct DevOpData {
  llg_ctx* ctx;
  double gamma;
  N_Vector y;
};
int dev_op(const double* cdev, const int* skip, N_Vector vl, N_Vector s, N_Vector vk, void* d) {
  const DevOpData* o = static_cast<const DevOpData*>(d);
  return launch_krylov(o->ctx, 0.0, cdev, skip, vl, s, vk, o->gamma, o->y);
}
}  // namespace

extern "C" {

int LLG_Krylov(sunrealtype c, N_Vector vl, N_Vector s, N_Vector vk, sunrealtype gamma,
               sunrealtype t, N_Vector y, N_Vector fy, void* user_data) {
  (void)t; (void)fy;
  llg_ctx* ctx = (llg_ctx*)user_data;
  const int rc = launch_krylov(ctx, c, nullptr, nullptr, vl, s, vk, gamma, y);
  if (rc == 0) {
    ctx->njv++;
    ctx->npsol++;
  }
  return rc;
}

int LLG_SpgmrSolve(N_Vector t0, N_Vector* V, int maxl, N_Vector s, N_Vector x,
                   sunrealtype delta, sunrealtype gamma, sunrealtype t, N_Vector y, N_Vector fy,
                   void* user_data, int* status, int* iters, sunrealtype* r_norm,
                   sunrealtype* resnorm) {
  (void)t; (void)fy;
  llg_ctx* ctx = (llg_ctx*)user_data;
  if (!ctx || ctx->lpc || !ctx->pd || !y) return -1;
  DevOpData d{ctx, gamma, y};
  const int rc = NVHip_SpgmrSolve(dev_op, &d, t0, V, maxl, s, x, delta, status, iters, r_norm,
                                  resnorm);
  if (rc == 0 && *status != NVHIP_SPGMR_FALLBACK) {
    ctx->njv += *iters;
    ctx->npsol += *iters;
  }
  return rc;
}

}  // extern "C"

namespace {
// queue one lagged pass (its inputs in `in`) and its finish; the kernel reads
// in.win, in.bx, s, y and U[j < l] at halo cells other workgroups own, so
// nothing it writes may alias them
// lag_step >= 0: the pass's sums go to the device-driven SPGMR step
// (NVHip_LagStep, last = lag_step) instead of a finish the host waits for
// (ticket unused)
int queue_pass(llg_ctx* ctx, int l, bool pre, ArnIn& in, const double* coef_dev,
               const int* skip_dev, N_Vector wout, N_Vector s, sunrealtype gamma, N_Vector y,
               unsigned* ticket, int lag_step = -1) {
  double* dw = N_VGetDeviceArrayPointer(wout);
  const double* ds = N_VGetDeviceArrayPointer(s);
  const double* dy = N_VGetDeviceArrayPointer(y);
  const double* reads[kArnMax + 6] = {in.win, ds, dy};
  int nr = 3;
  for (int j = 0; j < l; j++) reads[nr++] = in.U[j];
  if (l == 0 && !pre) reads[nr++] = in.U[0];
  for (int k = 0; k < in.nb; k++) reads[nr++] = in.bx[k];
  for (int i = 0; i < nr; i++)
    if (reads[i] && (reads[i] == dw || reads[i] == in.ul)) return -1;
  if (in.ul && in.ul == dw) return -1;
  int cap = 0;
  double* partials = NVHip_ReducePartials(wout, &cap);
  if (!partials) return -1;
  constexpr int BX = 32, BY = 8;
  const int K = l + 3 + (pre ? 1 : 0);
  const int ntx = (ctx->cfg.ng + BX - 1) / BX, nty = (ctx->cfg.ny + BY - 1) / BY;
  // one resident round (four 256-thread workgroups per CU on 256 CUs at
  // this kernel's register use); larger grids take several tiles each
  int nwg = ntx * nty;
  if (nwg > 1024) nwg = 1024;
  if (nwg * K > cap) nwg = cap / K;
  hipStream_t st = stream_of(wout, ctx);
  kprof::begin(kprof::JTV, st);
  const double mg = -gamma;
  const double* cd = coef_dev;
  const int* sk = skip_dev;
  switch (l) {
    case 0:
      if (pre) launch_arnoldi_t<0, true>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk);
      else launch_arnoldi_t<0, false>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk);
      break;
    case 1: launch_arnoldi_t<1, false>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk); break;
    case 2: launch_arnoldi_t<2, false>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk); break;
    case 3: launch_arnoldi_t<3, false>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk); break;
    case 4: launch_arnoldi_t<4, false>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk); break;
    case 5: launch_arnoldi_t<5, false>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk); break;
    case 6: launch_arnoldi_t<6, false>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk); break;
    default: launch_arnoldi_t<7, false>(ctx, st, nwg, ntx, nty, dy, ds, mg, in, dw, partials, cd, sk); break;
  }
  kprof::end(kprof::JTV, st);
  if (launch_status("arnoldi") != 0) return -1;
  if (lag_step >= 0) {
    if (NVHip_LagStep(wout, K, nwg, l, lag_step) != 0) return -1;
  } else if (NVHip_ReduceFinishAsync(wout, K, nwg, skip_dev, ticket) != 0) {
    return -1;
  }
  ctx->njv++;
  ctx->npsol += pre ? 2 : 1;
  return 0;
}

bool pass_ok(const llg_ctx* ctx) {
  return ctx && !ctx->lpc && ctx->pd && ctx->cfg.exec == LLG_EXEC_YMSK;
}
}  // namespace

extern "C" {

int LLG_ArnoldiQueue(int l, int pre, sunrealtype a, const sunrealtype* b, const double* coef_dev,
                     const int* skip_dev, N_Vector win, N_Vector* U, N_Vector wout, N_Vector s,
                     sunrealtype gamma, sunrealtype t, N_Vector y, N_Vector fy, void* user_data,
                     unsigned* ticket) {
  (void)t; (void)fy;
  llg_ctx* ctx = (llg_ctx*)user_data;
  if (!pass_ok(ctx) || !U || !wout || !s || !y || !ticket || l < 0 || l >= kArnMax ||
      (pre && l != 0) || (l > 0 && (!win || (!b && !coef_dev))) || (pre && !win))
    return -1;
  ArnIn in{};
  in.win = win ? N_VGetDeviceArrayPointer(win) : nullptr;
  in.a = a;
  for (int j = 0; j <= l; j++)
    if (!U[j]) return -1;
  for (int j = 0; j < l; j++) {
    in.U[j] = N_VGetDeviceArrayPointer(U[j]);
    in.b[j] = b ? b[j] : 0.0;
  }
  if (l == 0 && !pre) in.U[0] = N_VGetDeviceArrayPointer(U[0]);
  in.ul = (l > 0 || pre) ? N_VGetDeviceArrayPointer(U[l]) : nullptr;
  return queue_pass(ctx, l, pre != 0, in, coef_dev, skip_dev, wout, s, gamma, y, ticket);
}

int LLG_ArnoldiLag(int l, int pre, int last, const double* coef_dev, const int* skip_dev,
                   N_Vector win, N_Vector* U, N_Vector wout, N_Vector s, sunrealtype gamma,
                   sunrealtype t, N_Vector y, N_Vector fy, void* user_data) {
  (void)t; (void)fy;
  llg_ctx* ctx = (llg_ctx*)user_data;
  if (!pass_ok(ctx) || !U || !wout || !s || !y || !win || l < 0 || l >= kArnMax ||
      (pre != 0) != (l == 0) || (l > 0 && (!coef_dev || !skip_dev)))
    return -1;
  ArnIn in{};
  in.win = N_VGetDeviceArrayPointer(win);
  in.a = 1.0;
  for (int j = 0; j <= l; j++)
    if (!U[j]) return -1;
  for (int j = 0; j < l; j++) in.U[j] = N_VGetDeviceArrayPointer(U[j]);
  in.ul = N_VGetDeviceArrayPointer(U[l]);
  return queue_pass(ctx, l, pre != 0, in, l > 0 ? coef_dev : nullptr, l > 0 ? skip_dev : nullptr,
                    wout, s, gamma, y, nullptr, last != 0 ? 1 : 0);
}

int LLG_ArnoldiFirst(int nb, const sunrealtype* c, N_Vector* X, N_Vector* U, N_Vector wout,
                     N_Vector s, sunrealtype gamma, sunrealtype t, N_Vector y, N_Vector fy,
                     void* user_data, sunrealtype* r) {
  (void)t; (void)fy;
  llg_ctx* ctx = (llg_ctx*)user_data;
  if (!pass_ok(ctx) || nb < 1 || nb > 3 || !c || !X || !U || !U[0] || !wout || !s || !y || !r)
    return -1;
  ArnIn in{};
  in.nb = nb;
  for (int k = 0; k < nb; k++) {
    if (!X[k]) return -1;
    in.bx[k] = N_VGetDeviceArrayPointer(X[k]);
    in.bc[k] = c[k];
  }
  in.a = 1.0;
  in.ul = N_VGetDeviceArrayPointer(U[0]);
  unsigned tk = 0;
  if (queue_pass(ctx, 0, true, in, nullptr, nullptr, wout, s, gamma, y, &tk) != 0) return -1;
  return NVHip_ReduceWait(wout, tk, 4, r);
}

int LLG_Arnoldi(int l, int pre, sunrealtype a, const sunrealtype* b, N_Vector win, N_Vector* U,
                N_Vector wout, N_Vector s, sunrealtype gamma, sunrealtype t, N_Vector y,
                N_Vector fy, void* user_data, sunrealtype* r) {
  if (!r) return -1;
  unsigned tk = 0;
  if (LLG_ArnoldiQueue(l, pre, a, b, nullptr, nullptr, win, U, wout, s, gamma, t, y, fy,
                       user_data, &tk) != 0)
    return -1;
  return NVHip_ReduceWait(wout, tk, l + 3 + (pre ? 1 : 0), r);
}

int LLG_SetStencilBlock(llg_ctx* ctx, int bx, int by) {
  if (!ctx || bx < 1 || by < 1 || bx * by > 256 || (bx * by) % 64) return -1;
  ctx->bx = bx;
  ctx->by = by;
  return 0;
}

int LLG_Autotune(llg_ctx* ctx, int reps, int* choice) {
  if (!ctx || reps < 1) return -1;
  if (ctx->cfg.exec == LLG_EXEC_COMPACT) {   // 1D launches: nothing to tune
    if (choice) { choice[0] = k1D; choice[1] = 1; }
    return 0;
  }
  const int shapes[][2] = {{64, 4}, {32, 8}, {16, 16}, {128, 2}, {64, 2}, {32, 4}, {256, 1}};
  double *y = nullptr, *v = nullptr, *o = nullptr;
  const size_t nb = sizeof(double) * 3 * (size_t)ctx->ncell;
  if (hipMalloc(&y, nb) != hipSuccess) return -1;
  if (hipMalloc(&v, nb) != hipSuccess) { hipFree(y); return -1; }
  if (hipMalloc(&o, nb) != hipSuccess) { hipFree(y); hipFree(v); return -1; }
  hipMemsetAsync(y, 0, nb, ctx->stream);
  hipMemsetAsync(v, 0, nb, ctx->stream);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  float best = 1e30f;
  int bbx = ctx->bx, bby = ctx->by;
  for (const auto& sh : shapes) {
    ctx->bx = sh[0];
    ctx->by = sh[1];
    dim3 grid, block;
    grid_for(ctx, grid, block);
    // the Jv stencil stands for the stencil family (RHS, Jv, build/apply P)
    hipLaunchKernelGGL(k_jtv<false>, grid, block, 0, ctx->stream, geo_of(ctx),
                       to_dev(ctx->cfg.jv), (const double*)y, (const double*)v, o, 0.0);
    hipEventRecord(e0, ctx->stream);
    for (int i = 0; i < reps; i++)
      hipLaunchKernelGGL(k_jtv<false>, grid, block, 0, ctx->stream, geo_of(ctx),
                         to_dev(ctx->cfg.jv), (const double*)y, (const double*)v, o, 0.0);
    hipEventRecord(e1, ctx->stream);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) { best = ms; bbx = sh[0]; bby = sh[1]; }
  }
  ctx->bx = bbx;
  ctx->by = bby;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(y);
  hipFree(v);
  hipFree(o);
  if (choice) { choice[0] = bbx; choice[1] = bby; }
  if (ctx->demag) {
    int dc[6];
    if (Demag_Autotune(ctx->demag, reps, dc) != 0) return -1;
    if (choice)
      for (int i = 0; i < 6; i++) choice[2 + i] = dc[i];
  }
  return launch_status("LLG_Autotune");
}

int LLG_SetLearnedPrecond(llg_ctx* ctx, llg_lpc* p) {
  if (!ctx) return -1;
  ctx->lpc = p;
  return 0;
}

int LLG_GetGrid(llg_ctx* ctx, int* ng, int* ny) {
  if (!ctx) return -1;
  if (ng) *ng = ctx->cfg.ng;
  if (ny) *ny = ctx->cfg.ny;
  return 0;
}

const unsigned char* LLG_GetActiveDevice(llg_ctx* ctx) { return ctx ? ctx->d_active : nullptr; }

int LLG_SetAppliedField(llg_ctx* ctx, int which, const double* happ) {
  if (!ctx || !happ || which < 1 || which > 7) {
    set_err("LLG_SetAppliedField: NULL argument or term-set mask outside 1..7");
    return -1;
  }
  for (int a = 0; a < 3; a++) {
    if (which & LLG_TERMS_RHS) ctx->cfg.rhs.happ[a] = happ[a];
    if (which & LLG_TERMS_JV) ctx->cfg.jv.happ[a] = happ[a];
    if (which & LLG_TERMS_PC) ctx->cfg.pc.happ[a] = happ[a];
  }
  ctx->terms_ver++;
  return 0;
}

int LLG_GetAppliedField(llg_ctx* ctx, int which, double* happ) {
  if (!ctx || !happ) return -1;
  const llg_terms* t = which == LLG_TERMS_RHS ? &ctx->cfg.rhs
                     : which == LLG_TERMS_JV  ? &ctx->cfg.jv
                     : which == LLG_TERMS_PC  ? &ctx->cfg.pc : nullptr;
  if (!t) return -1;
  for (int a = 0; a < 3; a++) happ[a] = t->happ[a];
  return 0;
}

int LLG_MeanMagnetization(llg_ctx* ctx, N_Vector y, double* m) {
  if (!ctx || !y || !m || N_VGetLength_Hip(y) != (sunindextype)ctx->neq) return -1;
  if (!ctx->d_mean && hipMalloc(&ctx->d_mean, sizeof(double) * (3 * kMeanBlocks + 3)) != hipSuccess) {
    set_err("LLG_MeanMagnetization: hipMalloc failed");
    return -1;
  }
  hipStream_t s = stream_of(y, ctx);
  hipLaunchKernelGGL(k_mean_partial, dim3(kMeanBlocks), dim3(256), 0, s,
                     (const double*)N_VGetDeviceArrayPointer_Hip(y), ctx->d_active, ctx->ncell,
                     ctx->d_mean);
  hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(256), 0, s, ctx->d_mean);
  if (launch_status("LLG_MeanMagnetization")) return -1;
  double sum[3];
  if (hipMemcpyAsync(sum, ctx->d_mean + 3 * kMeanBlocks, sizeof(sum), hipMemcpyDeviceToHost, s) !=
          hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return -1;
  const double n = ctx->n_active > 0 ? (double)ctx->n_active : 1.0;
  for (int a = 0; a < 3; a++) m[a] = sum[a] / n;
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------- batches
struct llg_batch {
  std::vector<llg_ctx*> ctx;
  std::vector<long> ver;          // terms_ver of each context at the last upload
  std::vector<BatchStatic> h_st;
  BatchStatic* d_st = nullptr;
};

namespace {

BatchStatic static_of(const llg_ctx* c) {
  BatchStatic S{};
  S.g = geo_of(c);
  S.rhs = to_dev(c->cfg.rhs);
  S.jv = to_dev(c->cfg.jv);
  S.hd = c->demag ? c->d_hd : nullptr;
  S.ntx = (c->cfg.ng + kBatchBX - 1) / kBatchBX;
  S.nty = (c->cfg.ny + kBatchBY - 1) / kBatchBY;
  S.norm = c->cfg.normalize;
  return S;
}

// re-read the term sets of contexts changed since the last upload (a field
// ramp); the table is rewritten after the stream's earlier batches finished
int batch_refresh(llg_batch* b, hipStream_t s) {
  bool stale = false;
  for (size_t i = 0; i < b->ctx.size(); i++) stale |= b->ver[i] != b->ctx[i]->terms_ver;
  if (!stale) return 0;
  for (size_t i = 0; i < b->ctx.size(); i++) {
    b->h_st[i] = static_of(b->ctx[i]);
    b->ver[i] = b->ctx[i]->terms_ver;
  }
  if (hipStreamSynchronize(s) != hipSuccess ||
      hipMemcpy(b->d_st, b->h_st.data(), sizeof(BatchStatic) * b->h_st.size(),
                hipMemcpyHostToDevice) != hipSuccess) {
    set_err("LLG batch: table upload failed");
    return -1;
  }
  return 0;
}

// the stencil launch of instances `ids` (at most kBatchMax)
template <class Fill, class Launch>
void stencil_chunks(const llg_batch* b, const std::vector<int>& ids, Fill&& fill, Launch&& launch) {
  for (size_t c0 = 0; c0 < ids.size(); c0 += kBatchMax) {
    const int n = (int)std::min<size_t>(kBatchMax, ids.size() - c0);
    StencilBatch sb{};
    sb.n = n;
    int t = 0;
    for (int k = 0; k < n; k++) {
      const int i = ids[c0 + k];
      sb.t0[k] = t;
      sb.idx[k] = i;
      t += b->h_st[i].ntx * b->h_st[i].nty;
      fill(sb, k, i);
    }
    sb.t0[n] = t;
    launch(sb, t, c0, n);
  }
}

}  // namespace

extern "C" {

llg_batch* LLG_BatchCreate(int n, llg_ctx* const* ctx) {
  if (n < 1 || !ctx) {
    set_err("LLG_BatchCreate: no instances");
    return nullptr;
  }
  for (int i = 0; i < n; i++) {
    if (!ctx[i] || ctx[i]->cfg.exec != LLG_EXEC_YMSK) {
      set_err("LLG_BatchCreate: NULL context or compact execution (dense-mask instances only)");
      return nullptr;
    }
  }
  llg_batch* b = new llg_batch();
  b->ctx.assign(ctx, ctx + n);
  b->ver.assign(n, -1);
  b->h_st.resize(n);
  if (hipMalloc(&b->d_st, sizeof(BatchStatic) * n) != hipSuccess) {
    set_err("LLG_BatchCreate: hipMalloc failed");
    delete b;
    return nullptr;
  }
  if (batch_refresh(b, ctx[0]->stream) != 0) {
    LLG_BatchDestroy(b);
    return nullptr;
  }
  return b;
}

int LLG_BatchSize(const llg_batch* b) { return b ? (int)b->ctx.size() : -1; }

int LLG_RhsBatch(llg_batch* b, double* const* y, double* const* ydot,
                 const llg_res_in* const* res, void* stream) {
  if (!b || !y || !ydot) return -1;
  hipStream_t s = (hipStream_t)stream;
  if (batch_refresh(b, s) != 0) return -1;
  const int n = (int)b->ctx.size();
  std::vector<int> all, plain, resid, norm;
  for (int i = 0; i < n; i++) {
    if (!y[i]) continue;
    if (!ydot[i] || ydot[i] == y[i]) {
      set_err("LLG_RhsBatch: missing or aliased output");
      return -1;
    }
    all.push_back(i);
    (res && res[i] ? resid : plain).push_back(i);
    if (b->h_st[i].norm != LLG_NORM_NONE) norm.push_back(i);
    b->ctx[i]->nf++;
  }
  // [normalize] -> demag -> RHS, the order of f()
  for (size_t c0 = 0; c0 < norm.size(); c0 += kBatchMax) {
    const int m = (int)std::min<size_t>(kBatchMax, norm.size() - c0);
    StencilBatch sb{};
    sb.n = m;
    int maxn = 0;
    for (int k = 0; k < m; k++) {
      const int i = norm[c0 + k];
      sb.idx[k] = i;
      sb.y[k] = y[i];
      maxn = std::max(maxn, b->ctx[i]->ncell);
    }
    hipLaunchKernelGGL(k_normalize_batch, dim3((maxn + k1D - 1) / k1D, m), dim3(k1D), 0, s,
                       (const BatchStatic*)b->d_st, sb);
  }
  {
    std::vector<DemagData*> dd;
    std::vector<const double*> dy, dw;
    std::vector<double*> dh;
    for (int i : all) {
      llg_ctx* c = b->ctx[i];
      if (!c->demag) continue;
      dd.push_back(c->demag);
      dy.push_back(y[i]);
      dw.push_back(c->cfg.rhs.window && c->d_w ? c->d_w : nullptr);
      dh.push_back(c->d_hd);
    }
    if (!dd.empty() && Demag_ApplyBatch((int)dd.size(), dd.data(), dy.data(), dw.data(),
                                        dh.data(), s) != 0) {
      set_err(std::string("LLG_RhsBatch: ") + Demag_LastError());
      return -1;
    }
  }
  const dim3 block(kBatchBX, kBatchBY);
  stencil_chunks(b, plain,
                 [&](StencilBatch& sb, int k, int i) { sb.y[k] = y[i]; sb.out[k] = ydot[i]; },
                 [&](const StencilBatch& sb, int tiles, size_t, int) {
                   hipLaunchKernelGGL(k_rhs_batch<false>, dim3(tiles), block, 0, s,
                                      (const BatchStatic*)b->d_st, sb, ResBatch{});
                 });
  stencil_chunks(b, resid,
                 [&](StencilBatch& sb, int k, int i) { sb.y[k] = y[i]; sb.out[k] = ydot[i]; },
                 [&](const StencilBatch& sb, int tiles, size_t c0, int m) {
                   ResBatch rb{};
                   for (int k = 0; k < m; k++) {
                     const llg_res_in& r = *res[resid[c0 + k]];
                     rb.r[k] = ResIn{r.zn1, r.ycor, r.c0, r.c1, r.c2};
                   }
                   hipLaunchKernelGGL(k_rhs_batch<true>, dim3(tiles), block, 0, s,
                                      (const BatchStatic*)b->d_st, sb, rb);
                 });
  return launch_status("LLG_RhsBatch");
}

int LLG_JtvBatch(llg_batch* b, const double* const* v, double* const* Jv,
                 const double* const* y, void* stream) {
  if (!b || !v || !Jv || !y) return -1;
  hipStream_t s = (hipStream_t)stream;
  if (batch_refresh(b, s) != 0) return -1;
  std::vector<int> ids;
  for (int i = 0; i < (int)b->ctx.size(); i++) {
    if (!v[i]) continue;
    if (!Jv[i] || !y[i] || Jv[i] == v[i] || Jv[i] == y[i]) {
      set_err("LLG_JtvBatch: missing or aliased vector");
      return -1;
    }
    ids.push_back(i);
    b->ctx[i]->njv++;
  }
  stencil_chunks(b, ids,
                 [&](StencilBatch& sb, int k, int i) {
                   sb.y[k] = y[i];
                   sb.v[k] = v[i];
                   sb.out[k] = Jv[i];
                 },
                 [&](const StencilBatch& sb, int tiles, size_t, int) {
                   hipLaunchKernelGGL(k_jtv_batch, dim3(tiles), dim3(kBatchBX, kBatchBY), 0, s,
                                      (const BatchStatic*)b->d_st, sb);
                 });
  return launch_status("LLG_JtvBatch");
}

void LLG_BatchDestroy(llg_batch* b) {
  if (!b) return;
  hipFree(b->d_st);
  delete b;
}

int LLG_SetRhsHook(llg_ctx* ctx, llg_rhs_hook fn, void* data) {
  if (!ctx) return -1;
  ctx->rhs_hook = fn;
  ctx->rhs_hook_data = fn ? data : nullptr;
  return 0;
}

}  // extern "C"
