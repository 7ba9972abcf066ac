// learned_pc.hip -- a learned correction to the block-Jacobi preconditioner:
// a three-layer 3x3 convolution network (32 channels, BF16 weights and
// activations, FP32 accumulation on MFMA) with hand-written forward,
// backward and Adam kernels.  No reference counterpart.
//
// The preconditioner stays a fixed LINEAR operator during a linear solve,
// as left-preconditioned GMRES requires:
//
//   u = P_bj^-1 r                                  (block-Jacobi, k_apply_P)
//   f = gamma * [u, m x u, (m . u) m]              (9 channels, linear in u)
//   z = u + active * C3(C2(C1(f)))                 (no activations)
//
// m is the Newton iterate's magnetisation, fixed during the solve; the three
// convolutions are linear, so z is linear in r.  The factor gamma gives the
// correction the first-order scaling of (I - gamma J)^-1 - P_bj^-1.
//
// Layout: activations cell-major [cell][C] in BF16 (C a multiple of 16),
// so an MFMA A-fragment (16 cells x 16 channels) is one 8-byte load per
// lane.  Weights: FP32 master copy W[tap][cin][cout] (tap = (dy+1)*3+dx+1),
// and a BF16 copy in fragment order for v_mfma_f32_16x16x16_bf16:
// lane l of a 64-lane wave holds A[l%16][4(l/16)+j], B[4(l/16)+j][l%16]
// and D[4(l/16)+j][l%16], j = 0..3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "llg/learned_pc.h"
#include "llg/llg_plugin.h"
#include "llg/nvector_hip.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWg = 256;            // 4 waves
constexpr int kCellsPerWave = 16;
constexpr int kC1in = 16, kC1 = 32, kC2 = 32, kC3 = 16, kNout = 3;

thread_local std::string g_lerr;

// float -> bf16, round to nearest even (finite inputs)
__device__ __host__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __host__ inline float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// neighbour of (x, y) at tap t; -1 outside the grid (zero padding) unless
// periodic
__device__ inline int tap_cell(int x, int y, int t, int ng, int ny, int periodic) {
  int xx = x + (t % 3) - 1, yy = y + (t / 3) - 1;
  if (periodic) {
    xx = xx < 0 ? xx + ng : (xx >= ng ? xx - ng : xx);
    yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
  } else if (xx < 0 || xx >= ng || yy < 0 || yy >= ny) {
    return -1;
  }
  return yy * ng + xx;
}

// out[cell][co] = sum_t sum_ci in[tap(cell,t)][ci] W[t][ci][co]
// One wave: kGroups groups of 16 consecutive cells of one row (ng % 16 ==
// 0), all outputs; the weight fragments are loaded once into registers and
// reused for every group.  Output: BF16 [cell][COUT], or FP32
// [cell][ostride] for co < nstore.
constexpr int kGroups = 4;
template <int CIN, int COUT, bool F32OUT>
__global__ __launch_bounds__(kWg) void k_conv(const uint16_t* __restrict__ in,
                                              const bf16x4* __restrict__ wf, void* out,
                                              int ng, int ny, int periodic, int nstore,
                                              int ostride) {
  constexpr int KB = CIN / 16, NB = COUT / 16;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const long ncell = (long)ng * ny;
  const long base = ((long)blockIdx.x * (kWg / 64) + wave) * kCellsPerWave * kGroups;
  if (base >= ncell) return;   // whole wave: MFMA needs every lane
  bf16x4 w[9][KB][NB];
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int kb = 0; kb < KB; kb++)
#pragma unroll
      for (int nb = 0; nb < NB; nb++) w[t][kb][nb] = wf[((t * KB + kb) * NB + nb) * 64 + l];
  for (int grp = 0; grp < kGroups; grp++) {
    const long cell0 = base + (long)grp * kCellsPerWave;
    if (cell0 >= ncell) break;   // uniform across the wave
    const int y = (int)(cell0 / ng), x = (int)(cell0 % ng) + r;
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 9; t++) {
      const int c = tap_cell(x, y, t, ng, ny, periodic);
#pragma unroll
      for (int kb = 0; kb < KB; kb++) {
        bf16x4 a = bf16x4{0, 0, 0, 0};
        if (c >= 0) a = *reinterpret_cast<const bf16x4*>(in + (long)c * CIN + kb * 16 + 4 * q);
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
          acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w[t][kb][nb], acc[nb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long cell = cell0 + 4 * q + j;
        const int co = nb * 16 + r;
        if (F32OUT) {
          if (co < nstore) reinterpret_cast<float*>(out)[cell * ostride + co] = acc[nb][j];
        } else {
          reinterpret_cast<uint16_t*>(out)[cell * COUT + co] = f2bf(acc[nb][j]);
        }
      }
    }
  }
}

// Weight gradient partials: dW[t][ci][co] = sum_cells in[tap(cell,t)][ci] g[cell][co]
// Workgroup w sums cells [w*chunk, (w+1)*chunk); its 4 waves share the
// (t, kb) pairs.  part[w][t][kb][nb][16x16] (row ci, col co).
template <int CIN, int COUT>
__global__ __launch_bounds__(kWg) void k_wgrad(const uint16_t* __restrict__ in,
                                               const uint16_t* __restrict__ g,
                                               float* __restrict__ part, int ng, int ny,
                                               int periodic, int chunk) {
  constexpr int KB = CIN / 16, NB = COUT / 16, NP = 9 * KB;
  constexpr int PW = (NP + 3) / 4;   // pairs per wave
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const long ncell = (long)ng * ny;
  const long c0 = (long)blockIdx.x * chunk;
  long c1 = c0 + chunk;
  if (c1 > ncell) c1 = ncell;
  f32x4 acc[PW][NB];
#pragma unroll
  for (int p = 0; p < PW; p++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[p][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long gc = c0; gc < c1; gc += 16) {
    // B[k = cell][n = co]: lane holds cells gc+4q+j, channel nb*16+r
    bf16x4 b[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long cell = gc + 4 * q + j;
        b[nb][j] = (short)(cell < c1 ? g[cell * COUT + nb * 16 + r] : 0);
      }
#pragma unroll
    for (int p = 0; p < PW; p++) {
      const int pair = wave + 4 * p;
      if (pair >= NP) break;
      const int t = pair / KB, kb = pair % KB;
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long cell = gc + 4 * q + j;
        short v = 0;
        if (cell < c1) {
          const int c = tap_cell((int)(cell % ng), (int)(cell / ng), t, ng, ny, periodic);
          if (c >= 0) v = (short)in[(long)c * CIN + kb * 16 + r];
        }
        a[j] = v;
      }
#pragma unroll
      for (int nb = 0; nb < NB; nb++)
        acc[p][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nb], acc[p][nb], 0, 0, 0);
    }
  }
  float* pw = part + (long)blockIdx.x * 9 * CIN * COUT;
#pragma unroll
  for (int p = 0; p < PW; p++) {
    const int pair = wave + 4 * p;
    if (pair >= NP) break;
    const int t = pair / KB, kb = pair % KB;
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int ci = kb * 16 + 4 * q + j, co = nb * 16 + r;
        pw[((long)t * CIN + ci) * COUT + co] = acc[p][nb][j];
      }
  }
}

// dW = sum over workgroups (fixed order)
__global__ void k_wgrad_reduce(const float* __restrict__ part, int nwg, int nw,
                               float* __restrict__ dw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  float s = 0.f;
  for (int w = 0; w < nwg; w++) s += part[(long)w * nw + i];
  dw[i] = s;
}

// BF16 fragments from FP32 master weights W[9][cin][cout].  transpose: the
// adjoint convolution (taps flipped, cin <-> cout) for the input gradient.
__global__ void k_make_frag(const float* __restrict__ W, int cin, int cout, int transpose,
                            bf16x4* __restrict__ frag) {
  // fragment conv shape: CIN' = transpose ? cout : cin, COUT' = transpose ? cin : cout
  const int ci_n = transpose ? cout : cin, co_n = transpose ? cin : cout;
  const int KB = ci_n / 16, NB = co_n / 16;
  const int total = 9 * KB * NB * 64;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int l = idx & 63, blk = idx >> 6;
  const int nb = blk % NB, kb = (blk / NB) % KB, t = blk / (NB * KB);
  const int r = l & 15, q = l >> 4;
  bf16x4 v;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int ci = kb * 16 + 4 * q + j, co = nb * 16 + r;   // of the fragment conv
    float w;
    if (!transpose) w = W[((long)t * cin + ci) * cout + co];
    else w = W[((long)(8 - t) * cin + co) * cout + ci];
    v[j] = (short)f2bf(w);
  }
  frag[idx] = v;
}

// features f = scale * gamma * [u, m x u, (m.u) m, 0...] (16 channels, BF16)
__global__ void k_features(const double* __restrict__ u, const double* __restrict__ y,
                           const unsigned char* __restrict__ active, double sg, long n,
                           uint16_t* __restrict__ f) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const bool on = active ? active[c] != 0 : true;
  const double u1 = u[c], u2 = u[n + c], u3 = u[2 * n + c];
  const double m1 = y[c], m2 = y[n + c], m3 = y[2 * n + c];
  const double mu = m1 * u1 + m2 * u2 + m3 * u3;
  const double v[9] = {u1, u2, u3, m2 * u3 - m3 * u2, m3 * u1 - m1 * u3, m1 * u2 - m2 * u1,
                       mu * m1, mu * m2, mu * m3};
  uint16_t* o = f + c * kC1in;
#pragma unroll
  for (int k = 0; k < 9; k++) o[k] = on ? f2bf((float)(sg * v[k])) : (uint16_t)0;
#pragma unroll
  for (int k = 9; k < kC1in; k++) o[k] = 0;
}

// training loss: pred = s*u + active*o ; target s*x ; loss partials and the
// output gradient g = 2 active (pred - target) / n (BF16, channels >= 3 zero)
__global__ void k_loss(const float* __restrict__ o, const double* __restrict__ u,
                       const double* __restrict__ x, const unsigned char* __restrict__ active,
                       double s, long n, uint16_t* __restrict__ g, double* __restrict__ part) {
  __shared__ double sm[kWg / 64];
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (c < n) {
    const bool on = active ? active[c] != 0 : true;
    uint16_t* gc = g + c * kC3;
#pragma unroll
    for (int k = 0; k < kC3; k++) gc[k] = 0;
    if (on) {
#pragma unroll
      for (int k = 0; k < kNout; k++) {
        const double d = s * u[k * n + c] + (double)o[c * 4 + k] - s * x[k * n + c];
        e2 += d * d;
        gc[k] = f2bf((float)(2.0 * d / (double)n));
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int w = 0; w < kWg / 64; w++) a += sm[w];
    part[blockIdx.x] = a;
  }
}

// z = u + active * o / (scale*gamma-free): the network output is already in
// the units of u (features carried the factor gamma; no scale at inference)
__global__ void k_apply_out(const float* __restrict__ o, const unsigned char* __restrict__ active,
                            long n, double* __restrict__ z) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (active && !active[c]) return;
#pragma unroll
  for (int k = 0; k < kNout; k++) z[k * n + c] += (double)o[c * 4 + k];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------- fused inference pass
// features -> C1 -> C2 -> C3 -> z += active * out in ONE kernel: a
// workgroup owns a 16 x 16 output tile; the features of the tile plus a
// 3-cell halo, the 32-channel layer-1 activations (2-cell halo) and the
// layer-2 activations (1-cell halo) stay in LDS (BF16), so only u, y, z and
// the active flags cross HBM.  The products run on gfx950's
// v_mfma_f32_16x16x32_bf16: 16 cells x 32 reduction values x 16 output
// channels per instruction; the reduction index is tap * C_in + channel
// (layer 1: 9 taps x 16 channels padded to 5 x 32, layers 2 and 3: one tap
// of 32 channels per instruction).  Operand layout (lane l, q = l / 16):
// A[l % 16][8q + j], B[8q + j][l % 16], D[4q + j][l % 16], j = 0..7 / 0..3;
// the weights are the A operand and the activations the B operand, so D
// is [output channel][cell].
// The weight fragments (FP32 master weights rounded to BF16) are built by
// k_make_frag32 and held in registers; every wave holds all output blocks
// of a layer.  Activations outside the grid are zero (zero padding) or wrapped
// (periodic), as in the layer-by-layer kernels; the FP32 accumulation order
// differs from theirs (K = 32 per instruction), so results agree to BF16
// rounding of the hidden layers, not bit for bit.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
constexpr int kT = 16;                        // output tile edge
constexpr int kWF = kT + 6, kW1 = kT + 4, kW2 = kT + 2;
constexpr int kNF = kWF * kWF, kN1 = kW1 * kW1, kN2 = kW2 * kW2, kNO = kT * kT;
// LDS layout of an activation box: [chunk][cell][8 channels] (BF16), one
// plane per 8 channels.  The 16 lanes of an operand read take 16
// consecutive cells of one plane: 256 contiguous bytes.  The plane stride
// (in 16-bit units) is padded to 64 bytes modulo 256, so lanes of the four
// planes that the hardware may serve in one pass start on distinct banks.
constexpr int plane_stride(int ncell) { return ncell * 8 + ((32 - (ncell * 8) % 128) % 128 + 128) % 128; }
constexpr int kPF = plane_stride(kNF), kP1 = plane_stride(kN1), kP2 = plane_stride(kN2);
constexpr int kFusedLds = 2 * (4 * kP1 + (2 * kPF > 4 * kP2 ? 2 * kPF : 4 * kP2));
// fragment steps per layer and 16-channel output blocks
constexpr int kS1steps = 5, kS2steps = 9, kS3steps = 9;
static_assert((kPF * 2) % 256 == 64 && (kP1 * 2) % 256 == 64 && (kP2 * 2) % 256 == 64,
              "plane strides 64 B modulo 256 B");

__device__ inline bf16x8 lds8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(p));
}

// Weight fragments for the 16x16x32 instruction from W[9][cin][cout] (FP32),
// lane l holding W[k = 8q + j][channel l % 16] (used in A position):
// layer with cin == 16: step s covers taps 2s, 2s+1; cin == 32: step = tap
__global__ void k_make_frag32(const float* __restrict__ W, int cin, int cout,
                              u16x8* __restrict__ frag) {
  const int NB = cout / 16, steps = cin == 16 ? 5 : 9;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= steps * NB * 64) return;
  const int l = idx & 63, blk = idx >> 6;
  const int nb = blk % NB, st = blk / NB;
  const int q = l >> 4, co = nb * 16 + (l & 15);
  u16x8 v;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    int tap, ci;
    if (cin == 16) { tap = 2 * st + (q >> 1); ci = 8 * (q & 1) + j; }
    else { tap = st; ci = 8 * q + j; }
    v[j] = tap < 9 ? f2bf(W[((long)tap * cin + ci) * cout + co]) : (uint16_t)0;
  }
  frag[idx] = v;
}

// the grid cell at tile-box coordinates, -1 outside a non-periodic grid
__device__ inline int box_cell(int gx, int gy, int ng, int ny, int periodic) {
  if (periodic) {
    gx = ((gx % ng) + ng) % ng;
    gy = ((gy % ny) + ny) % ny;
  } else if (gx < 0 || gx >= ng || gy < 0 || gy >= ny) {
    return -1;
  }
  return gy * ng + gx;
}

// One layer on the tile boxes: output box WO x WO (halo HO around the tile)
// from the input box (WO + 2) x (WO + 2), input plane stride PIN, output
// plane stride POUT.  The product is formed transposed, D[channel][cell] =
// W^T[channel][k] x A^T[k][cell], so a lane ends with 4 consecutive
// channels of ONE cell (D[4q + j][l % 16]) and stores them as one 8-byte
// word; the weight fragments in A position have the lane layout they have
// in B position (lane l: W[k = 8q + j][channel l % 16]).
template <int CIN, int STEPS, int NB, int WO, int HO, int PIN, int POUT>
__device__ inline void conv_layer(const uint16_t* __restrict__ in,
                                  const u16x8* __restrict__ frag, int x0, int y0, int ng, int ny,
                                  int periodic, uint16_t* __restrict__ out) {
  constexpr int WIN = WO + 2, NO = WO * WO;
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  // every wave holds the weights of all NB output blocks, so one operand
  // read from LDS feeds NB instructions
  bf16x8 w[NB][STEPS];
#pragma unroll
  for (int nb = 0; nb < NB; nb++)
#pragma unroll
    for (int st = 0; st < STEPS; st++)
      w[nb][st] = __builtin_bit_cast(bf16x8, frag[(st * NB + nb) * 64 + l]);
  // this lane's operand plane and, for 16 input channels, its tap parity
  const uint16_t* inq = in + (CIN == 16 ? (q & 1) : q) * PIN;
  const int todd = CIN == 16 ? (q >> 1) : 0;
  constexpr int ngrp = (NO + 15) / 16;
  for (int g = wave; g < ngrp; g += 4) {
    const int c = g * 16 + r;                  // this lane's cell (B column)
    const int cc = c < NO ? c : NO - 1;        // computed, not stored
    const int bx = cc % WO, by = cc / WO;
    const uint16_t* base = inq + (by * WIN + bx) * 8;
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int st = 0; st < STEPS; st++) {
      // tap of this step (layer 1: 2 st + parity, the tenth tap's weights
      // are zero and it reads tap 8 again)
      const int t0 = CIN == 16 ? 2 * st : st;
      const int off0 = ((t0 / 3) * WIN + t0 % 3) * 8;
      const int t1 = t0 + 1 > 8 ? 8 : t0 + 1;
      const int off1 = ((t1 / 3) * WIN + t1 % 3) * 8;
      const bf16x8 b = lds8(base + (todd ? off1 : off0));
#pragma unroll
      for (int nb = 0; nb < NB; nb++)
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[nb][st], b, acc[nb], 0, 0, 0);
    }
    if (c < NO) {
      const int gx = x0 - HO + bx, gy = y0 - HO + by;
      const bool inside = periodic || (gx >= 0 && gx < ng && gy >= 0 && gy < ny);
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        bf16x4v v;
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = (__bf16)(inside ? acc[nb][j] : 0.f);
        // channels nb*16 + 4q .. +3: plane 2 nb + q / 2, offset 4 (q % 2)
        *reinterpret_cast<bf16x4v*>(out + (2 * nb + (q >> 1)) * POUT + c * 8 + 4 * (q & 1)) = v;
      }
    }
  }
}

// layer 3 and z = u + active * o (channels 0..2 of 16) on every cell of the
// tile: the lanes with q = 0 hold channels 0..3 of cell l % 16
__device__ inline void conv_out(const uint16_t* __restrict__ in, const u16x8* __restrict__ frag,
                                int x0, int y0, int ng, int ny, long n,
                                const unsigned char* __restrict__ active,
                                const double* __restrict__ u, double* __restrict__ z) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  bf16x8 w[kS3steps];
#pragma unroll
  for (int st = 0; st < kS3steps; st++) w[st] = __builtin_bit_cast(bf16x8, frag[st * 64 + l]);
  const uint16_t* inq = in + q * kP2;
  constexpr int ngrp = kNO / 16;
  for (int g = wave; g < ngrp; g += 4) {
    const int c = g * 16 + r;
    const int bx = c % kT, by = c / kT;
    const uint16_t* base = inq + (by * kW2 + bx) * 8;
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int st = 0; st < kS3steps; st++)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
          w[st], lds8(base + ((st / 3) * kW2 + st % 3) * 8), acc, 0, 0, 0);
    const int gx = x0 + bx, gy = y0 + by;
    if (q == 0 && gx < ng && gy < ny) {
      const long cell = (long)gy * ng + gx;
      const bool on = !active || active[cell];
#pragma unroll
      for (int k = 0; k < kNout; k++) {
        const double uc = u[k * n + cell];
        z[k * n + cell] = on ? uc + (double)acc[k] : uc;
      }
    }
  }
}

__global__ __launch_bounds__(kWg) void k_lpc_fused(const double* __restrict__ y,
                                                   const unsigned char* __restrict__ active,
                                                   double sg, int ng, int ny, int periodic,
                                                   const u16x8* __restrict__ f1,
                                                   const u16x8* __restrict__ f2,
                                                   const u16x8* __restrict__ f3,
                                                   const double* __restrict__ u,
                                                   double* __restrict__ z) {
  extern __shared__ uint16_t lds[];
  uint16_t* H1 = lds;                          // 4 planes of kP1
  uint16_t* FA = lds + 4 * kP1;                // features (2 planes of kPF), later layer 2
  const long n = (long)ng * ny;
  const int tiles_x = ng / kT;
  const int x0 = (blockIdx.x % tiles_x) * kT, y0 = (blockIdx.x / tiles_x) * kT;
  // features of the tile plus a 3-cell halo: gamma * [u, m x u, (m.u) m],
  // zero off the mask and outside a non-periodic grid
  for (int i = threadIdx.x; i < kNF; i += kWg) {
    const int bx = i % kWF, by = i / kWF;
    const int c = box_cell(x0 - 3 + bx, y0 - 3 + by, ng, ny, periodic);
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = 0.f;
    if (c >= 0 && (!active || active[c])) {
      // FP32 is ample ahead of the BF16 rounding of the features
      const float s = (float)sg;
      const float u1 = s * (float)u[c], u2 = s * (float)u[n + c], u3 = s * (float)u[2 * n + c];
      const float m1 = (float)y[c], m2 = (float)y[n + c], m3 = (float)y[2 * n + c];
      const float mu = m1 * u1 + m2 * u2 + m3 * u3;
      v[0] = u1; v[1] = u2; v[2] = u3;
      v[3] = m2 * u3 - m3 * u2; v[4] = m3 * u1 - m1 * u3; v[5] = m1 * u2 - m2 * u1;
      v[6] = mu * m1; v[7] = mu * m2; v[8] = mu * m3;
    }
    bf16x8 lo, hi;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      lo[k] = (__bf16)v[k];
      hi[k] = (__bf16)v[8 + k];
    }
    *reinterpret_cast<bf16x8*>(FA + i * 8) = lo;
    *reinterpret_cast<bf16x8*>(FA + kPF + i * 8) = hi;
  }
  __syncthreads();
  conv_layer<16, kS1steps, 2, kW1, 2, kPF, kP1>(FA, f1, x0, y0, ng, ny, periodic, H1);
  __syncthreads();
  // every wave has read the features (one barrier), then H1 -> layer 2 in FA
  conv_layer<32, kS2steps, 2, kW2, 1, kP1, kP2>(H1, f2, x0, y0, ng, ny, periodic, FA);
  __syncthreads();
  // each tile writes z on its own cells only; the halos read u
  conv_out(FA, f3, x0, y0, ng, ny, n, active, u, z);
}

}  // namespace

struct llg_lpc {
  int ng, ny, periodic;
  long n;
  hipStream_t s;
  const unsigned char* d_active;   // borrowed from the context (may be null)
  // FP32 master weights, gradients, Adam moments: layers 1..3 concatenated
  float *W, *G, *M, *V, *part;
  double* d_lpart;
  int nw[3], off[3], nwtot, nwg_grad, chunk;
  long step;
  bf16x4 *frag[3], *fragT[3];
  uint16_t *f, *h1, *h2, *g3, *g2, *g1;
  float* o;
  int enabled;
  u16x8* frag32[3];   // weight fragments of the 16x16x32 instruction (fused pass)
  double* ubuf;       // 3n: the block-Jacobi output the correction reads
  int fused;          // 1: the fused inference kernel; 0: layer by layer
};

namespace {

const int kCin[3] = {kC1in, kC1, kC2};
const int kCout[3] = {kC1, kC2, kC3};

bool alloc_all(llg_lpc* p) {
  const long n = p->n;
  p->nwtot = 0;
  for (int k = 0; k < 3; k++) {
    p->nw[k] = 9 * kCin[k] * kCout[k];
    p->off[k] = p->nwtot;
    p->nwtot += p->nw[k];
  }
  p->chunk = 4096;
  p->nwg_grad = cdiv(n, p->chunk);
  bool ok = true;
  auto A = [&](void** ptr, size_t bytes) {
    ok = ok && hipMalloc(ptr, bytes) == hipSuccess;
    if (ok) ok = hipMemset(*ptr, 0, bytes) == hipSuccess;
  };
  A((void**)&p->W, sizeof(float) * p->nwtot);
  A((void**)&p->G, sizeof(float) * p->nwtot);
  A((void**)&p->M, sizeof(float) * p->nwtot);
  A((void**)&p->V, sizeof(float) * p->nwtot);
  A((void**)&p->part, sizeof(float) * (size_t)p->nwg_grad * 9 * 32 * 32);
  A((void**)&p->d_lpart, sizeof(double) * cdiv(n, kWg));
  for (int k = 0; k < 3; k++) {
    A((void**)&p->frag[k], sizeof(uint16_t) * p->nw[k]);
    A((void**)&p->fragT[k], sizeof(uint16_t) * p->nw[k]);
  }
  A((void**)&p->f, sizeof(uint16_t) * n * kC1in);
  A((void**)&p->h1, sizeof(uint16_t) * n * kC1);
  A((void**)&p->h2, sizeof(uint16_t) * n * kC2);
  A((void**)&p->g3, sizeof(uint16_t) * n * kC3);
  A((void**)&p->g2, sizeof(uint16_t) * n * kC2);
  A((void**)&p->g1, sizeof(uint16_t) * n * kC1);
  A((void**)&p->o, sizeof(float) * n * 4);
  A((void**)&p->ubuf, sizeof(double) * 3 * n);
  for (int k = 0; k < 3; k++)
    A((void**)&p->frag32[k], sizeof(u16x8) * (k == 0 ? 5 : 9) * (kCout[k] / 16) * 64);
  return ok;
}

void make_frags(llg_lpc* p) {
  for (int k = 0; k < 3; k++) {
    const int nf = (k == 0 ? 5 : 9) * (kCout[k] / 16) * 64;
    hipLaunchKernelGGL(k_make_frag32, dim3(cdiv(nf, 256)), dim3(256), 0, p->s, p->W + p->off[k],
                       kCin[k], kCout[k], p->frag32[k]);
  }
  for (int k = 0; k < 3; k++) {
    const int tot = p->nw[k] / 4;
    hipLaunchKernelGGL(k_make_frag, dim3(cdiv(tot, 256)), dim3(256), 0, p->s, p->W + p->off[k],
                       kCin[k], kCout[k], 0, p->frag[k]);
    hipLaunchKernelGGL(k_make_frag, dim3(cdiv(tot, 256)), dim3(256), 0, p->s, p->W + p->off[k],
                       kCin[k], kCout[k], 1, p->fragT[k]);
  }
}

template <int CIN, int COUT, bool F32>
void conv(llg_lpc* p, const uint16_t* in, const bf16x4* wf, void* out, int nstore = 0,
          int ostride = 0) {
  hipLaunchKernelGGL((k_conv<CIN, COUT, F32>), dim3(cdiv(p->n, 64 * kGroups)), dim3(kWg), 0,
                     p->s, in, wf,
                     out, p->ng, p->ny, p->periodic, nstore, ostride);
}

template <int CIN, int COUT>
void wgrad(llg_lpc* p, const uint16_t* in, const uint16_t* g, float* dw) {
  hipLaunchKernelGGL((k_wgrad<CIN, COUT>), dim3(p->nwg_grad), dim3(kWg), 0, p->s, in, g, p->part,
                     p->ng, p->ny, p->periodic, p->chunk);
  const int nw = 9 * CIN * COUT;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(cdiv(nw, 256)), dim3(256), 0, p->s, p->part,
                     p->nwg_grad, nw, dw);
}

// forward: features -> h1 -> h2 -> o (FP32, 3 of 4 channels)
void forward(llg_lpc* p, const double* u, const double* y, double sg) {
  hipLaunchKernelGGL(k_features, dim3(cdiv(p->n, 256)), dim3(256), 0, p->s, u, y, p->d_active,
                     sg, p->n, p->f);
  conv<kC1in, kC1, false>(p, p->f, p->frag[0], p->h1);
  conv<kC1, kC2, false>(p, p->h1, p->frag[1], p->h2);
  conv<kC2, kC3, true>(p, p->h2, p->frag[2], p->o, kNout, 4);
}

// z = u + active * N(u) at state y (u and z distinct)
void apply_correction(llg_lpc* p, const double* y, double gamma, const double* u, double* z) {
  if (p->fused) {
    const int tiles = (p->ng / kT) * cdiv(p->ny, kT);
    hipLaunchKernelGGL(k_lpc_fused, dim3(tiles), dim3(kWg), kFusedLds, p->s, y, p->d_active,
                       gamma, p->ng, p->ny, p->periodic, (const u16x8*)p->frag32[0],
                       (const u16x8*)p->frag32[1], (const u16x8*)p->frag32[2], u, z);
    return;
  }
  forward(p, u, y, gamma);
  hipMemcpyAsync(z, u, sizeof(double) * 3 * p->n, hipMemcpyDeviceToDevice, p->s);
  hipLaunchKernelGGL(k_apply_out, dim3(cdiv(p->n, 256)), dim3(256), 0, p->s, p->o, p->d_active,
                     p->n, z);
}

bool ok_launch(const char* what) {
  const hipError_t e = hipPeekAtLastError();
  if (e != hipSuccess) {
    g_lerr = std::string(what) + ": " + hipGetErrorString(e);
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

const char* LLG_LpcLastError(void) { return g_lerr.c_str(); }

llg_lpc* LLG_LpcCreate(llg_ctx* ctx, int periodic) {
  g_lerr.clear();
  if (!ctx) { g_lerr = "null context"; return nullptr; }
  int ng = 0, ny = 0;
  LLG_GetGrid(ctx, &ng, &ny);
  if (ng <= 0 || ny <= 0 || ng % 16) {
    g_lerr = "learned preconditioner needs ng a multiple of 16";
    return nullptr;
  }
  llg_lpc* p = new llg_lpc();
  std::memset((void*)p, 0, sizeof(*p));
  p->ng = ng;
  p->ny = ny;
  p->n = (long)ng * ny;
  p->periodic = periodic ? 1 : 0;
  p->s = (hipStream_t)LLG_GetStream(ctx);
  p->d_active = LLG_GetActiveDevice(ctx);
  p->fused = 1;
  if (!alloc_all(p)) {
    g_lerr = "learned preconditioner: device allocation failed";
    LLG_LpcDestroy(p);
    return nullptr;
  }
  // zero weights: the correction starts as exactly zero (plain block-Jacobi)
  make_frags(p);
  hipStreamSynchronize(p->s);
  return p;
}

void LLG_LpcDestroy(llg_lpc* p) {
  if (!p) return;
  void* bufs[] = {p->W, p->G, p->M, p->V, p->part, p->d_lpart, p->frag[0], p->frag[1],
                  p->frag[2], p->fragT[0], p->fragT[1], p->fragT[2], p->f, p->h1, p->h2,
                  p->g3, p->g2, p->g1, p->o, p->ubuf, p->frag32[0], p->frag32[1],
                  p->frag32[2]};
  for (void* b : bufs)
    if (b) hipFree(b);
  delete p;
}

int LLG_LpcNumWeights(llg_lpc* p) { return p ? p->nwtot : -1; }

int LLG_LpcSetWeights(llg_lpc* p, const float* w_host) {
  if (!p || !w_host) return -1;
  if (hipMemcpyAsync(p->W, w_host, sizeof(float) * p->nwtot, hipMemcpyHostToDevice, p->s) !=
      hipSuccess)
    return -1;
  hipMemsetAsync(p->M, 0, sizeof(float) * p->nwtot, p->s);
  hipMemsetAsync(p->V, 0, sizeof(float) * p->nwtot, p->s);
  p->step = 0;
  make_frags(p);
  hipStreamSynchronize(p->s);
  return ok_launch("LLG_LpcSetWeights") ? 0 : -1;
}

int LLG_LpcGetWeights(llg_lpc* p, float* w_host) {
  if (!p || !w_host) return -1;
  hipStreamSynchronize(p->s);
  return hipMemcpy(w_host, p->W, sizeof(float) * p->nwtot, hipMemcpyDeviceToHost) == hipSuccess
             ? 0 : -1;
}

int LLG_LpcGetGrad(llg_lpc* p, float* g_host) {
  if (!p || !g_host) return -1;
  hipStreamSynchronize(p->s);
  return hipMemcpy(g_host, p->G, sizeof(float) * p->nwtot, hipMemcpyDeviceToHost) == hipSuccess
             ? 0 : -1;
}

int LLG_LpcForward(llg_lpc* p, const double* u_dev, const double* y_dev, double gamma,
                   double scale, float* out_host) {
  if (!p || !u_dev || !y_dev) return -1;
  forward(p, u_dev, y_dev, scale * gamma);
  if (!ok_launch("LLG_LpcForward")) return -1;
  hipStreamSynchronize(p->s);
  if (out_host &&
      hipMemcpy(out_host, p->o, sizeof(float) * 4 * p->n, hipMemcpyDeviceToHost) != hipSuccess)
    return -1;
  return 0;
}

int LLG_LpcLossGrad(llg_lpc* p, const double* u_dev, const double* y_dev, const double* x_dev,
                    double gamma, double scale, double* loss) {
  if (!p || !u_dev || !y_dev || !x_dev) return -1;
  forward(p, u_dev, y_dev, scale * gamma);
  const int nb = cdiv(p->n, kWg);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(kWg), 0, p->s, p->o, u_dev, x_dev, p->d_active,
                     scale, p->n, p->g3, p->d_lpart);
  // backward: g3 -> dW3, g2 = C3^T g3 -> dW2, g1 = C2^T g2 -> dW1
  wgrad<kC2, kC3>(p, p->h2, p->g3, p->G + p->off[2]);
  conv<kC3, kC2, false>(p, p->g3, p->fragT[2], p->g2);
  wgrad<kC1, kC2>(p, p->h1, p->g2, p->G + p->off[1]);
  conv<kC2, kC1, false>(p, p->g2, p->fragT[1], p->g1);
  wgrad<kC1in, kC1>(p, p->f, p->g1, p->G + p->off[0]);
  if (!ok_launch("LLG_LpcLossGrad")) return -1;
  std::vector<double> part(nb);
  hipStreamSynchronize(p->s);
  hipMemcpy(part.data(), p->d_lpart, sizeof(double) * nb, hipMemcpyDeviceToHost);
  double s = 0.0;
  for (double v : part) s += v;
  if (loss) *loss = s / (double)p->n;
  return 0;
}

int LLG_LpcAdamStep(llg_lpc* p, float lr) {
  if (!p) return -1;
  p->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)p->step), c2 = 1.f - powf(b2, (float)p->step);
  hipLaunchKernelGGL(k_adam, dim3(cdiv(p->nwtot, 256)), dim3(256), 0, p->s, p->W, p->G, p->M,
                     p->V, p->nwtot, lr, b1, b2, 1e-8f, c1, c2);
  make_frags(p);
  return ok_launch("LLG_LpcAdamStep") ? 0 : -1;
}

int LLG_LpcTrainStep(llg_lpc* p, const double* u_dev, const double* y_dev,
                     const double* x_dev, double gamma, double scale, float lr, double* loss) {
  if (LLG_LpcLossGrad(p, u_dev, y_dev, x_dev, gamma, scale, loss) != 0) return -1;
  return LLG_LpcAdamStep(p, lr);
}

int LLG_LpcCorrect(llg_lpc* p, const double* y_dev, double gamma, double* z_dev) {
  if (!p || !y_dev || !z_dev) return -1;
  hipMemcpyAsync(p->ubuf, z_dev, sizeof(double) * 3 * p->n, hipMemcpyDeviceToDevice, p->s);
  apply_correction(p, y_dev, gamma, p->ubuf, z_dev);
  return ok_launch("LLG_LpcCorrect") ? 0 : -1;
}

int LLG_LpcApply(llg_lpc* p, const double* y_dev, double gamma, const double* u_dev,
                 double* z_dev) {
  if (!p || !y_dev || !u_dev || !z_dev || u_dev == z_dev) return -1;
  apply_correction(p, y_dev, gamma, u_dev, z_dev);
  return ok_launch("LLG_LpcApply") ? 0 : -1;
}

double* LLG_LpcInputBuffer(llg_lpc* p) { return p ? p->ubuf : nullptr; }

int LLG_LpcSetFused(llg_lpc* p, int on) {
  if (!p) return -1;
  p->fused = on ? 1 : 0;
  return 0;
}

}  // extern "C"
