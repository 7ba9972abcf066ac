// geometry.cpp -- host-side geometry and initial-condition builders.
//
// Reference builders restated (paths under src/irregular/2d_fft/):
//   square hole   2d_i4/2d_fft.cu:497-523
//   rectangle     2d_i5/2d_fft.cu:516-535 (active inside a centred rectangle)
//   ellipse       2d_i6/2d_i6.cu:462-533
//   polycrystal   2d_i3/2d_fft.cu:253-387 (glibc rand() call order kept;
//                 the O(N * holes) window loop is replaced by an exact
//                 periodic squared-distance transform giving the same
//                 integer d^2, hence the same w)
//   ICs           two-domain 2d_i4/2d_fft.cu:567-610, uniform
//                 2d_i6/2d_i6.cu:671-711, head-on stripes 2d_i1/2d_fft.cu:625-657,
//                 grain bumps 2d_i3/2d_fft.cu:916-984
// Builders without a reference counterpart (ring, tilted-z, split-x,
// random-sphere ICs) draw from splitmix64 so they are reproducible on any
// platform.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "llg/llg_geometry.h"

namespace {

double rand_unit() { return (double)rand() / (double)RAND_MAX; }

struct SplitMix {
  uint64_t s;
  double next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

// nearest periodic seed image per cell, block-pruned (see the caller)
void voronoi_blocks(int ng, int ny, int num_grains, const llg_grain* grains, int* grain_id) {
  constexpr int B = 32;
  const int nbx = (ng + B - 1) / B, nby = (ny + B - 1) / B;
  const int nblk = nbx * nby;
  const int nimg = 9 * num_grains;
  auto work = [&](int b0, int b1) {
    std::vector<int> cand;
    cand.reserve(nimg);
    for (int b = b0; b < b1; b++) {
      const int bi = b % nbx, bj = b / nbx;
      const int i0 = bi * B, j0 = bj * B;
      const int i1 = std::min(ng, i0 + B), j1 = std::min(ny, j0 + B);
      // cell centres of the block span [x0, x1] x [y0, y1]
      const double x0 = i0 + 0.5, x1 = (i1 - 1) + 0.5, y0 = j0 + 0.5, y1 = (j1 - 1) + 0.5;
      double U2 = 1.0e300;
      for (int g = 0; g < num_grains; g++)
        for (int dyp = -1; dyp <= 1; dyp++)
          for (int dxp = -1; dxp <= 1; dxp++) {
            const double sx = grains[g].sx + (double)dxp * (double)ng;
            const double sy = grains[g].sy + (double)dyp * (double)ny;
            const double fx = std::max(std::fabs(sx - x0), std::fabs(sx - x1));
            const double fy = std::max(std::fabs(sy - y0), std::fabs(sy - y1));
            U2 = std::min(U2, fx * fx + fy * fy);
          }
      const double lim = U2 * (1.0 + 1e-9) + 1e-6;
      cand.clear();
      for (int g = 0; g < num_grains; g++)
        for (int dyp = -1; dyp <= 1; dyp++)
          for (int dxp = -1; dxp <= 1; dxp++) {
            const double sx = grains[g].sx + (double)dxp * (double)ng;
            const double sy = grains[g].sy + (double)dyp * (double)ny;
            const double nx_ = sx < x0 ? x0 - sx : (sx > x1 ? sx - x1 : 0.0);
            const double ny_ = sy < y0 ? y0 - sy : (sy > y1 ? sy - y1 : 0.0);
            if (nx_ * nx_ + ny_ * ny_ <= lim) cand.push_back((g * 3 + (dyp + 1)) * 3 + (dxp + 1));
          }
      for (int j = j0; j < j1; j++)
        for (int i = i0; i < i1; i++) {
          const double x = (double)i + 0.5, y = (double)j + 0.5;
          int best = 0;
          double best_d2 = 1.0e30;
          for (const int c : cand) {
            const int g = c / 9, dyp = (c / 3) % 3 - 1, dxp = c % 3 - 1;
            const double sx = grains[g].sx + (double)dxp * (double)ng;
            const double sy = grains[g].sy + (double)dyp * (double)ny;
            const double dx = x - sx, dy = y - sy;
            const double d2 = dx * dx + dy * dy;
            if (d2 < best_d2) { best_d2 = d2; best = g; }
          }
          grain_id[j * ng + i] = best;
        }
    }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const int nth = (int)std::max(1u, std::min(16u, hw ? hw : 1u));
  if (nth == 1 || nblk < 64) {
    work(0, nblk);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nth; t++)
    th.emplace_back(work, (int)((long)nblk * t / nth), (int)((long)nblk * (t + 1) / nth));
  for (auto& t : th) t.join();
}

// 1D squared distance transform, lower envelope of parabolas
// (Felzenszwalb & Huttenlocher); entries >= 1e299 are "no feature".
void dt1d(const std::vector<double>& f, std::vector<double>& d) {
  const int n = (int)f.size();
  d.assign(n, 1e300);
  std::vector<int> v(n);
  std::vector<double> z(n + 1);
  int k = -1;
  for (int q = 0; q < n; q++) {
    if (!(f[q] < 1e299)) continue;
    if (k < 0) {
      k = 0;
      v[0] = q;
      z[0] = -1e300;
      z[1] = 1e300;
      continue;
    }
    double s = ((f[q] + (double)q * q) - (f[v[k]] + (double)v[k] * v[k])) / (2.0 * q - 2.0 * v[k]);
    while (s <= z[k]) {
      k--;
      s = ((f[q] + (double)q * q) - (f[v[k]] + (double)v[k] * v[k])) / (2.0 * q - 2.0 * v[k]);
    }
    k++;
    v[k] = q;
    z[k] = s;
    z[k + 1] = 1e300;
  }
  if (k < 0) return;
  int j = 0;
  for (int q = 0; q < n; q++) {
    while (z[j + 1] < q) j++;
    const double dq = (double)q - v[j];
    d[q] = dq * dq + f[v[j]];
  }
}

}  // namespace

extern "C" {

int LLG_GeomSquareHole(int ng, int ny, double cxf, double cyf, double halfside_frac,
                       unsigned char* active) {
  if (ng <= 0 || ny <= 0 || !active) return -1;
  const double cx = cxf * (double)(ng - 1);
  const double cy = cyf * (double)(ny - 1);
  const double halfside = halfside_frac * (double)ny;
  int na = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < ng; ++i) {
      const double ddx = std::fabs((double)i - cx);
      const double ddy = std::fabs((double)j - cy);
      const int hole = (ddx <= halfside && ddy <= halfside);
      active[j * ng + i] = hole ? 0 : 1;
      na += !hole;
    }
  return na;
}

int LLG_GeomRectangle(int ng, int ny, double wf, double hf, unsigned char* active) {
  if (ng <= 0 || ny <= 0 || !active) return -1;
  const double cx = 0.5 * (double)ng, cy = 0.5 * (double)ny;
  const double hw = 0.5 * wf * (double)ng, hh = 0.5 * hf * (double)ny;
  int na = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < ng; ++i) {
      const double dx = std::fabs((double)i + 0.5 - cx);
      const double dy = std::fabs((double)j + 0.5 - cy);
      const int a = (dx <= hw && dy <= hh);
      active[j * ng + i] = (unsigned char)a;
      na += a;
    }
  return na;
}

int LLG_GeomRing(int ng, int ny, double wf, double hf, double inner_wf, double inner_hf,
                 unsigned char* active) {
  if (ng <= 0 || ny <= 0 || !active || inner_wf >= 1.0 || inner_hf >= 1.0) return -1;
  const double cx = 0.5 * (double)ng, cy = 0.5 * (double)ny;
  const double hw = 0.5 * wf * (double)ng, hh = 0.5 * hf * (double)ny;
  const double iw = hw * inner_wf, ih = hh * inner_hf;
  int na = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < ng; ++i) {
      const double dx = std::fabs((double)i + 0.5 - cx);
      const double dy = std::fabs((double)j + 0.5 - cy);
      const int outer = (dx <= hw && dy <= hh);
      const int inner = (dx < iw && dy < ih);
      const int a = outer && !inner;
      active[j * ng + i] = (unsigned char)a;
      na += a;
    }
  return na;
}

int LLG_GeomEllipse(int ng, int ny, double rxf, double ryf, unsigned char* active) {
  if (ng <= 0 || ny <= 0 || !active) return -1;
  const double cx = 0.5 * (double)ng, cy = 0.5 * (double)ny;
  const double rx = rxf * (double)ng, ry = ryf * (double)ny;
  if (rx <= 0.0 || ry <= 0.0) return -1;
  const double irx2 = 1.0 / (rx * rx), iry2 = 1.0 / (ry * ry);
  int na = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < ng; ++i) {
      const double dx = (double)i + 0.5 - cx;
      const double dy = (double)j + 0.5 - cy;
      const double q = dx * dx * irx2 + dy * dy * iry2;
      const int a = q <= 1.0;
      active[j * ng + i] = (unsigned char)a;
      na += a;
    }
  return na;
}

int LLG_GeomPolycrystal(int ng, int ny, int num_grains, double dead_frac, unsigned seed,
                        double z_bias, double eps_cells, llg_grain* grains, int* grain_id,
                        unsigned char* active, double* w, double* msk, double* nsk) {
  if (ng <= 0 || ny <= 0 || num_grains < 1 || !grains || !grain_id) return -1;
  const int N = ng * ny;
  srand(seed);
  int gx = (int)std::ceil(std::sqrt((double)num_grains * (double)ng / (double)ny));
  if (gx < 1) gx = 1;
  const int gy = (num_grains + gx - 1) / gx;
  const double dxs = (double)ng / (double)gx, dys = (double)ny / (double)gy;
  int idx = 0;
  for (int j = 0; j < gy && idx < num_grains; j++)
    for (int i = 0; i < gx && idx < num_grains; i++) {
      grains[idx].sx = ((double)i + 0.15 + 0.7 * rand_unit()) * dxs;
      grains[idx].sy = ((double)j + 0.15 + 0.7 * rand_unit()) * dys;
      while (grains[idx].sx < 0.0) grains[idx].sx += (double)ng;
      while (grains[idx].sx >= (double)ng) grains[idx].sx -= (double)ng;
      while (grains[idx].sy < 0.0) grains[idx].sy += (double)ny;
      while (grains[idx].sy >= (double)ny) grains[idx].sy -= (double)ny;
      idx++;
    }
  while (idx < num_grains) {
    grains[idx].sx = rand_unit() * (double)ng;
    grains[idx].sy = rand_unit() * (double)ny;
    idx++;
  }
  for (int g = 0; g < num_grains; g++) {
    const double u = rand_unit();
    const double sign = (rand_unit() < 0.5) ? -1.0 : 1.0;
    const double z = sign * std::pow(u, 1.0 / z_bias);
    const double phi = 2.0 * M_PI * rand_unit();
    const double rxy = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    grains[g].ax = rxy * std::cos(phi);
    grains[g].ay = rxy * std::sin(phi);
    grains[g].az = z;
    const double psi = 2.0 * M_PI * rand_unit();
    grains[g].nx = std::cos(psi);
    grains[g].ny = std::sin(psi);
    grains[g].nz = 0.0;
    grains[g].dead = (rand_unit() < dead_frac) ? 1 : 0;
  }
  // Voronoi with the 3x3 periodic images (2d_i3/2d_fft.cu:300-315): the
  // nearest image by the reference's d2 arithmetic, ties to the first in its
  // (grain, dyp, dxp) scan order.  Cells are taken in 32x32 blocks; a block
  // scans, in that same order, only the images that can be nearest to one
  // of its cells: U = min over images of the largest distance from the
  // image to the block rectangle bounds every cell's nearest distance, and
  // an image whose distance to the rectangle exceeds U (with a margin far
  // above rounding) is strictly farther than that cell's nearest, so it can
  // neither win nor tie.  Same ids as the full scan in O(N) work.
  voronoi_blocks(ng, ny, num_grains, grains, grain_id);
  int n_active = 0;
  for (int k = 0; k < N; k++) {
    const int dead = grains[grain_id[k]].dead;
    if (active) active[k] = dead ? 0 : 1;
    n_active += !dead;
  }
  if (w) {
    if (n_active == N) {
      for (int k = 0; k < N; k++) w[k] = 1.0;
    } else {
      // exact periodic squared distance to the nearest dead cell:
      // rows (tripled for wrap) then columns (tripled for wrap)
      std::vector<double> g2((size_t)N);
      std::vector<double> f, d;
      for (int j = 0; j < ny; j++) {
        f.assign(3 * ng, 1e300);
        for (int r = 0; r < 3; r++)
          for (int i = 0; i < ng; i++)
            if (grains[grain_id[j * ng + i]].dead) f[r * ng + i] = 0.0;
        dt1d(f, d);
        for (int i = 0; i < ng; i++) g2[(size_t)j * ng + i] = d[ng + i];
      }
      for (int i = 0; i < ng; i++) {
        f.assign(3 * ny, 1e300);
        for (int r = 0; r < 3; r++)
          for (int j = 0; j < ny; j++) f[r * ny + j] = g2[(size_t)j * ng + i];
        dt1d(f, d);
        for (int j = 0; j < ny; j++) {
          const int cell = j * ng + i;
          if (grains[grain_id[cell]].dead) { w[cell] = 0.0; continue; }
          const double d2_min = d[ny + j];
          const double dd = std::sqrt(d2_min) - 0.5;
          double ww = 0.5 * (1.0 + std::tanh(dd / eps_cells));
          if (ww < 1.0e-3) ww = 0.0;
          if (ww > 1.0 - 1.0e-3) ww = 1.0;
          w[cell] = ww;
        }
      }
    }
  }
  for (int k = 0; k < N; k++) {
    const llg_grain& G = grains[grain_id[k]];
    if (msk) { msk[k] = G.ax; msk[N + k] = G.ay; msk[2 * N + k] = G.az; }
    if (nsk) { nsk[k] = G.nx; nsk[N + k] = G.ny; nsk[2 * N + k] = G.nz; }
  }
  return n_active;
}

int LLG_ICUniform(int ng, int ny, const unsigned char* active, double mx, double my, double mz,
                  double* y) {
  if (!y) return -1;
  const int N = ng * ny;
  const double n0 = std::sqrt(mx * mx + my * my + mz * mz);
  if (n0 > 1.0e-30) { mx /= n0; my /= n0; mz /= n0; }
  for (int k = 0; k < N; k++) {
    const int a = active ? active[k] != 0 : 1;
    y[k] = a ? mx : 0.0;
    y[N + k] = a ? my : 0.0;
    y[2 * N + k] = a ? mz : 0.0;
  }
  return 0;
}

int LLG_ICTwoDomain(int ng, int ny, const unsigned char* active, int j_split, double tail_my,
                    double* y) {
  if (!y) return -1;
  const int N = ng * ny;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < ng; i++) {
      const int cell = j * ng + i;
      const int a = active ? active[cell] != 0 : 1;
      if (a) {
        double mx0, my0, mz0 = 0.0;
        if (j >= j_split) { mx0 = -1.0; my0 = 0.0; }
        else { mx0 = +1.0; my0 = (i < ng / 2) ? -tail_my : +tail_my; }
        const double n0 = std::sqrt(mx0 * mx0 + my0 * my0 + mz0 * mz0);
        if (n0 > 1e-30) { mx0 /= n0; my0 /= n0; mz0 /= n0; }
        y[cell] = mx0; y[N + cell] = my0; y[2 * N + cell] = mz0;
      } else {
        y[cell] = y[N + cell] = y[2 * N + cell] = 0.0;
      }
    }
  return 0;
}

int LLG_ICStripes(int ng, int ny, double left_frac, double right_frac, double eps,
                  unsigned seed, double* y) {
  if (!y) return -1;
  const int N = ng * ny;
  const int q1 = (int)(left_frac * ng), q3 = (int)(right_frac * ng);
  srand(seed);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < ng; i++) {
      const int cell = j * ng + i;
      double mx0 = (i >= q1 && i < q3) ? 1.0 : -1.0;
      double my0 = eps * (2.0 * (double)rand() / (double)RAND_MAX - 1.0);
      double mz0 = eps * (2.0 * (double)rand() / (double)RAND_MAX - 1.0);
      const double nrm = std::sqrt(mx0 * mx0 + my0 * my0 + mz0 * mz0);
      mx0 /= nrm; my0 /= nrm; mz0 /= nrm;
      y[cell] = mx0; y[N + cell] = my0; y[2 * N + cell] = mz0;
    }
  return 0;
}

int LLG_ICGrainBumps(int ng, int ny, const llg_grain* grains, const int* grain_id,
                     const double* w, double core_mz, unsigned seed, double* y) {
  if (!y || !grains || !grain_id || !w) return -1;
  const int N = ng * ny;
  srand(seed);
  const double eps_n = 0.05;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < ng; i++) {
      const int k = j * ng + i;
      const llg_grain& G = grains[grain_id[k]];
      if (w[k] <= 1.0e-3) { y[k] = y[N + k] = y[2 * N + k] = 0.0; continue; }
      double ddx = (double)i + 0.5 - G.sx, ddy = (double)j + 0.5 - G.sy;
      if (ddx > ng / 2.0) ddx -= ng;
      if (ddx < -ng / 2.0) ddx += ng;
      if (ddy > ny / 2.0) ddy -= ny;
      if (ddy < -ny / 2.0) ddy += ny;
      const double r = std::sqrt(ddx * ddx + ddy * ddy);
      double s = r / 14.0;
      if (s > 1.0) s = 1.0;
      double sign_z = (G.az >= 0.0) ? 1.0 : -1.0;
      if (std::fabs(G.az) < 0.05) sign_z = 0.0;
      double mz0 = sign_z * core_mz * (1.0 - s * s);
      const double mperp = std::sqrt(std::fmax(0.0, 1.0 - mz0 * mz0));
      double tx = -ddy, ty = ddx;
      const double tn = std::sqrt(tx * tx + ty * ty);
      if (tn > 1.0e-12) { tx /= tn; ty /= tn; } else { tx = 1.0; ty = 0.0; }
      const double bn = std::sqrt(G.ax * G.ax + G.ay * G.ay);
      double ex, ey;
      if (bn > 1.0e-6) { ex = G.ax / bn; ey = G.ay / bn; } else { ex = 1.0; ey = 0.0; }
      double pxx = (1.0 - s) * tx + s * ex, pyy = (1.0 - s) * ty + s * ey;
      const double pn = std::sqrt(pxx * pxx + pyy * pyy);
      if (pn > 1.0e-12) { pxx /= pn; pyy /= pn; }
      double mx0 = mperp * pxx + eps_n * (2.0 * rand_unit() - 1.0);
      double my0 = mperp * pyy + eps_n * (2.0 * rand_unit() - 1.0);
      mz0 += eps_n * (2.0 * rand_unit() - 1.0);
      const double n = std::sqrt(mx0 * mx0 + my0 * my0 + mz0 * mz0);
      if (n > 1.0e-12) { mx0 /= n; my0 /= n; mz0 /= n; }
      y[k] = mx0; y[N + k] = my0; y[2 * N + k] = mz0;
    }
  return 0;
}

int LLG_ICTiltedZ(int ng, int ny, const unsigned char* active, double max_angle, uint64_t seed,
                  double* y) {
  if (!y) return -1;
  const int N = ng * ny;
  SplitMix r{seed};
  for (int k = 0; k < N; k++) {
    const double th = max_angle * r.next();
    const double ph = 2.0 * M_PI * r.next();
    const int a = active ? active[k] != 0 : 1;
    double mx = std::sin(th) * std::cos(ph), my = std::sin(th) * std::sin(ph), mz = std::cos(th);
    const double n = std::sqrt(mx * mx + my * my + mz * mz);
    mx /= n; my /= n; mz /= n;
    y[k] = a ? mx : 0.0; y[N + k] = a ? my : 0.0; y[2 * N + k] = a ? mz : 0.0;
  }
  return 0;
}

int LLG_ICSplitX(int ng, int ny, const unsigned char* active, double max_angle, uint64_t seed,
                 double* y) {
  if (!y) return -1;
  const int N = ng * ny;
  SplitMix r{seed};
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < ng; i++) {
      const int k = j * ng + i;
      const double th = max_angle * r.next();
      const double ph = 2.0 * M_PI * r.next();
      const int a = active ? active[k] != 0 : 1;
      const double sx = (i < ng / 2) ? 1.0 : -1.0;
      double mx = sx * std::cos(th), my = std::sin(th) * std::cos(ph), mz = std::sin(th) * std::sin(ph);
      const double n = std::sqrt(mx * mx + my * my + mz * mz);
      mx /= n; my /= n; mz /= n;
      y[k] = a ? mx : 0.0; y[N + k] = a ? my : 0.0; y[2 * N + k] = a ? mz : 0.0;
    }
  return 0;
}

int LLG_ICRandomSphere(int ng, int ny, const unsigned char* active, uint64_t seed, double* y) {
  if (!y) return -1;
  const int N = ng * ny;
  SplitMix r{seed};
  for (int k = 0; k < N; k++) {
    const double z = 2.0 * r.next() - 1.0;
    const double ph = 2.0 * M_PI * r.next();
    const double rho = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    const int a = active ? active[k] != 0 : 1;
    y[k] = a ? rho * std::cos(ph) : 0.0;
    y[N + k] = a ? rho * std::sin(ph) : 0.0;
    y[2 * N + k] = a ? z : 0.0;
  }
  return 0;
}

}  // extern "C"
