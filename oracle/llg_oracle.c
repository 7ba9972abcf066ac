/*
 * llg_oracle.c -- CPU FP64 restatement of the reference's LLG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path never links it.
 *
 * Functions and the reference lines they restate (paths relative to the
 * reference root; i4 = src/irregular/2d_fft/2d_i4, i6 = .../2d_i6,
 * i3 = .../2d_i3, i1 = .../2d_i1, heter = src/2d/2d_heter):
 *   orc_normalize      exact i1/2d_fft.cu:253-275, regularised i2/2d_fft.cu:262-283
 *   orc_rhs            i4/2d_fft.cu:262-319 (ymsk), i6/2d_i6.cu:250-343 (compact),
 *                      i1/2d_fft.cu:296-360 (linear aniso + DMI),
 *                      i3/2d_fft.cu:451-627 (windowed polycrystal); wrapper order
 *                      i4/2d_fft.cu:328-366, i6/2d_i6.cu:352-396, i3/2d_fft.cu:656-726
 *   orc_jtv            i4/jtv.cu:47-120, i6/jtv.cu:47-130, i1/jtv.cu:43-96
 *   orc_psetup         i4/precond.cu:120-252,324-358 (always rebuild, jcur=TRUE),
 *                      i6/precond.cu:118-249, i1/precond.cu:108-241,
 *                      two-phase jok protocol heter/precond.cu:154-192,265-314
 *   orc_psolve         i4/precond.cu:255-286,361-385, i6/precond.cu:98-110,252-285
 *   orc_ctt/orc_calt   i4/demag_fft.cu:27-91 (81-point face-averaged Newell tensor)
 *   orc_demag_direct   i4/demag_direct.cu:86-131,231-281 (O(N^2) periodic sum);
 *                      padded: same sum without wrap (no reference counterpart)
 *   orc_demag_fft      i4/demag_fft.cu:184-363 (9 complex tensor spectra,
 *                      forward FFT of m, 3x3 product, inverse, shift, scale)
 *   orc_set_happ       field-sweep stage (the applied field changed between
 *                      integrator restarts; no reference counterpart)
 *   orc_nv_*           SUNDIALS serial N_Vector semantics; the multi-dot and
 *                      linear combination of i4/deferred_nvector.cu:95-180 and
 *                      heter/deferred_nvector.cu:135-151
 * Parity pins for every function are listed in tests/ (see the test files'
 * headers); anything without one is marked "parity unpinned" there.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>

#include "llg/llg_plugin.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================== Newell tensor */
void orc_ctt(double b, double a, double sx, double sy, double dm[10]) {
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
  dm[3] = log((dnnn - yn) / (dnpn - yp)) - log((dpnn - yn) / (dppn - yp)) -
          log((dnnp - yn) / (dnpp - yp)) + log((dpnp - yn) / (dppp - yp));
  dm[4] = log((dnnn - zn) / (dnnp - zp)) - log((dnpn - zn) / (dnpp - zp)) -
          log((dpnn - zn) / (dpnp - zp)) + log((dppn - zn) / (dppp - zp));
  dm[5] = atan(zn * xn / (yn * dnnn)) - atan(zp * xn / (yn * dnnp)) -
          atan(zn * xp / (yn * dpnn)) + atan(zp * xp / (yn * dpnp)) -
          atan(zn * xn / (yp * dnpn)) + atan(zp * xn / (yp * dnpp)) +
          atan(zn * xp / (yp * dppn)) - atan(zp * xp / (yp * dppp));
  dm[6] = log((dnnn - xn) / (dpnn - xp)) - log((dnpn - xn) / (dppn - xp)) -
          log((dnnp - xn) / (dpnp - xp)) + log((dnpp - xn) / (dppp - xp));
  dm[7] = log((dnnn - yn) / (dnpn - yp)) - log((dnnp - yn) / (dnpp - yp)) -
          log((dpnn - yn) / (dppn - yp)) + log((dpnp - yn) / (dppp - yp));
  dm[8] = log((dnnn - xn) / (dpnn - xp)) - log((dnnp - xn) / (dpnp - xp)) -
          log((dnpn - xn) / (dppn - xp)) + log((dnpp - xn) / (dppp - xp));
  dm[9] = atan(xn * yn / (zn * dnnn)) - atan(xp * yn / (zn * dpnn)) -
          atan(xn * yp / (zn * dnpn)) + atan(xp * yp / (zn * dppn)) -
          atan(xn * yn / (zp * dnnp)) + atan(xp * yn / (zp * dpnp)) +
          atan(xn * yp / (zp * dnpp)) - atan(xp * yp / (zp * dppp));
}

/* tensor at one displacement (sx, sy) in cells: 9-component average */
void orc_tensor_at(double thick, int dx, int dy, double out9[9]) {
  const double a = 0.49999, b = 0.5 * thick;
  double dm[10];
  for (int c = 0; c < 9; c++) out9[c] = 0.0;
  for (int jy = -4; jy <= 4; jy++) {
    const double sy = (double)dy + 0.1 * (double)jy;
    for (int ix = -4; ix <= 4; ix++) {
      const double sx = (double)dx + 0.1 * (double)ix;
      orc_ctt(b, a, sx, sy, dm);
      for (int c = 0; c < 9; c++) out9[c] += dm[c + 1];
    }
  }
  for (int c = 0; c < 9; c++) out9[c] /= 81.;
}

/* calt: planes9[c*mdx*mdy + j*mdx + i] = N_c(i - mdx/2, j - mdy/2) */
void orc_calt(double thick, int mdx, int mdy, double* planes9) {
  const int mdx2 = mdx / 2, mdy2 = mdy / 2;
  const size_t plane = (size_t)mdx * mdy;
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 0; j < mdy; j++) {
    double t9[9];
    for (int i = 0; i < mdx; i++) {
      orc_tensor_at(thick, i - mdx2, j - mdy2, t9);
      for (int c = 0; c < 9; c++) planes9[c * plane + (size_t)j * mdx + i] = t9[c];
    }
  }
}

/* ========================================================== demag */
/* O(N^2) direct sum.  periodic: displacement wrapped into [-n/2, n-n/2)
 * exactly as the reference's ((d + n/2) % n + n) % n indexing; padded:
 * plain displacement (open boundary).  w may be NULL (no window). */
void orc_demag_direct(int nx, int ny, double thick, double strength, int padded,
                      const double* y, const double* w, double* h) {
  const int ncell = nx * ny;
  const int mdx = padded ? 2 * nx : nx, mdy = padded ? 2 * ny : ny;
  const size_t plane = (size_t)mdx * mdy;
  double* T = (double*)malloc(9 * plane * sizeof(double));
  orc_calt(thick, mdx, mdy, T);
  const int mdx2 = mdx / 2, mdy2 = mdy / 2;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < ncell; t++) {
    const int ti = t % nx, tj = t / nx;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int sj = 0; sj < ny; sj++) {
      for (int si = 0; si < nx; si++) {
        const int src = sj * nx + si;
        const double ww = w ? w[src] : 1.0;
        const double vx = ww * y[src], vy = ww * y[ncell + src], vz = ww * y[2 * ncell + src];
        const int di = ti - si, dj = tj - sj;
        int tx, ty;
        if (padded) {
          tx = di + mdx2;
          ty = dj + mdy2;
        } else {
          tx = ((di + mdx2) % mdx + mdx) % mdx;
          ty = ((dj + mdy2) % mdy + mdy) % mdy;
        }
        const size_t k = (size_t)ty * mdx + tx;
        s0 += T[0 * plane + k] * vx + T[1 * plane + k] * vy + T[2 * plane + k] * vz;
        s1 += T[3 * plane + k] * vx + T[4 * plane + k] * vy + T[5 * plane + k] * vz;
        s2 += T[6 * plane + k] * vx + T[7 * plane + k] * vy + T[8 * plane + k] * vz;
      }
    }
    h[t] = strength * s0;
    h[ncell + t] = strength * s1;
    h[2 * ncell + t] = strength * s2;
  }
  free(T);
}

/* --- mixed-radix complex FFT (2,3,4,5 + generic odd), recursive
 * decimation in time, exp(sign*2*pi*i*j*k/n) kernel, unnormalised. */
typedef double complex cplx;

/* twiddle table exp(sign 2 pi i k / N), k in [0, N), N the top-level length */
static void dft_small(const cplx* in, size_t istride, cplx* out, int n, const cplx* tw,
                      int twstride, int N) {
  for (int k = 0; k < n; k++) {
    cplx s = 0;
    for (int j = 0; j < n; j++) s += in[j * istride] * tw[(size_t)((long)j * k % n) * twstride % N];
    out[k] = s;
  }
}

/* recursive mixed-radix decimation in time; tw holds exp(sign 2 pi i k/N)
 * and a transform of length n uses every (N/n)-th entry */
static void fft_rec(const cplx* in, size_t istride, cplx* out, int n, const cplx* tw, int N,
                    cplx* scratch) {
  if (n == 1) { out[0] = in[0]; return; }
  int r = 0;
  if (n % 4 == 0) r = 4;
  else if (n % 2 == 0) r = 2;
  else if (n % 3 == 0) r = 3;
  else if (n % 5 == 0) r = 5;
  const int twstride = N / n;
  if (r == 0) { dft_small(in, istride, out, n, tw, twstride, N); return; }
  const int m = n / r;
  for (int q = 0; q < r; q++) fft_rec(in + q * istride, istride * r, out + q * m, m, tw, N, scratch);
  for (int k = 0; k < m; k++) {
    cplx t[5];
    for (int q = 0; q < r; q++) t[q] = out[q * m + k] * tw[(size_t)q * k * twstride];
    for (int s2 = 0; s2 < r; s2++) {
      cplx acc = 0;
      for (int q = 0; q < r; q++) acc += t[q] * tw[(size_t)((q * s2) % r) * (N / r)];
      scratch[s2 * m + k] = acc;
    }
  }
  memcpy(out, scratch, (size_t)n * sizeof(cplx));
}

void orc_fft1d(cplx* data, int n, int sign) {
  cplx* out = (cplx*)malloc((size_t)n * sizeof(cplx));
  cplx* in = (cplx*)malloc((size_t)n * sizeof(cplx));
  cplx* scratch = (cplx*)malloc((size_t)n * sizeof(cplx));
  cplx* tw = (cplx*)malloc((size_t)n * sizeof(cplx));
  for (int k = 0; k < n; k++) {
    const long double ang = sign * 2.0L * 3.14159265358979323846264338327950288L * k / n;
    tw[k] = (double)cosl(ang) + I * (double)sinl(ang);
  }
  memcpy(in, data, (size_t)n * sizeof(cplx));
  fft_rec(in, 1, out, n, tw, n, scratch);
  memcpy(data, out, (size_t)n * sizeof(cplx));
  free(out); free(in); free(scratch); free(tw);
}

static cplx* make_tw(int n, int sign) {
  cplx* tw = (cplx*)malloc((size_t)n * sizeof(cplx));
  for (int k = 0; k < n; k++) {
    const long double ang = sign * 2.0L * 3.14159265358979323846264338327950288L * k / n;
    tw[k] = (double)cosl(ang) + I * (double)sinl(ang);
  }
  return tw;
}

/* 2D FFT of an ny x nx row-major complex array (rows then columns) */
void orc_fft2d(cplx* a, int nx, int ny, int sign) {
  cplx* twx = make_tw(nx, sign);
  cplx* twy = make_tw(ny, sign);
#pragma omp parallel
  {
    const int nmax = nx > ny ? nx : ny;
    cplx* in = (cplx*)malloc((size_t)nmax * sizeof(cplx));
    cplx* out = (cplx*)malloc((size_t)nmax * sizeof(cplx));
    cplx* scratch = (cplx*)malloc((size_t)nmax * sizeof(cplx));
#pragma omp for schedule(static)
    for (int j = 0; j < ny; j++) {
      memcpy(in, a + (size_t)j * nx, (size_t)nx * sizeof(cplx));
      fft_rec(in, 1, out, nx, twx, nx, scratch);
      memcpy(a + (size_t)j * nx, out, (size_t)nx * sizeof(cplx));
    }
#pragma omp for schedule(static)
    for (int i = 0; i < nx; i++) {
      for (int j = 0; j < ny; j++) in[j] = a[(size_t)j * nx + i];
      fft_rec(in, 1, out, ny, twy, ny, scratch);
      for (int j = 0; j < ny; j++) a[(size_t)j * nx + i] = out[j];
    }
    free(in); free(out); free(scratch);
  }
  free(twx); free(twy);
}

typedef struct orc_demag {
  int nx, ny, px, py, padded;
  double strength;
  double n0[3];
  cplx* spec9;   /* 9 full complex spectra, py*px each */
  cplx* work3;   /* 3 planes */
} orc_demag;

orc_demag* orc_demag_create(int nx, int ny, double thick, double strength, int padded) {
  orc_demag* d = (orc_demag*)calloc(1, sizeof(orc_demag));
  d->nx = nx; d->ny = ny; d->padded = padded; d->strength = strength;
  d->px = padded ? 2 * nx : nx;
  d->py = padded ? 2 * ny : ny;
  const size_t plane = (size_t)d->px * d->py;
  double* T = (double*)malloc(9 * plane * sizeof(double));
  orc_calt(thick, d->px, d->py, T);
  const size_t i0 = (size_t)(d->py / 2) * d->px + d->px / 2;
  d->n0[0] = T[0 * plane + i0] * strength;
  d->n0[1] = T[4 * plane + i0] * strength;
  d->n0[2] = T[8 * plane + i0] * strength;
  d->spec9 = (cplx*)malloc(9 * plane * sizeof(cplx));
  d->work3 = (cplx*)malloc(3 * plane * sizeof(cplx));
  for (int c = 0; c < 9; c++) {
    cplx* s = d->spec9 + c * plane;
    for (size_t k = 0; k < plane; k++) s[k] = T[c * plane + k];
    if (padded) {
      /* displacement -n (index 0 of the centred array) never occurs in a
       * linear convolution of n samples; clear it so the padded kernel is
       * exactly the open-boundary operator. */
      for (int i = 0; i < d->px; i++) s[i] = 0;
      for (int j = 0; j < d->py; j++) s[(size_t)j * d->px] = 0;
    }
    orc_fft2d(s, d->px, d->py, -1);
  }
  free(T);
  return d;
}

void orc_demag_destroy(orc_demag* d) {
  if (!d) return;
  free(d->spec9); free(d->work3); free(d);
}

void orc_demag_self(orc_demag* d, double* n3) {
  if (!d) { n3[0] = n3[1] = n3[2] = 0; return; }
  n3[0] = d->n0[0]; n3[1] = d->n0[1]; n3[2] = d->n0[2];
}

/* gather (x w) -> forward FFT -> 3x3 spectral product -> inverse -> shift
 * by (px/2, py/2) -> scale strength/(px*py) (i4/demag_fft.cu:325-363) */
void orc_demag_apply(orc_demag* d, const double* y, const double* w, double* h) {
  const int nx = d->nx, ny = d->ny, px = d->px, py = d->py, ncell = nx * ny;
  const size_t plane = (size_t)px * py;
  cplx* M = d->work3;
  memset(M, 0, 3 * plane * sizeof(cplx));
  for (int c = 0; c < 3; c++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int cell = j * nx + i;
        const double ww = w ? w[cell] : 1.0;
        M[c * plane + (size_t)j * px + i] = ww * y[c * ncell + cell];
      }
  for (int c = 0; c < 3; c++) orc_fft2d(M + c * plane, px, py, -1);
  cplx* H = (cplx*)malloc(3 * plane * sizeof(cplx));
  const cplx* F = d->spec9;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < plane; k++) {
    const cplx ma = M[k], mb = M[plane + k], mc = M[2 * plane + k];
    H[k] = F[0 * plane + k] * ma + F[1 * plane + k] * mb + F[2 * plane + k] * mc;
    H[plane + k] = F[3 * plane + k] * ma + F[4 * plane + k] * mb + F[5 * plane + k] * mc;
    H[2 * plane + k] = F[6 * plane + k] * ma + F[7 * plane + k] * mb + F[8 * plane + k] * mc;
  }
  for (int c = 0; c < 3; c++) orc_fft2d(H + c * plane, px, py, +1);
  const double scale = d->strength / (double)plane;
  const int px2 = px / 2, py2 = py / 2;
  for (int c = 0; c < 3; c++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int si = (i + px2) % px, sj = (j + py2) % py;
        h[c * ncell + j * nx + i] = creal(H[c * plane + (size_t)sj * px + si]) * scale;
      }
  free(H);
}

/* =========================================================== context */
typedef struct orc_ctx {
  llg_config cfg;
  int ng, ny, ncell, neq;
  unsigned char* active;
  double* ymsk3;
  int *active_ids, *inactive_ids;
  int n_active, n_inactive;
  double* w;
  double *cmsk, *cnsk;
  double* hd;     /* field of the last rhs call */
  double* J;      /* 9*ncell (two-phase jok protocol) */
  double* P;      /* 9*ncell SoA inverse blocks */
  orc_demag* demag;
  double n0[3];
  long nf, njv, npset, npsol;
} orc_ctx;

orc_ctx* orc_create(const llg_config* cfg, const unsigned char* active, const double* w,
                    const double* cmsk, const double* cnsk) {
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->cfg = *cfg;
  c->ng = cfg->ng; c->ny = cfg->ny;
  c->ncell = cfg->ng * cfg->ny;
  c->neq = 3 * c->ncell;
  const int N = c->ncell;
  c->active = (unsigned char*)malloc(N);
  for (int i = 0; i < N; i++) c->active[i] = active ? (active[i] != 0) : 1;
  /* ymsk: FP64 1/0 replicated in three planes (i4/2d_fft.cu:497-518) */
  c->ymsk3 = (double*)malloc(3 * N * sizeof(double));
  for (int i = 0; i < N; i++) {
    const double m = c->active[i] ? 1.0 : 0.0;
    c->ymsk3[i] = c->ymsk3[N + i] = c->ymsk3[2 * N + i] = m;
  }
  /* ascending id lists (i6/2d_i6.cu:514-518) */
  c->active_ids = (int*)malloc((N + 1) * sizeof(int));
  c->inactive_ids = (int*)malloc((N + 1) * sizeof(int));
  for (int i = 0; i < N; i++) {
    if (c->active[i]) c->active_ids[c->n_active++] = i;
    else c->inactive_ids[c->n_inactive++] = i;
  }
  if (w) { c->w = (double*)malloc(N * sizeof(double)); memcpy(c->w, w, N * sizeof(double)); }
  if (cmsk) { c->cmsk = (double*)malloc(3 * N * sizeof(double)); memcpy(c->cmsk, cmsk, 3 * N * sizeof(double)); }
  if (cnsk) { c->cnsk = (double*)malloc(3 * N * sizeof(double)); memcpy(c->cnsk, cnsk, 3 * N * sizeof(double)); }
  c->hd = (double*)calloc(3 * N, sizeof(double));
  c->J = (double*)calloc(9 * N, sizeof(double));
  c->P = (double*)calloc(9 * N, sizeof(double));
  if (cfg->demag != LLG_DEMAG_OFF && cfg->demag_strength > 0.0) {
    c->demag = orc_demag_create(c->ng, c->ny, cfg->demag_thick, cfg->demag_strength,
                                cfg->demag == LLG_DEMAG_PADDED);
    orc_demag_self(c->demag, c->n0);
  }
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  free(c->active); free(c->ymsk3); free(c->active_ids); free(c->inactive_ids);
  free(c->w); free(c->cmsk); free(c->cnsk); free(c->hd); free(c->J); free(c->P);
  orc_demag_destroy(c->demag);
  free(c);
}

int orc_counts(orc_ctx* c, int* ncell, int* na, int* ni) {
  *ncell = c->ncell; *na = c->n_active; *ni = c->n_inactive;
  return 0;
}
const double* orc_ymsk(orc_ctx* c) { return c->ymsk3; }
const int* orc_active_ids(orc_ctx* c) { return c->active_ids; }
const int* orc_inactive_ids(orc_ctx* c) { return c->inactive_ids; }
const double* orc_hd(orc_ctx* c) { return c->hd; }
const double* orc_pinv(orc_ctx* c) { return c->P; }
/* applied field of the term sets in `which` (1 RHS, 2 Jv, 4 PC), the
 * oracle side of a field-sweep stage (no reference counterpart: the
 * reference's field is a compile-time constant, i4/2d_fft.cu:178-180) */
void orc_set_happ(orc_ctx* c, int which, const double* happ) {
  for (int a = 0; a < 3; a++) {
    if (which & 1) c->cfg.rhs.happ[a] = happ[a];
    if (which & 2) c->cfg.jv.happ[a] = happ[a];
    if (which & 4) c->cfg.pc.happ[a] = happ[a];
  }
}
void orc_self_coupling(orc_ctx* c, double* n3) { n3[0] = c->n0[0]; n3[1] = c->n0[1]; n3[2] = c->n0[2]; }

/* neighbour indices with the configured boundary; ghost flags for Neumann */
static inline void neighbours(const orc_ctx* c, int gx, int gy, int* lc, int* rc, int* uc,
                              int* dc) {
  const int ng = c->ng, ny = c->ny;
  if (c->cfg.boundary == LLG_BC_NEUMANN) {
    const int self = gy * ng + gx;
    *lc = gx > 0 ? self - 1 : self;
    *rc = gx < ng - 1 ? self + 1 : self;
    *uc = gy > 0 ? self - ng : self;
    *dc = gy < ny - 1 ? self + ng : self;
  } else {
    const int xl = (gx - 1 < 0) ? gx - 1 + ng : gx - 1;
    const int xr = (gx + 1 >= ng) ? gx + 1 - ng : gx + 1;
    const int yu = (gy - 1 < 0) ? gy - 1 + ny : gy - 1;
    const int yd = (gy + 1 >= ny) ? gy + 1 - ny : gy + 1;
    *lc = gy * ng + xl; *rc = gy * ng + xr; *uc = yu * ng + gx; *dc = yd * ng + gx;
  }
}

static inline void axes(const orc_ctx* c, const llg_terms* T, int cell, double msk[3],
                        double nsk[3]) {
  const int N = c->ncell;
  if (T->cell_axes && c->cmsk) {
    msk[0] = c->cmsk[cell]; msk[1] = c->cmsk[N + cell]; msk[2] = c->cmsk[2 * N + cell];
  } else {
    msk[0] = T->msk[0]; msk[1] = T->msk[1]; msk[2] = T->msk[2];
  }
  if (T->cell_axes && c->cnsk) {
    nsk[0] = c->cnsk[cell]; nsk[1] = c->cnsk[N + cell]; nsk[2] = c->cnsk[2 * N + cell];
  } else {
    nsk[0] = T->nsk[0]; nsk[1] = T->nsk[1]; nsk[2] = T->nsk[2];
  }
}

/* effective field h at one cell from the terms T (without demag) */
static void field_at(const orc_ctx* c, const llg_terms* T, const double* y, int cell, int gx,
                     int gy, double h[3]) {
  const int N = c->ncell;
  int lc, rc, uc, dc;
  neighbours(c, gx, gy, &lc, &rc, &uc, &dc);
  double msk[3], nsk[3];
  axes(c, T, cell, msk, nsk);
  const double m[3] = {y[cell], y[N + cell], y[2 * N + cell]};
  const int win = T->window && c->w;
  const double wl = win ? c->w[lc] : 1.0, wr = win ? c->w[rc] : 1.0;
  const double wu = win ? c->w[uc] : 1.0, wd = win ? c->w[dc] : 1.0;
  const double mdotmsk = m[0] * msk[0] + m[1] * msk[1] + m[2] * msk[2];
  for (int a = 0; a < 3; a++) {
    const double L = wl * y[a * N + lc], R = wr * y[a * N + rc];
    const double U = wu * y[a * N + uc], D = wd * y[a * N + dc];
    double ex = L + R + U + D;
    if (c->cfg.exch_self) ex = ex - 4.0 * m[a];
    double an;
    if (T->aniso == LLG_ANISO_CUBIC) an = msk[a] * T->chk * m[a] * (m[a] * m[a] - 1.0);
    else an = msk[a] * (T->chk * mdotmsk + T->cha);
    h[a] = T->happ[a] + T->che * ex + an + T->chb * nsk[a] * (L + R);
  }
}

static inline double gate_of(const orc_ctx* c, int gate, int cell) {
  if (gate == LLG_GATE_MASK) return c->active[cell] ? 1.0 : 0.0;
  if (gate == LLG_GATE_WINDOW) return c->w ? c->w[cell] : 1.0;
  return 1.0;
}

void orc_normalize(orc_ctx* c, double* y) {
  const int N = c->ncell;
  if (c->cfg.normalize == LLG_NORM_NONE) return;
  for (int i = 0; i < N; i++) {
    const double m1 = y[i], m2 = y[N + i], m3 = y[2 * N + i];
    const double ymp = sqrt(m1 * m1 + m2 * m2 + m3 * m3);
    if (c->cfg.normalize == LLG_NORM_EXACT) {
      if (ymp > 1.0e-30) {
        const double inv = 1.0 / ymp;
        y[i] = m1 * inv; y[N + i] = m2 * inv; y[2 * N + i] = m3 * inv;
      }
    } else {
      const double inv = 1.0 / (ymp + 1.0e-3);
      y[i] = m1 * inv; y[N + i] = m2 * inv; y[2 * N + i] = m3 * inv;
    }
  }
}

/* f(): [normalize] -> demag (or zero) -> [zero holes] -> RHS */
void orc_rhs_arrays(orc_ctx* c, double* y, double* yd) {
  const int N = c->ncell;
  const llg_config* cf = &c->cfg;
  c->nf++;
  orc_normalize(c, y);
  if (c->demag) orc_demag_apply(c->demag, y, cf->rhs.window ? c->w : NULL, c->hd);
  else memset(c->hd, 0, 3 * N * sizeof(double));
  const int compact = cf->exec == LLG_EXEC_COMPACT;
  if (compact)
    for (int t = 0; t < c->n_inactive; t++) {
      const int cell = c->inactive_ids[t];
      yd[cell] = yd[N + cell] = yd[2 * N + cell] = 0.0;
    }
  const int ncomp = compact ? c->n_active : N;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < ncomp; t++) {
    const int cell = compact ? c->active_ids[t] : t;
    const int gx = cell % c->ng, gy = cell / c->ng;
    double h[3];
    field_at(c, &cf->rhs, y, cell, gx, gy, h);
    h[0] += c->hd[cell]; h[1] += c->hd[N + cell]; h[2] += c->hd[2 * N + cell];
    const double m1 = y[cell], m2 = y[N + cell], m3 = y[2 * N + cell];
    const double mh = m1 * h[0] + m2 * h[1] + m3 * h[2];
    const double g = compact ? 1.0 : gate_of(c, cf->rhs.gate, cell);
    const double f1 = cf->chg * (m3 * h[1] - m2 * h[2]) + cf->alpha * (h[0] - mh * m1);
    const double f2 = cf->chg * (m1 * h[2] - m3 * h[0]) + cf->alpha * (h[1] - mh * m2);
    const double f3 = cf->chg * (m2 * h[0] - m1 * h[1]) + cf->alpha * (h[2] - mh * m3);
    if (compact) { yd[cell] = f1; yd[N + cell] = f2; yd[2 * N + cell] = f3; }
    else { yd[cell] = g * f1; yd[N + cell] = g * f2; yd[2 * N + cell] = g * f3; }
  }
}

/* analytic Jv (no demag), i4/jtv.cu:47-120 */
void orc_jtv_arrays(orc_ctx* c, const double* y, const double* v, double* Jv) {
  const int N = c->ncell;
  const llg_config* cf = &c->cfg;
  const llg_terms* T = &cf->jv;
  c->njv++;
  const int compact = cf->exec == LLG_EXEC_COMPACT;
  if (compact)
    for (int t = 0; t < c->n_inactive; t++) {
      const int cell = c->inactive_ids[t];
      Jv[cell] = Jv[N + cell] = Jv[2 * N + cell] = 0.0;
    }
  const int ncomp = compact ? c->n_active : N;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < ncomp; t++) {
    const int cell = compact ? c->active_ids[t] : t;
    const int gx = cell % c->ng, gy = cell / c->ng;
    double h[3], dh[3];
    field_at(c, T, y, cell, gx, gy, h);
    /* dh = same linear operator applied to v, plus the anisotropy derivative */
    int lc, rc, uc, dc;
    neighbours(c, gx, gy, &lc, &rc, &uc, &dc);
    double msk[3], nsk[3];
    axes(c, T, cell, msk, nsk);
    const double m[3] = {y[cell], y[N + cell], y[2 * N + cell]};
    const double vv[3] = {v[cell], v[N + cell], v[2 * N + cell]};
    const int win = T->window && c->w;
    const double wl = win ? c->w[lc] : 1.0, wr = win ? c->w[rc] : 1.0;
    const double wu = win ? c->w[uc] : 1.0, wd = win ? c->w[dc] : 1.0;
    const double vdotmsk = vv[0] * msk[0] + vv[1] * msk[1] + vv[2] * msk[2];
    for (int a = 0; a < 3; a++) {
      const double L = wl * v[a * N + lc], R = wr * v[a * N + rc];
      const double U = wu * v[a * N + uc], D = wd * v[a * N + dc];
      double ex = L + R + U + D;
      if (cf->exch_self) ex = ex - 4.0 * vv[a];
      double dan;
      if (T->aniso == LLG_ANISO_CUBIC) dan = (msk[a] * T->chk * (3.0 * m[a] * m[a] - 1.0)) * vv[a];
      else dan = msk[a] * T->chk * vdotmsk;
      dh[a] = T->che * ex + dan + T->chb * nsk[a] * (L + R);
    }
    const double m1 = m[0], m2 = m[1], m3 = m[2], v1 = vv[0], v2 = vv[1], v3 = vv[2];
    const double mh = m1 * h[0] + m2 * h[1] + m3 * h[2];
    const double dmh = (v1 * h[0] + v2 * h[1] + v3 * h[2]) + (m1 * dh[0] + m2 * dh[1] + m3 * dh[2]);
    const double g = compact ? 1.0 : gate_of(c, T->gate, cell);
    const double j1 = cf->chg * (v3 * h[1] + m3 * dh[1] - v2 * h[2] - m2 * dh[2]) +
                      cf->alpha * (dh[0] - dmh * m1 - mh * v1);
    const double j2 = cf->chg * (v1 * h[2] + m1 * dh[2] - v3 * h[0] - m3 * dh[0]) +
                      cf->alpha * (dh[1] - dmh * m2 - mh * v2);
    const double j3 = cf->chg * (v2 * h[0] + m2 * dh[0] - v1 * h[1] - m1 * dh[1]) +
                      cf->alpha * (dh[2] - dmh * m3 - mh * v3);
    if (compact) { Jv[cell] = j1; Jv[N + cell] = j2; Jv[2 * N + cell] = j3; }
    else { Jv[cell] = g * j1; Jv[N + cell] = g * j2; Jv[2 * N + cell] = g * j3; }
  }
}

/* local 3x3 Jacobian of the LLG RHS w.r.t. the cell's own m
 * (i4/precond.cu:42-50,174-197) */
static void local_J(const orc_ctx* c, const double* y, int cell, int gx, int gy, double J[9]) {
  const int N = c->ncell;
  const llg_config* cf = &c->cfg;
  const llg_terms* T = &cf->pc;
  double h[3];
  field_at(c, T, y, cell, gx, gy, h);
  h[0] += c->hd[cell]; h[1] += c->hd[N + cell]; h[2] += c->hd[2 * N + cell];
  double msk[3], nsk[3];
  axes(c, T, cell, msk, nsk);
  const double m1 = y[cell], m2 = y[N + cell], m3 = y[2 * N + cell];
  const double m[3] = {m1, m2, m3};
  double k[3];
  /* self derivative of the exchange: Neumann ghosts are the cell itself */
  double self_ex = 0.0;
  if (cf->boundary == LLG_BC_NEUMANN || cf->exch_self) {
    int ghosts = 0;
    if (cf->boundary == LLG_BC_NEUMANN) {
      ghosts += (gx == 0) + (gx == c->ng - 1) + (gy == 0) + (gy == c->ny - 1);
    }
    self_ex = T->che * ((double)ghosts - (cf->exch_self ? 4.0 : 0.0));
  }
  for (int a = 0; a < 3; a++) {
    double dan;
    if (T->aniso == LLG_ANISO_CUBIC) dan = T->chk * (3.0 * m[a] * m[a] - 1.0) * msk[a];
    else dan = msk[a] * msk[a] * T->chk;
    k[a] = dan + c->n0[a] + self_ex;
  }
  const double e1 = h[0] + m1 * k[0], e2 = h[1] + m2 * k[1], e3 = h[2] + m3 * k[2];
  const double mh = m1 * h[0] + m2 * h[1] + m3 * h[2];
  const double al = cf->alpha, cg = cf->chg;
  J[0] = al * (k[0] - e1 * m1 - mh);
  J[1] = cg * (m3 * k[1] - h[2]) - al * e2 * m1;
  J[2] = cg * (h[1] - m2 * k[2]) - al * e3 * m1;
  J[3] = cg * (h[2] - m3 * k[0]) - al * e1 * m2;
  J[4] = al * (k[1] - e2 * m2 - mh);
  J[5] = cg * (m1 * k[2] - h[0]) - al * e3 * m2;
  J[6] = cg * (m2 * k[0] - h[1]) - al * e1 * m3;
  J[7] = cg * (h[0] - m1 * k[1]) - al * e2 * m3;
  J[8] = al * (k[2] - e3 * m3 - mh);
}

/* A = I - gamma J, inverse by adjugate, identity when |det| <= 1e-30 */
static void invert_block(const double J[9], double gamma, double P[9]) {
  const double A00 = 1.0 - gamma * J[0], A01 = -gamma * J[1], A02 = -gamma * J[2];
  const double A10 = -gamma * J[3], A11 = 1.0 - gamma * J[4], A12 = -gamma * J[5];
  const double A20 = -gamma * J[6], A21 = -gamma * J[7], A22 = 1.0 - gamma * J[8];
  const double c00 = A11 * A22 - A12 * A21;
  const double c10 = A12 * A20 - A10 * A22;
  const double c20 = A10 * A21 - A11 * A20;
  const double det = A00 * c00 + A01 * c10 + A02 * c20;
  if (!(fabs(det) > 1e-30)) {
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

int orc_psetup_arrays(orc_ctx* c, const double* y, int jok, int* jcur, double gamma) {
  const int N = c->ncell;
  const llg_config* cf = &c->cfg;
  c->npset++;
  const int compact = cf->exec == LLG_EXEC_COMPACT;
  const int ncomp = compact ? c->n_active : N;
  const int two_phase = cf->pc_mode == LLG_PC_JOK;
  const int rebuild_J = !two_phase || !jok;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < ncomp; t++) {
    const int cell = compact ? c->active_ids[t] : t;
    const int gx = cell % c->ng, gy = cell / c->ng;
    double J[9], P[9];
    if (rebuild_J) {
      local_J(c, y, cell, gx, gy, J);
      for (int s = 0; s < 9; s++) c->J[s * N + cell] = J[s];
    } else {
      for (int s = 0; s < 9; s++) J[s] = c->J[s * N + cell];
    }
    invert_block(J, gamma, P);
    for (int s = 0; s < 9; s++) c->P[s * N + cell] = P[s];
  }
  *jcur = rebuild_J ? 1 : 0;
  return 0;
}

void orc_psolve_arrays(orc_ctx* c, const double* r, double* z) {
  const int N = c->ncell;
  const llg_config* cf = &c->cfg;
  c->npsol++;
  const int compact = cf->exec == LLG_EXEC_COMPACT;
  if (compact)
    for (int t = 0; t < c->n_inactive; t++) {
      const int cell = c->inactive_ids[t];
      z[cell] = z[N + cell] = z[2 * N + cell] = 0.0;
    }
  const int ncomp = compact ? c->n_active : N;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < ncomp; t++) {
    const int cell = compact ? c->active_ids[t] : t;
    const double* P = c->P;
    const double r0 = r[cell], r1 = r[N + cell], r2 = r[2 * N + cell];
    const double z0 = P[0 * N + cell] * r0 + P[1 * N + cell] * r1 + P[2 * N + cell] * r2;
    const double z1 = P[3 * N + cell] * r0 + P[4 * N + cell] * r1 + P[5 * N + cell] * r2;
    const double z2 = P[6 * N + cell] * r0 + P[7 * N + cell] * r1 + P[8 * N + cell] * r2;
    const double g = compact ? 1.0 : gate_of(c, cf->pc.gate, cell);
    if (compact) { z[cell] = z0; z[N + cell] = z1; z[2 * N + cell] = z2; }
    else { z[cell] = g * z0; z[N + cell] = g * z1; z[2 * N + cell] = g * z2; }
  }
}

/* ==================================================== serial N_Vector */
typedef struct { sunindextype n; double* d; int own; } orc_nv_content;
#define NVC(v) ((orc_nv_content*)(v)->content)
#define NVD(v) (NVC(v)->d)
#define NVN(v) (NVC(v)->n)

static long g_orc_nv_dots = 0;
long orc_nv_num_reductions(void) { return g_orc_nv_dots; }

static N_Vector orc_nv_alloc(sunindextype n, int alloc_data);
static N_Vector nv_clone(N_Vector w) { return orc_nv_alloc(NVN(w), 1); }
static N_Vector nv_cloneempty(N_Vector w) { return orc_nv_alloc(NVN(w), 0); }
static void nv_destroy(N_Vector v) {
  if (!v) return;
  if (NVC(v)->own) free(NVD(v));
  free(v->content); free(v->ops); free(v);
}
static N_Vector_ID nv_id(N_Vector v) { (void)v; return SUNDIALS_NVEC_SERIAL; }
static void nv_space(N_Vector v, sunindextype* lrw, sunindextype* liw) { *lrw = NVN(v); *liw = 1; }
static sunrealtype* nv_data(N_Vector v) { return NVD(v); }
static void nv_setdata(sunrealtype* d, N_Vector v) { NVD(v) = d; }
static sunindextype nv_len(N_Vector v) { return NVN(v); }
static void nv_linearsum(double a, N_Vector x, double b, N_Vector y, N_Vector z) {
  const double *X = NVD(x), *Y = NVD(y); double* Z = NVD(z);
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = a * X[i] + b * Y[i];
}
static void nv_const(double c, N_Vector z) {
  double* Z = NVD(z);
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = c;
}
static void nv_prod(N_Vector x, N_Vector y, N_Vector z) {
  const double *X = NVD(x), *Y = NVD(y); double* Z = NVD(z);
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = X[i] * Y[i];
}
static void nv_div(N_Vector x, N_Vector y, N_Vector z) {
  const double *X = NVD(x), *Y = NVD(y); double* Z = NVD(z);
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = X[i] / Y[i];
}
static void nv_scale(double c, N_Vector x, N_Vector z) {
  const double* X = NVD(x); double* Z = NVD(z);
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = c * X[i];
}
static void nv_abs(N_Vector x, N_Vector z) {
  const double* X = NVD(x); double* Z = NVD(z);
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = fabs(X[i]);
}
static void nv_inv(N_Vector x, N_Vector z) {
  const double* X = NVD(x); double* Z = NVD(z);
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = 1.0 / X[i];
}
static void nv_addconst(N_Vector x, double b, N_Vector z) {
  const double* X = NVD(x); double* Z = NVD(z);
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = X[i] + b;
}
/* reductions use long double accumulators (the Kahan/long-double pin) */
static double nv_dot(N_Vector x, N_Vector y) {
  const double *X = NVD(x), *Y = NVD(y);
  long double s = 0;
  g_orc_nv_dots++;
  for (sunindextype i = 0; i < NVN(x); i++) s += (long double)X[i] * Y[i];
  return (double)s;
}
static double nv_maxnorm(N_Vector x) {
  const double* X = NVD(x); double m = 0;
  g_orc_nv_dots++;
  for (sunindextype i = 0; i < NVN(x); i++) if (fabs(X[i]) > m) m = fabs(X[i]);
  return m;
}
static double nv_wrms(N_Vector x, N_Vector w) {
  const double *X = NVD(x), *W = NVD(w);
  long double s = 0;
  g_orc_nv_dots++;
  for (sunindextype i = 0; i < NVN(x); i++) { const long double p = (long double)X[i] * W[i]; s += p * p; }
  return sqrt((double)(s / NVN(x)));
}
static double nv_wrmsmask(N_Vector x, N_Vector w, N_Vector id) {
  const double *X = NVD(x), *W = NVD(w), *ID = NVD(id);
  long double s = 0;
  g_orc_nv_dots++;
  for (sunindextype i = 0; i < NVN(x); i++)
    if (ID[i] > 0.0) { const long double p = (long double)X[i] * W[i]; s += p * p; }
  return sqrt((double)(s / NVN(x)));
}
static double nv_min(N_Vector x) {
  const double* X = NVD(x); double m = X[0];
  g_orc_nv_dots++;
  for (sunindextype i = 1; i < NVN(x); i++) if (X[i] < m) m = X[i];
  return m;
}
static double nv_wl2(N_Vector x, N_Vector w) {
  const double *X = NVD(x), *W = NVD(w);
  long double s = 0;
  for (sunindextype i = 0; i < NVN(x); i++) { const long double p = (long double)X[i] * W[i]; s += p * p; }
  return sqrt((double)s);
}
static double nv_l1(N_Vector x) {
  const double* X = NVD(x); long double s = 0;
  for (sunindextype i = 0; i < NVN(x); i++) s += fabs(X[i]);
  return (double)s;
}
static void nv_compare(double c, N_Vector x, N_Vector z) {
  const double* X = NVD(x); double* Z = NVD(z);
  for (sunindextype i = 0; i < NVN(z); i++) Z[i] = (fabs(X[i]) >= c) ? 1.0 : 0.0;
}
static sunbooleantype nv_invtest(N_Vector x, N_Vector z) {
  const double* X = NVD(x); double* Z = NVD(z); int ok = 1;
  for (sunindextype i = 0; i < NVN(z); i++) {
    if (X[i] == 0.0) ok = 0; else Z[i] = 1.0 / X[i];
  }
  return ok;
}
static sunbooleantype nv_constrmask(N_Vector c, N_Vector x, N_Vector m) {
  const double *C = NVD(c), *X = NVD(x); double* M = NVD(m); int ok = 1;
  for (sunindextype i = 0; i < NVN(x); i++) {
    M[i] = 0.0;
    if (C[i] == 0.0) continue;
    const int bad = (fabs(C[i]) > 1.5 && X[i] * C[i] <= 0.0) || (fabs(C[i]) > 0.5 && X[i] * C[i] < 0.0);
    if (bad) { ok = 0; M[i] = 1.0; }
  }
  return ok;
}
static double nv_minquotient(N_Vector num, N_Vector den) {
  const double *A = NVD(num), *B = NVD(den); double m = 1.7976931348623157e308;
  for (sunindextype i = 0; i < NVN(num); i++) if (B[i] != 0.0 && A[i] / B[i] < m) m = A[i] / B[i];
  return m;
}
static SUNErrCode nv_lincomb(int nv, double* c, N_Vector* X, N_Vector z) {
  double* Z = NVD(z);
  const sunindextype n = NVN(z);
  /* z may alias X[0]: accumulate per element */
#pragma omp parallel for schedule(static)
  for (sunindextype i = 0; i < n; i++) {
    double s = 0.0;
    for (int j = 0; j < nv; j++) s += c[j] * NVD(X[j])[i];
    Z[i] = s;
  }
  return SUN_SUCCESS;
}
static SUNErrCode nv_scaleaddmulti(int nv, double* a, N_Vector x, N_Vector* Y, N_Vector* Z) {
  const double* X = NVD(x);
  for (int j = 0; j < nv; j++) {
    const double* Yj = NVD(Y[j]); double* Zj = NVD(Z[j]);
    for (sunindextype i = 0; i < NVN(x); i++) Zj[i] = a[j] * X[i] + Yj[i];
  }
  return SUN_SUCCESS;
}
static SUNErrCode nv_dotmulti(int nv, N_Vector x, N_Vector* Y, double* d) {
  for (int j = 0; j < nv; j++) d[j] = nv_dot(x, Y[j]);
  return SUN_SUCCESS;
}
static SUNErrCode nv_scalevecarray(int nv, double* c, N_Vector* X, N_Vector* Z) {
  for (int j = 0; j < nv; j++) nv_scale(c[j], X[j], Z[j]);
  return SUN_SUCCESS;
}
static SUNErrCode nv_lsumvecarray(int nv, double a, N_Vector* X, double b, N_Vector* Y, N_Vector* Z) {
  for (int j = 0; j < nv; j++) nv_linearsum(a, X[j], b, Y[j], Z[j]);
  return SUN_SUCCESS;
}
static SUNErrCode nv_constvecarray(int nv, double c, N_Vector* Z) {
  for (int j = 0; j < nv; j++) nv_const(c, Z[j]);
  return SUN_SUCCESS;
}
static SUNErrCode nv_wrmsvecarray(int nv, N_Vector* X, N_Vector* W, double* nrm) {
  for (int j = 0; j < nv; j++) nrm[j] = nv_wrms(X[j], W[j]);
  return SUN_SUCCESS;
}

static N_Vector orc_nv_alloc(sunindextype n, int alloc_data) {
  N_Vector v = (N_Vector)calloc(1, sizeof(*v));
  v->ops = (N_Vector_Ops)calloc(1, sizeof(struct _generic_N_Vector_Ops));
  orc_nv_content* c = (orc_nv_content*)calloc(1, sizeof(orc_nv_content));
  c->n = n;
  if (alloc_data) { c->d = (double*)calloc((size_t)n, sizeof(double)); c->own = 1; }
  v->content = c;
  N_Vector_Ops o = v->ops;
  o->nvgetvectorid = nv_id; o->nvclone = nv_clone; o->nvcloneempty = nv_cloneempty;
  o->nvdestroy = nv_destroy; o->nvspace = nv_space; o->nvgetarraypointer = nv_data;
  o->nvsetarraypointer = nv_setdata; o->nvgetlength = nv_len; o->nvgetlocallength = nv_len;
  o->nvlinearsum = nv_linearsum; o->nvconst = nv_const; o->nvprod = nv_prod; o->nvdiv = nv_div;
  o->nvscale = nv_scale; o->nvabs = nv_abs; o->nvinv = nv_inv; o->nvaddconst = nv_addconst;
  o->nvdotprod = nv_dot; o->nvmaxnorm = nv_maxnorm; o->nvwrmsnorm = nv_wrms;
  o->nvwrmsnormmask = nv_wrmsmask; o->nvmin = nv_min; o->nvwl2norm = nv_wl2; o->nvl1norm = nv_l1;
  o->nvcompare = nv_compare; o->nvinvtest = nv_invtest; o->nvconstrmask = nv_constrmask;
  o->nvminquotient = nv_minquotient;
  o->nvlinearcombination = nv_lincomb; o->nvscaleaddmulti = nv_scaleaddmulti;
  o->nvdotprodmulti = nv_dotmulti;
  o->nvlinearsumvectorarray = nv_lsumvecarray; o->nvscalevectorarray = nv_scalevecarray;
  o->nvconstvectorarray = nv_constvecarray; o->nvwrmsnormvectorarray = nv_wrmsvecarray;
  return v;
}

N_Vector orc_nv_new(sunindextype n) { return orc_nv_alloc(n, 1); }
double* orc_nv_data(N_Vector v) { return NVD(v); }
void orc_nv_destroy(N_Vector v) { nv_destroy(v); }

/* ============================================ CVODE callbacks on the oracle */
int orc_cv_rhs(double t, N_Vector y, N_Vector ydot, void* ud) {
  (void)t;
  orc_rhs_arrays((orc_ctx*)ud, NVD(y), NVD(ydot));
  return 0;
}
int orc_cv_jtv(N_Vector v, N_Vector Jv, double t, N_Vector y, N_Vector fy, void* ud, N_Vector tmp) {
  (void)t; (void)fy; (void)tmp;
  orc_jtv_arrays((orc_ctx*)ud, NVD(y), NVD(v), NVD(Jv));
  return 0;
}
int orc_cv_psetup(double t, N_Vector y, N_Vector fy, sunbooleantype jok, sunbooleantype* jcur,
                  double gamma, void* ud) {
  (void)t; (void)fy;
  int jc = 0;
  orc_psetup_arrays((orc_ctx*)ud, NVD(y), jok, &jc, gamma);
  *jcur = jc ? SUNTRUE : SUNFALSE;
  return 0;
}
int orc_cv_psolve(double t, N_Vector y, N_Vector fy, N_Vector r, N_Vector z, double gamma,
                  double delta, int lr, void* ud) {
  (void)t; (void)y; (void)fy; (void)gamma; (void)delta; (void)lr;
  orc_psolve_arrays((orc_ctx*)ud, NVD(r), NVD(z));
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
