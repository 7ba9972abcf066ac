// CPU check of the register-resident FFT (csrc/fft_regs.h): the lanes of a
// transform are emulated one step at a time (every lane finishes a step
// before any lane starts the next, which is what the barriers give on the
// GPU) and the results are compared with a direct O(L^2) DFT in long double.
// Prints one line per plan and case; exit status 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "fft_regs.h"

using rfft::Cplx;

template <class T>
struct Cases {
  int fails = 0;

  template <class F>
  void run_plan(const char* name, bool half) {
    constexpr int L = F::L, U = F::U, V = F::V;
    std::vector<Cplx<T>> tw(L);
    for (int q = 0; q < L; q++) {
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * q / L;
      tw[q] = rfft::mk<T>((T)cosl(a), (T)sinl(a));
    }
    std::mt19937_64 g(12345 + L + (half ? 7 : 0));
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    std::vector<Cplx<T>> x(L);
    for (int n = 0; n < L; n++)
      x[n] = (half && n >= L / 2) ? rfft::mk<T>(0, 0) : rfft::mk<T>((T)dist(g), (T)dist(g));
    // direct DFT in long double (angle table indexed by n k mod L)
    std::vector<long double> cw(L), sw(L);
    for (int q = 0; q < L; q++) {
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * q / L;
      cw[q] = cosl(a);
      sw[q] = sinl(a);
    }
    std::vector<long double> Xr(L), Xi(L);
    for (int k = 0; k < L; k++) {
      long double sr = 0, si = 0;
      for (int n = 0; n < L; n++) {
        const int q = (int)((long)n * k % L);
        sr += x[n].x * cw[q] - x[n].y * sw[q];
        si += x[n].x * sw[q] + x[n].y * cw[q];
      }
      Xr[k] = sr;
      Xi[k] = si;
    }
    std::vector<Cplx<T>> lds(F::kSlots);
    std::vector<std::vector<Cplx<T>>> r(U, std::vector<Cplx<T>>(V));
    for (int u = 0; u < U; u++)
      for (int b = 0; b < V; b++) r[u][b] = x[u + U * b];
    // forward, step by step
    for (int u = 0; u < U; u++) F::template f1_fwd<false>(r[u].data(), u, tw.data());
    if (half) {
      // the half-input F1 must equal the full one on zero-padded data
      std::vector<std::vector<Cplx<T>>> r2(U, std::vector<Cplx<T>>(V));
      for (int u = 0; u < U; u++) {
        for (int b = 0; b < V; b++) r2[u][b] = x[u + U * b];
        F::template f1_fwd<true>(r2[u].data(), u, tw.data());
        for (int b = 0; b < V; b++) r[u][b] = r2[u][b];
      }
    }
    for (int u = 0; u < U; u++) F::put_f1(lds.data(), r[u].data(), u);
    for (int u = 0; u < U; u++) F::template Stage<0>::get_in(lds.data(), r[u].data(), u);
    for (int u = 0; u < U; u++) F::template Stage<0>::fwd(r[u].data(), u, tw.data());
    rfft::sfor<1, F::M>([&](auto D) {
      for (int u = 0; u < U; u++) F::template Stage<D - 1>::put_out(lds.data(), r[u].data(), u);
      for (int u = 0; u < U; u++) F::template Stage<D>::get_in(lds.data(), r[u].data(), u);
      for (int u = 0; u < U; u++) F::template Stage<D>::fwd(r[u].data(), u, tw.data());
    });
    using Last = typename F::Last;
    std::vector<int> seen(L, 0);
    double err = 0, mag = 0;
    for (int u = 0; u < U; u++)
      for (int j = 0; j < Last::NG; j++)
        for (int t = 0; t < Last::Q; t++) {
          const int k = Last::k_of(u, j, t);
          seen[k]++;
          const Cplx<T> v = r[u][j * Last::Q + t];
          err = std::fmax(err, (double)std::hypot((long double)v.x - Xr[k], (long double)v.y - Xi[k]));
          mag = std::fmax(mag, (double)std::hypot(Xr[k], Xi[k]));
        }
    bool cover = true;
    for (int k = 0; k < L; k++) cover = cover && seen[k] == 1;
    const double tol = (sizeof(T) == 8 ? 1e-14 : 2e-6) * std::log2((double)L);
    const bool ok_f = cover && err <= tol * mag;
    std::printf("%-28s %s forward: rel err %.3e  cover %d  %s\n", name, half ? "half" : "full",
                err / mag, (int)cover, ok_f ? "ok" : "FAIL");
    const std::vector<std::vector<Cplx<T>>> rf = r;   // forward result
    // inverse back to x (times L); half: only the first half of the outputs
    rfft::sfor<0, F::M - 1>([&](auto I) {
      constexpr int D = F::M - 1 - I;
      for (int u = 0; u < U; u++) F::template Stage<D>::inv(r[u].data(), u, tw.data());
      for (int u = 0; u < U; u++) F::template Stage<D>::put_in(lds.data(), r[u].data(), u);
      for (int u = 0; u < U; u++) F::template Stage<D - 1>::get_out(lds.data(), r[u].data(), u);
    });
    for (int u = 0; u < U; u++) F::template Stage<0>::inv(r[u].data(), u, tw.data());
    for (int u = 0; u < U; u++) F::template Stage<0>::put_in(lds.data(), r[u].data(), u);
    for (int u = 0; u < U; u++) F::get_f1(lds.data(), r[u].data(), u);
    double ierr = 0, imag = 0;
    for (int u = 0; u < U; u++) {
      if (half) F::template f1_inv<true>(r[u].data(), u, tw.data());
      else F::template f1_inv<false>(r[u].data(), u, tw.data());
      for (int b = 0; b < (half ? V / 2 : V); b++) {
        const int n = u + U * b;
        const Cplx<T> v = r[u][b];
        ierr = std::fmax(ierr, std::hypot((double)v.x / L - x[n].x, (double)v.y / L - x[n].y));
        imag = std::fmax(imag, std::hypot((double)x[n].x, (double)x[n].y));
      }
    }
    const bool ok_i = ierr <= tol * imag;
    std::printf("%-28s %s inverse: rel err %.3e  %s\n", name, half ? "half" : "full", ierr / imag,
                ok_i ? "ok" : "FAIL");
    // split exchanges (real parts, then imaginary parts, through a buffer of
    // T, as forward_split / inverse_split do) give the whole-value results
    // bit for bit, forward and inverse
    std::vector<T> sbuf(F::kSlots);
    std::vector<std::vector<Cplx<T>>> s(U, std::vector<Cplx<T>>(V));
    for (int u = 0; u < U; u++) {
      for (int b = 0; b < V; b++) s[u][b] = x[u + U * b];
      if (half) F::template f1_fwd<true>(s[u].data(), u, tw.data());
      else F::template f1_fwd<false>(s[u].data(), u, tw.data());
    }
    rfft::sfor<0, 2>([&](auto Pt) {
      for (int u = 0; u < U; u++) F::template put_f1_part<Pt>(sbuf.data(), s[u].data(), u);
      for (int u = 0; u < U; u++) F::template Stage<0>::template get_in_part<Pt>(sbuf.data(), s[u].data(), u);
    });
    for (int u = 0; u < U; u++) F::template Stage<0>::fwd(s[u].data(), u, tw.data());
    rfft::sfor<1, F::M>([&](auto D) {
      rfft::sfor<0, 2>([&](auto Pt) {
        for (int u = 0; u < U; u++)
          F::template Stage<D - 1>::template put_out_part<Pt>(sbuf.data(), s[u].data(), u);
        for (int u = 0; u < U; u++)
          F::template Stage<D>::template get_in_part<Pt>(sbuf.data(), s[u].data(), u);
      });
      for (int u = 0; u < U; u++) F::template Stage<D>::fwd(s[u].data(), u, tw.data());
    });
    bool ok_s = true;
    for (int u = 0; u < U; u++)
      for (int b = 0; b < V; b++)
        ok_s = ok_s && s[u][b].x == rf[u][b].x && s[u][b].y == rf[u][b].y;
    rfft::sfor<0, F::M - 1>([&](auto I) {
      constexpr int D = F::M - 1 - I;
      for (int u = 0; u < U; u++) F::template Stage<D>::inv(s[u].data(), u, tw.data());
      rfft::sfor<0, 2>([&](auto Pt) {
        for (int u = 0; u < U; u++)
          F::template Stage<D>::template put_in_part<Pt>(sbuf.data(), s[u].data(), u);
        for (int u = 0; u < U; u++)
          F::template Stage<D - 1>::template get_out_part<Pt>(sbuf.data(), s[u].data(), u);
      });
    });
    for (int u = 0; u < U; u++) F::template Stage<0>::inv(s[u].data(), u, tw.data());
    rfft::sfor<0, 2>([&](auto Pt) {
      for (int u = 0; u < U; u++) F::template Stage<0>::template put_in_part<Pt>(sbuf.data(), s[u].data(), u);
      for (int u = 0; u < U; u++) F::template get_f1_part<Pt>(sbuf.data(), s[u].data(), u);
    });
    for (int u = 0; u < U; u++) {
      if (half) F::template f1_inv<true>(s[u].data(), u, tw.data());
      else F::template f1_inv<false>(s[u].data(), u, tw.data());
      for (int b = 0; b < (half ? V / 2 : V); b++)
        ok_s = ok_s && s[u][b].x == r[u][b].x && s[u][b].y == r[u][b].y;
    }
    std::printf("%-28s %s split exchange == whole: %s\n", name, half ? "half" : "full",
                ok_s ? "ok" : "FAIL");
    if (!ok_f || !ok_i || !ok_s) fails++;
  }
};

#define PLAN(T, ...) rfft::RegFft<T, __VA_ARGS__>

int main() {
  Cases<double> d;
  Cases<float> f;
#define BOTH(name, ...)                                   \
  d.run_plan<PLAN(double, __VA_ARGS__)>(name, false);     \
  d.run_plan<PLAN(double, __VA_ARGS__)>(name, true);      \
  f.run_plan<PLAN(float, __VA_ARGS__)>(name " f32", false); \
  f.run_plan<PLAN(float, __VA_ARGS__)>(name " f32", true);
  BOTH("L=128 U=16 (4,4)", 128, 16, 4, 4)
  BOTH("L=256 U=32 (8,4)", 256, 32, 8, 4)
  BOTH("L=384 U=16 (4,4)", 384, 16, 4, 4)
  BOTH("L=512 U=64 (8,8)", 512, 64, 8, 8)
  BOTH("L=1024 U=64 (8,8)", 1024, 64, 8, 8)
  BOTH("L=1536 U=64 (8,8)", 1536, 64, 8, 8)
  BOTH("L=2048 U=64 (8,8)", 2048, 64, 8, 8)
  BOTH("L=2048 U=128 (16,8)", 2048, 128, 16, 8)
  BOTH("L=4096 U=256 (16,16)", 4096, 256, 16, 16)
  BOTH("L=8192 U=512 (8,8,8)", 8192, 512, 8, 8, 8)
  BOTH("L=1024 U=128 (8,4,4)", 1024, 128, 8, 4, 4)
  BOTH("L=2048 U=256 (8,8,4)", 2048, 256, 8, 8, 4)
  BOTH("L=4096 U=512 (8,8,8)", 4096, 512, 8, 8, 8)
  BOTH("L=8192 U=1024 (8,8,4,4)", 8192, 1024, 8, 8, 4, 4)
  const int fails = d.fails + f.fails;
  std::printf("%s (%d failing)\n", fails ? "FAIL" : "ALL OK", fails);
  return fails ? 1 : 0;
}
