// fft_regs.h -- register-resident mixed-radix FFT for the demag passes.
//
// A transform of length L = U * V runs on U lanes that each hold V complex
// values in registers.  With n = u + U b (lane u, register b) and
// k = kb + V ku:
//
//   X[kb + V ku] = sum_u W_U^(u ku) [ W_L^(u kb) sum_b x[u + U b] W_V^(b kb) ]
//
// Stage F1 is the inner V-point DFT in each lane's registers followed by the
// twiddle W_L^(u kb).  The outer U-point DFTs (one per kb) run as M more
// in-register stages of radices Q_0 .. Q_{M-1} (product U, each dividing V):
// stage d takes the most significant remaining digit of u.  Between stages
// the values move through LDS once (one store, one load); the first stage
// needs no LDS on the way in (each lane loads its own inputs) and the last
// forward stage leaves every value in registers at a known output index, so
// a pointwise product (the demag tensor) needs no exchange at all and the
// inverse transform starts from the same registers.  The inverse runs the
// stages backwards with conjugate twiddles (unnormalised).
//
// LDS layout (per transform): the canonical index of a value at any depth is
// c = kb + V * X; it is stored at kb + (V + 1) * X (one spare slot per V
// values), so the stride-V stores of stage F1 spread over the banks while
// every other access is contiguous in kb across lanes.
//
// Zero-padded transforms: with only the first half of the inputs non-zero
// (HALF_IN) stage F1 is two V/2-point DFTs; when only the first half of the
// outputs is wanted (HALF_OUT) the inverse F1 is two V/2-point DFTs.
//
// All functions are host+device so a CPU build can emulate the lanes.
#pragma once

#include <type_traits>
#include <utility>

#if defined(__HIPCC__)
#define RF_HD __host__ __device__
#define RF_INL __forceinline__
#else
#define RF_HD
#define RF_INL inline __attribute__((always_inline))
#endif

namespace rfft {

// ------------------------------------------------------------ complex
template <class T>
struct alignas(2 * sizeof(T)) Cplx {
  T x, y;
};

template <class T>
RF_HD RF_INL Cplx<T> mk(T x, T y) {
  Cplx<T> r;
  r.x = x;
  r.y = y;
  return r;
}

// Global-memory access of a complex value as ONE native two-element vector
// (a single 16- or 8-byte load/store).  Accessing the two members
// separately lets the compiler split every complex load into two scalar
// loads, and the gfx950 backend has been seen to merge the imaginary-part
// loads of two different elements of a table into one (the FP64 384-point
// inverse row kernel read another element's imaginary parts); whole-vector
// accesses leave nothing to merge.
#if defined(__HIPCC__)
template <class T>
struct VecOf;
template <>
struct VecOf<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct VecOf<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <class T>
RF_HD RF_INL Cplx<T> ldc(const Cplx<T>* p) {
  const typename VecOf<T>::type v = *reinterpret_cast<const typename VecOf<T>::type*>(p);
  Cplx<T> r;
  r.x = v.x;
  r.y = v.y;
  return r;
}
template <class T>
RF_HD RF_INL void stc(Cplx<T>* p, Cplx<T> c) {
  typename VecOf<T>::type v;
  v.x = c.x;
  v.y = c.y;
  *reinterpret_cast<typename VecOf<T>::type*>(p) = v;
}
#else
template <class T>
inline Cplx<T> ldc(const Cplx<T>* p) { return *p; }
template <class T>
inline void stc(Cplx<T>* p, Cplx<T> c) { *p = c; }
#endif

RF_HD RF_INL double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
RF_HD RF_INL float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <class C2> RF_HD RF_INL C2 cmul(C2 a, C2 b) {
  C2 r;
  r.x = fmad(a.x, b.x, -(a.y * b.y));
  r.y = fmad(a.x, b.y, a.y * b.x);
  return r;
}
// a * conj(b)
template <class C2> RF_HD RF_INL C2 cmulc(C2 a, C2 b) {
  C2 r;
  r.x = fmad(a.x, b.x, a.y * b.y);
  r.y = fmad(a.y, b.x, -(a.x * b.y));
  return r;
}
template <class C2> RF_HD RF_INL C2 cadd(C2 a, C2 b) { C2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <class C2> RF_HD RF_INL C2 csub(C2 a, C2 b) { C2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <class C2> RF_HD RF_INL C2 cconj(C2 a) { C2 r; r.x = a.x; r.y = -a.y; return r; }
// sign * i * a
template <class C2, class T> RF_HD RF_INL C2 cmuli(C2 a, T sign) {
  C2 r;
  r.x = -sign * a.y;
  r.y = sign * a.x;
  return r;
}

// ------------------------------------------- compile-time cos / sin
// cos and sin of 2 pi p / q, reduced with exact integer arithmetic to an
// angle in [0, pi/4] and summed as Taylor series (about 1 ulp).
struct CS {
  double c, s;
};
constexpr double kTwoPi = 6.28318530717958647692528676655900577;
constexpr double taylor_sin(double x) {
  double x2 = x * x, term = x, s = x;
  for (int n = 1; n < 14; n++) {
    term *= -x2 / (double)((2 * n) * (2 * n + 1));
    s += term;
  }
  return s;
}
constexpr double taylor_cos(double x) {
  double x2 = x * x, term = 1.0, s = 1.0;
  for (int n = 1; n < 14; n++) {
    term *= -x2 / (double)((2 * n - 1) * (2 * n));
    s += term;
  }
  return s;
}
constexpr CS cs_turn(long p, long q) {
  p %= q;
  if (p < 0) p += q;
  bool neg_s = false, neg_c = false, swap = false;
  if (2 * p > q) { p = q - p; neg_s = true; }             // r -> 1 - r
  if (4 * p > q) { p = q - 2 * p; q = 2 * q; neg_c = true; }   // r -> 1/2 - r
  if (8 * p > q) { p = q - 4 * p; q = 4 * q; swap = true; }    // r -> 1/4 - r
  const double th = kTwoPi * (double)p / (double)q;
  double c = taylor_cos(th), s = taylor_sin(th);
  if (swap) { const double t = c; c = s; s = t; }
  if (neg_c) c = -c;
  if (neg_s) s = -s;
  return CS{c, s};
}

// x * exp(S 2 pi i E / N) with E, N, S compile-time
template <int N, int S, int E, class T>
RF_HD RF_INL Cplx<T> twc(Cplx<T> x) {
  constexpr int e = ((E % N) + N) % N;
  if constexpr (e == 0) {
    return x;
  } else if constexpr ((8 * e) % N == 0) {
    constexpr int o = 8 * e / N;   // octant 1..7 of the turn
    if constexpr (o == 4) {
      return mk<T>(-x.x, -x.y);
    } else if constexpr (o == 2 || o == 6) {
      // exp(S i pi/2) = S i ; exp(S i 3pi/2) = -S i
      constexpr int sg = (o == 2) ? S : -S;
      if constexpr (sg > 0) return mk<T>(-x.y, x.x);
      else return mk<T>(x.y, -x.x);
    } else {
      constexpr CS w = cs_turn((long)S * e, N);
      constexpr int sc = w.c > 0 ? 1 : -1, ss = w.s > 0 ? 1 : -1;
      const T r = T(0.70710678118654752440);
      // (x.x c - x.y s, x.x s + x.y c), |c| = |s| = r
      const T re = (sc > 0 ? x.x : -x.x) - (ss > 0 ? x.y : -x.y);
      const T im = (ss > 0 ? x.x : -x.x) + (sc > 0 ? x.y : -x.y);
      return mk<T>(r * re, r * im);
    }
  } else {
    constexpr CS w = cs_turn((long)S * e, N);
    const T c = T(w.c), s = T(w.s);
    return mk<T>(fmad(x.x, c, -(x.y * s)), fmad(x.x, s, x.y * c));
  }
}

// ------------------------------------------------------ static for
template <int B, int E, class F>
RF_HD RF_INL void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// ---------------------------------------------- in-register DFTs
// v[k] = sum_n v[n] exp(S 2 pi i n k / N), natural order in and out
template <int N, int S, class T>
RF_HD RF_INL void rdft(Cplx<T>* v);

template <int S, class T>
RF_HD RF_INL void dft2(Cplx<T>* v) {
  const Cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <int S, class T>
RF_HD RF_INL void dft4(Cplx<T>* v) {
  const T sign = T(S);
  const Cplx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  const Cplx<T> c = cadd(v[1], v[3]), d = cmuli(csub(v[1], v[3]), sign);
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}
template <int S, class T>
RF_HD RF_INL void dft8(Cplx<T>* v) {
  Cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
  Cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<S, T>(e);
  dft4<S, T>(o);
  const Cplx<T> t0 = o[0];
  const Cplx<T> t1 = twc<8, S, 1>(o[1]);
  const Cplx<T> t2 = twc<8, S, 2>(o[2]);
  const Cplx<T> t3 = twc<8, S, 3>(o[3]);
  v[0] = cadd(e[0], t0); v[4] = csub(e[0], t0);
  v[1] = cadd(e[1], t1); v[5] = csub(e[1], t1);
  v[2] = cadd(e[2], t2); v[6] = csub(e[2], t2);
  v[3] = cadd(e[3], t3); v[7] = csub(e[3], t3);
}
template <int S, class T>
RF_HD RF_INL void dft3(Cplx<T>* v) {
  const T c = T(-0.5), s = T(S) * T(0.86602540378443864676);
  const Cplx<T> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
  const Cplx<T> a = mk<T>(fmad(c, t1.x, v[0].x), fmad(c, t1.y, v[0].y));
  const Cplx<T> b = mk<T>(-s * t2.y, s * t2.x);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(a, b);
  v[2] = csub(a, b);
}

// inner radix of a composite length (n = n1 + N1 n2, k = k2 + N2 k1)
template <int N>
constexpr int inner_radix() {
  if (N % 3 == 0) return 3;
  if (N == 16) return 4;
  if (N % 8 == 0) return 8;
  if (N % 4 == 0) return 4;
  return 2;
}

template <int N, int S, class T>
RF_HD RF_INL void rdft(Cplx<T>* v) {
  if constexpr (N == 1) {
  } else if constexpr (N == 2) {
    dft2<S, T>(v);
  } else if constexpr (N == 3) {
    dft3<S, T>(v);
  } else if constexpr (N == 4) {
    dft4<S, T>(v);
  } else if constexpr (N == 8) {
    dft8<S, T>(v);
  } else {
    constexpr int N1 = inner_radix<N>(), N2 = N / N1;
    static_assert(N1 * N2 == N && N2 > 1, "unsupported register DFT length");
    Cplx<T> a[N1][N2];
    sfor<0, N1>([&](auto I1) {
      constexpr int n1 = I1;
      sfor<0, N2>([&](auto I2) { a[n1][I2] = v[n1 + N1 * I2]; });
      rdft<N2, S, T>(a[n1]);
      sfor<1, N2>([&](auto K2) { a[n1][K2] = twc<N, S, n1 * K2>(a[n1][K2]); });
    });
    sfor<0, N2>([&](auto K2) {
      Cplx<T> b[N1];
      sfor<0, N1>([&](auto I1) { b[I1] = a[I1][K2]; });
      rdft<N1, S, T>(b);
      sfor<0, N1>([&](auto K1) { v[K2 + N2 * K1] = b[K1]; });
    });
  }
}

// DFT of N values whose second half is zero (v[N/2 .. N) are not read)
template <int N, int S, class T>
RF_HD RF_INL void rdft_half_in(Cplx<T>* v) {
  constexpr int M = N / 2;
  Cplx<T> e[M], o[M];
  sfor<0, M>([&](auto B) {
    e[B] = v[B];
    o[B] = twc<N, S, B>(v[B]);
  });
  rdft<M, S, T>(e);
  rdft<M, S, T>(o);
  sfor<0, M>([&](auto K) {
    v[2 * K] = e[K];
    v[2 * K + 1] = o[K];
  });
}

// first half of the outputs of a DFT of N values (written to v[0 .. N/2))
template <int N, int S, class T>
RF_HD RF_INL void rdft_half_out(Cplx<T>* v) {
  constexpr int M = N / 2;
  Cplx<T> e[M], o[M];
  sfor<0, M>([&](auto K) {
    e[K] = v[2 * K];
    o[K] = v[2 * K + 1];
  });
  rdft<M, S, T>(e);
  rdft<M, S, T>(o);
  sfor<0, M>([&](auto B) { v[B] = cadd(e[B], twc<N, S, B>(o[B])); });
}

// An opaque copy of the lane index: the index arithmetic of each stage (LDS
// slots, twiddle offsets) is recomputed from it where the stage runs, so
// the compiler neither keeps per-lane offsets of every transform of a
// kernel live in registers nor merges the twiddle loads of different
// transforms (which would hold all of them in registers at once).
RF_HD RF_INL int opaque(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(v) : : "memory");
#endif
  return v;
}

// ------------------------------------------------------ the plan
// tw[q] = exp(-2 pi i q / L), q in [0, L)
template <class T, int L_, int U_, int... Qs>
struct RegFft {
  using C = Cplx<T>;
  static constexpr int L = L_, U = U_, V = L_ / U_;
  static constexpr int M = sizeof...(Qs);
  static constexpr int kQ[M] = {Qs...};
  static constexpr int kSlots = L + L / V;   // padded LDS slots per transform
  static constexpr int kThreads = U >= 64 ? U : 64;
  static constexpr int kPerGroup = U >= 64 ? 1 : 64 / U;   // transforms per workgroup

  static constexpr int qprod(int d) {
    int p = 1;
    for (int i = 0; i < d; i++) p *= kQ[i];
    return p;
  }
  static_assert(L_ % U_ == 0, "L must be U * V");
  static_assert((U_ & (U_ - 1)) == 0, "U must be a power of two");
  static_assert(qprod(M) == U_, "stage radices must multiply to U");

  // ---------------- twiddle runs
  // w_t = tw[base t] = W_L^(base t) for t in [1, N) (base t < L), handed to
  // fn(t, w_t) in increasing t.  Each run of four consecutive t reads its
  // first value from the table and forms the other three as products with
  // w_1, w_2 = w_1 w_1 and w_3 = w_2 w_1 (at most three roundings deep, a
  // few ulp): one table read per four twiddles.  The table reads go through
  // the vector-memory address unit, one per CU, which every lane-varying
  // read occupies for 16 cycles per wave (16-byte values); the products run
  // on the four SIMDs.
  template <int N, class Fn>
  RF_HD RF_INL static void twiddle_run(const C* tw, unsigned base, Fn&& fn) {
    if constexpr (N > 1) {
      const C w1 = ldc(tw + base);
      fn(std::integral_constant<int, 1>{}, w1);
      if constexpr (N > 2) {
        const C w2 = cmul(w1, w1);
        fn(std::integral_constant<int, 2>{}, w2);
        if constexpr (N > 3) {
          const C w3 = cmul(w2, w1);
          fn(std::integral_constant<int, 3>{}, w3);
          sfor<1, (N + 3) / 4>([&](auto Cq) {
            constexpr int c = 4 * Cq;
            const C wc = ldc(tw + base * (unsigned)c);
            fn(std::integral_constant<int, c>{}, wc);
            if constexpr (c + 1 < N) fn(std::integral_constant<int, c + 1>{}, cmul(wc, w1));
            if constexpr (c + 2 < N) fn(std::integral_constant<int, c + 2>{}, cmul(wc, w2));
            if constexpr (c + 3 < N) fn(std::integral_constant<int, c + 3>{}, cmul(wc, w3));
          });
        }
      }
    }
  }

  // w_t = W_L^(u t) for t in [1, N) from the one table read w_1 = tw[u]:
  // w_2 = w_1 w_1, w_3 = w_2 w_1, w_4 = w_2 w_2, and each further run of
  // four from its first value w_c = w_(c-4) w_4 times w_1, w_2, w_3 (at
  // most N/4 + 3 roundings deep).  Stage F1's table index u t strides the
  // lanes by t values, so every read past t = 1 would touch one cache
  // sector per lane; the read of tw[u] is contiguous across the lanes.
  template <int N, class Fn>
  RF_HD RF_INL static void twiddle_pow(const C* tw, unsigned u, Fn&& fn) {
    if constexpr (N > 1) {
      const C w1 = ldc(tw + u);
      fn(std::integral_constant<int, 1>{}, w1);
      if constexpr (N > 2) {
        const C w2 = cmul(w1, w1);
        fn(std::integral_constant<int, 2>{}, w2);
        if constexpr (N > 3) {
          const C w3 = cmul(w2, w1);
          fn(std::integral_constant<int, 3>{}, w3);
          if constexpr (N > 4) {
            const C w4 = cmul(w2, w2);
            C wc = w4;
            sfor<1, (N + 3) / 4>([&](auto Cq) {
              constexpr int c = 4 * Cq;
              if constexpr (c > 4) wc = cmul(wc, w4);
              fn(std::integral_constant<int, c>{}, wc);
              if constexpr (c + 1 < N) fn(std::integral_constant<int, c + 1>{}, cmul(wc, w1));
              if constexpr (c + 2 < N) fn(std::integral_constant<int, c + 2>{}, cmul(wc, w2));
              if constexpr (c + 3 < N) fn(std::integral_constant<int, c + 3>{}, cmul(wc, w3));
            });
          }
        }
      }
    }
  }

  // ---------------- stage F1 (lane u holds x[u + U b], b in [0, V))
  template <bool HALF_IN>
  RF_HD RF_INL static void f1_fwd(C* r, int u, const C* tw) {
    if constexpr (HALF_IN) rdft_half_in<V, -1, T>(r);
    else rdft<V, -1, T>(r);
    twiddle_pow<V>(tw, (unsigned)opaque(u), [&](auto KB, const C& w) { r[KB] = cmul(r[KB], w); });
  }
  template <bool HALF_OUT>
  RF_HD RF_INL static void f1_inv(C* r, int u, const C* tw) {
    twiddle_pow<V>(tw, (unsigned)opaque(u), [&](auto KB, const C& w) { r[KB] = cmulc(r[KB], w); });
    if constexpr (HALF_OUT) rdft_half_out<V, 1, T>(r);
    else rdft<V, 1, T>(r);
  }
  RF_HD RF_INL static void put_f1(C* lds, const C* r, int u) {
    sfor<0, V>([&](auto KB) { lds[KB + (V + 1) * u] = r[KB]; });
  }
  RF_HD RF_INL static void get_f1(const C* lds, C* r, int u) {
    sfor<0, V>([&](auto KB) { r[KB] = lds[KB + (V + 1) * u]; });
  }

  // ---------------- stage D of the outer DFT
  // lane's group j: g = lane + U j -> (kb, kpre, a1); values t in [0, Q)
  //   in  (depth D)  : c = kb + V (kpre + QP (a1 + P t))
  //   out (depth D+1): c = kb + V (kpre + QP t + QP Q a1)
  template <int D>
  struct Stage {
    static constexpr int Q = kQ[D], QP = qprod(D), P = U / (QP * Q), NG = V / Q;
    static_assert(V % Q == 0, "stage radix must divide V");
    RF_HD RF_INL static void group(int lane, int j, int& kb, int& kpre, int& a1) {
      const int g = lane + U * j;
      kb = g % V;
      const int r = g / V;
      kpre = r % QP;
      a1 = r / QP;
    }
    // padded LDS slot of canonical kb + V X
    RF_HD RF_INL static int slot(int kb, int X) { return kb + (V + 1) * X; }
    RF_HD RF_INL static int x_in(int kpre, int a1, int t) { return kpre + QP * (a1 + P * t); }
    RF_HD RF_INL static int x_out(int kpre, int a1, int t) { return kpre + QP * t + QP * Q * a1; }

    RF_HD RF_INL static void get_in(const C* lds, C* r, int lane) {
      sfor<0, NG>([&](auto J) {
        int kb, kp, a1;
        group(lane, J, kb, kp, a1);
        sfor<0, Q>([&](auto Tt) { r[J * Q + Tt] = lds[slot(kb, x_in(kp, a1, Tt))]; });
      });
    }
    RF_HD RF_INL static void put_in(C* lds, const C* r, int lane) {
      sfor<0, NG>([&](auto J) {
        int kb, kp, a1;
        group(lane, J, kb, kp, a1);
        sfor<0, Q>([&](auto Tt) { lds[slot(kb, x_in(kp, a1, Tt))] = r[J * Q + Tt]; });
      });
    }
    RF_HD RF_INL static void get_out(const C* lds, C* r, int lane) {
      sfor<0, NG>([&](auto J) {
        int kb, kp, a1;
        group(lane, J, kb, kp, a1);
        sfor<0, Q>([&](auto Tt) { r[J * Q + Tt] = lds[slot(kb, x_out(kp, a1, Tt))]; });
      });
    }
    RF_HD RF_INL static void put_out(C* lds, const C* r, int lane) {
      sfor<0, NG>([&](auto J) {
        int kb, kp, a1;
        group(lane, J, kb, kp, a1);
        sfor<0, Q>([&](auto Tt) { lds[slot(kb, x_out(kp, a1, Tt))] = r[J * Q + Tt]; });
      });
    }
    // forward: DFT over t, then W_{P Q}^(a1 t) = W_L^(a1 t V QP)
    RF_HD RF_INL static void fwd(C* r, int lane, const C* tw) {
      sfor<0, NG>([&](auto J) {
        rdft<Q, -1, T>(r + J * Q);
        if constexpr (P > 1) {
          int kb, kp, a1;
          group(opaque(lane), J, kb, kp, a1);
          twiddle_run<Q>(tw, (unsigned)(a1 * V * QP),
                         [&](auto Tt, const C& w) { r[J * Q + Tt] = cmul(r[J * Q + Tt], w); });
        }
      });
    }
    RF_HD RF_INL static void inv(C* r, int lane, const C* tw) {
      sfor<0, NG>([&](auto J) {
        if constexpr (P > 1) {
          int kb, kp, a1;
          group(opaque(lane), J, kb, kp, a1);
          twiddle_run<Q>(tw, (unsigned)(a1 * V * QP),
                         [&](auto Tt, const C& w) { r[J * Q + Tt] = cmulc(r[J * Q + Tt], w); });
        }
        rdft<Q, 1, T>(r + J * Q);
      });
    }
    // one part (0: real, 1: imaginary) of every value through a buffer of
    // T, at the same slots as the whole-value forms above
    template <int PART>
    RF_HD RF_INL static void put_in_part(T* buf, const C* r, int lane) {
      sfor<0, NG>([&](auto J) {
        int kb, kp, a1;
        group(lane, J, kb, kp, a1);
        sfor<0, Q>([&](auto Tt) { buf[slot(kb, x_in(kp, a1, Tt))] = PART ? r[J * Q + Tt].y : r[J * Q + Tt].x; });
      });
    }
    template <int PART>
    RF_HD RF_INL static void get_in_part(const T* buf, C* r, int lane) {
      sfor<0, NG>([&](auto J) {
        int kb, kp, a1;
        group(lane, J, kb, kp, a1);
        sfor<0, Q>([&](auto Tt) {
          const T v = buf[slot(kb, x_in(kp, a1, Tt))];
          if constexpr (PART) r[J * Q + Tt].y = v;
          else r[J * Q + Tt].x = v;
        });
      });
    }
    template <int PART>
    RF_HD RF_INL static void put_out_part(T* buf, const C* r, int lane) {
      sfor<0, NG>([&](auto J) {
        int kb, kp, a1;
        group(lane, J, kb, kp, a1);
        sfor<0, Q>([&](auto Tt) { buf[slot(kb, x_out(kp, a1, Tt))] = PART ? r[J * Q + Tt].y : r[J * Q + Tt].x; });
      });
    }
    template <int PART>
    RF_HD RF_INL static void get_out_part(const T* buf, C* r, int lane) {
      sfor<0, NG>([&](auto J) {
        int kb, kp, a1;
        group(lane, J, kb, kp, a1);
        sfor<0, Q>([&](auto Tt) {
          const T v = buf[slot(kb, x_out(kp, a1, Tt))];
          if constexpr (PART) r[J * Q + Tt].y = v;
          else r[J * Q + Tt].x = v;
        });
      });
    }
    // output index k of register J*Q + t after the last forward stage
    RF_HD RF_INL static int k_of(int lane, int J, int t) {
      int kb, kp, a1;
      group(lane, J, kb, kp, a1);
      return kb + V * x_out(kp, a1, t);
    }
  };
  using Last = Stage<M - 1>;

  // ---------------- whole transforms; `bar` synchronises the lanes
  // forward: registers in F1 layout -> registers in the last stage's layout
  template <bool HALF_IN, class Bar>
  RF_HD RF_INL static void forward(C* r, C* lds, int lane, const C* tw, Bar&& bar) {
    int ln = opaque(lane);
    f1_fwd<HALF_IN>(r, ln, tw);
    bar();
    put_f1(lds, r, ln);
    bar();
    ln = opaque(lane);
    Stage<0>::get_in(lds, r, ln);
    Stage<0>::fwd(r, ln, tw);
    sfor<1, M>([&](auto D) {
      bar();
      Stage<D - 1>::put_out(lds, r, ln);
      bar();
      ln = opaque(lane);
      Stage<D>::get_in(lds, r, ln);
      Stage<D>::fwd(r, ln, tw);
    });
  }
  // inverse: registers in the last stage's layout -> F1 layout
  template <bool HALF_OUT, class Bar>
  RF_HD RF_INL static void inverse(C* r, C* lds, int lane, const C* tw, Bar&& bar) {
    int ln = opaque(lane);
    sfor<0, M - 1>([&](auto I) {
      constexpr int D = M - 1 - I;
      Stage<D>::inv(r, ln, tw);
      bar();
      Stage<D>::put_in(lds, r, ln);
      bar();
      ln = opaque(lane);
      Stage<D - 1>::get_out(lds, r, ln);
    });
    Stage<0>::inv(r, ln, tw);
    bar();
    Stage<0>::put_in(lds, r, ln);
    bar();
    ln = opaque(lane);
    get_f1(lds, r, ln);
    f1_inv<HALF_OUT>(r, ln, tw);
  }

  // ---------------- the same transforms with split exchanges
  // Each exchange moves the real parts of all values, then the imaginary
  // parts, through a buffer of kSlots values of T: half the LDS of a
  // whole-value exchange, twice the accesses at half the width.  Same
  // arithmetic and slots, so the results equal forward / inverse bit for bit.
  template <int PART>
  RF_HD RF_INL static void put_f1_part(T* buf, const C* r, int u) {
    sfor<0, V>([&](auto KB) { buf[KB + (V + 1) * u] = PART ? r[KB].y : r[KB].x; });
  }
  template <int PART>
  RF_HD RF_INL static void get_f1_part(const T* buf, C* r, int u) {
    sfor<0, V>([&](auto KB) {
      const T v = buf[KB + (V + 1) * u];
      if constexpr (PART) r[KB].y = v;
      else r[KB].x = v;
    });
  }
  template <bool HALF_IN, class Bar>
  RF_HD RF_INL static void forward_split(C* r, T* buf, int lane, const C* tw, Bar&& bar) {
    int ln = opaque(lane);
    f1_fwd<HALF_IN>(r, ln, tw);
    sfor<0, 2>([&](auto Pt) {
      bar();
      put_f1_part<Pt>(buf, r, ln);
      bar();
      Stage<0>::template get_in_part<Pt>(buf, r, ln);
    });
    ln = opaque(lane);
    Stage<0>::fwd(r, ln, tw);
    sfor<1, M>([&](auto D) {
      sfor<0, 2>([&](auto Pt) {
        bar();
        Stage<D - 1>::template put_out_part<Pt>(buf, r, ln);
        bar();
        Stage<D>::template get_in_part<Pt>(buf, r, ln);
      });
      ln = opaque(lane);
      Stage<D>::fwd(r, ln, tw);
    });
  }
  template <bool HALF_OUT, class Bar>
  RF_HD RF_INL static void inverse_split(C* r, T* buf, int lane, const C* tw, Bar&& bar) {
    int ln = opaque(lane);
    sfor<0, M - 1>([&](auto I) {
      constexpr int D = M - 1 - I;
      Stage<D>::inv(r, ln, tw);
      sfor<0, 2>([&](auto Pt) {
        bar();
        Stage<D>::template put_in_part<Pt>(buf, r, ln);
        bar();
        Stage<D - 1>::template get_out_part<Pt>(buf, r, ln);
      });
      ln = opaque(lane);
    });
    Stage<0>::inv(r, ln, tw);
    sfor<0, 2>([&](auto Pt) {
      bar();
      Stage<0>::template put_in_part<Pt>(buf, r, ln);
      bar();
      get_f1_part<Pt>(buf, r, ln);
    });
    ln = opaque(lane);
    f1_inv<HALF_OUT>(r, ln, tw);
  }
};

}  // namespace rfft
