// Register-resident DFTs of compile-time size R (1 <= R <= 20, factors 2,3,5,7)
// for one work item.  Composite sizes use one split R = P*M.  When P and M are
// coprime (6, 10, 12, 14, 15, 20) it is the prime-factor (Good-Thomas) split,
// which needs no twiddles: with n = (M n1 + P n2) mod R and
// k = (M (M^-1 mod P) k1 + P (P^-1 mod M) k2) mod R,
//   X[k] = sum_{n1<P} W_P^{n1 k1} sum_{n2<M} x[n] W_M^{n2 k2}
// (both index maps are compile-time register renamings).  Otherwise (8 is a
// base size; 9, 16, 18) one Cooley-Tukey split:
//   X[M k1 + k2] = sum_{n1<P} W_P^{n1 k1} W_R^{n1 k2} sum_{n2<M} x[n1 + P n2] W_M^{n2 k2}
// with the twiddles W_R^{e} folded to compile-time constants (the quarter
// turns as swaps and sign changes).
#pragma once

#include "fft_device.hpp"

namespace clat {

// ---- compile-time trigonometry: cos/sin(2 pi e / R) ------------------------
constexpr double kHalfPi = 1.57079632679489661923132169163975144;

constexpr double taylor_sin(double x) {
  double x2 = x * x, t = x, s = x;
  for (int i = 1; i < 14; ++i) {
    t *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
    s += t;
  }
  return s;
}
constexpr double taylor_cos(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int i = 1; i < 14; ++i) {
    t *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
    s += t;
  }
  return s;
}
// angle 2 pi e / R, reduced exactly with integers to [0, pi/4]
constexpr void cossin2pi(int e, int R, double& c, double& s) {
  e = ((e % R) + R) % R;
  const int q = (4 * e) / R;       // quadrant
  const int r = 4 * e - q * R;     // angle within quadrant = (pi/2) r / R
  double cf = 0.0, sf = 0.0;
  if (2 * r <= R) {
    const double phi = kHalfPi * (double)r / (double)R;
    cf = taylor_cos(phi);
    sf = taylor_sin(phi);
  } else {
    const double phi = kHalfPi * (double)(R - r) / (double)R;
    cf = taylor_sin(phi);
    sf = taylor_cos(phi);
  }
  switch (q) {
    case 0: c = cf; s = sf; break;
    case 1: c = -sf; s = cf; break;
    case 2: c = -cf; s = -sf; break;
    default: c = sf; s = -cf; break;
  }
}

template <int R>
struct TwTable {
  double c[R > 0 ? R : 1];
  double s[R > 0 ? R : 1];
  constexpr TwTable() : c(), s() {
    for (int e = 0; e < R; ++e) cossin2pi(e, R, c[e], s[e]);
  }
};

// ---- size decomposition -----------------------------------------------------
template <int R>
struct Split {  // first factor P (a base butterfly), rest M
  static constexpr int P = (R % 4 == 0 && R != 8) ? 4 : (R % 5 == 0) ? 5 : (R % 3 == 0) ? 3
                           : (R % 7 == 0) ? 7 : (R % 2 == 0) ? 2 : R;
  static constexpr int M = R / P;
};

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
// a^-1 mod m (a and m coprime; m = 1 gives 0)
constexpr int inv_mod_c(int a, int m) {
  for (int x = 0; x < m; ++x)
    if ((a * x) % m == 1 % m) return x;
  return 0;
}

template <int R, bool INV, typename T>
__device__ __forceinline__ void rdft(C2<T>* v);

template <int R, bool INV, typename T>
__device__ __forceinline__ void rdft_base(C2<T>* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) bfly2<T, INV>(v);
  else if constexpr (R == 3) bfly3<T, INV>(v);
  else if constexpr (R == 4) bfly4<T, INV>(v);
  else if constexpr (R == 5) bfly5<T, INV>(v);
  else if constexpr (R == 7) bfly7<T, INV>(v);
  else if constexpr (R == 8) bfly8<T, INV>(v);
}

template <int R, bool INV, typename T>
__device__ __forceinline__ void rdft(C2<T>* v) {
  constexpr int P = Split<R>::P, M = Split<R>::M;
  if constexpr (R <= 5 || R == 7 || R == 8) {
    rdft_base<R, INV, T>(v);
  } else if constexpr (gcd_c(P, M) == 1) {
    constexpr int MI = inv_mod_c(M, P), PI = inv_mod_c(P, M);
    C2<T> y[R];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) {
      C2<T> sub[M];
#pragma unroll
      for (int n2 = 0; n2 < M; ++n2) sub[n2] = v[(M * n1 + P * n2) % R];
      rdft<M, INV, T>(sub);
#pragma unroll
      for (int k2 = 0; k2 < M; ++k2) y[n1 * M + k2] = sub[k2];
    }
#pragma unroll
    for (int k2 = 0; k2 < M; ++k2) {
      C2<T> t[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) t[n1] = y[n1 * M + k2];
      rdft_base<P, INV, T>(t);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) v[(k1 * M * MI + k2 * P * PI) % R] = t[k1];
    }
  } else {
    constexpr TwTable<R> tw{};
    C2<T> y[R];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) {
      C2<T> sub[M];
#pragma unroll
      for (int n2 = 0; n2 < M; ++n2) sub[n2] = v[n1 + P * n2];
      rdft<M, INV, T>(sub);
#pragma unroll
      for (int k2 = 0; k2 < M; ++k2) {
        const int e = (n1 * k2) % R;
        const C2<T> a = sub[k2];
        if (e == 0) {
          y[n1 * M + k2] = a;
        } else if (2 * e == R) {
          y[n1 * M + k2] = mk<T>(-a.x, -a.y);
        } else if (4 * e == R) {  // W^{R/4} = -i (forward), +i (inverse)
          y[n1 * M + k2] = INV ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
        } else if (4 * e == 3 * R) {
          y[n1 * M + k2] = INV ? mk<T>(a.y, -a.x) : mk<T>(-a.y, a.x);
        } else {
          const T c = (T)tw.c[e];
          const T s = INV ? (T)tw.s[e] : (T)(-tw.s[e]);
          y[n1 * M + k2] = mk<T>(a.x * c - a.y * s, a.x * s + a.y * c);
        }
      }
    }
#pragma unroll
    for (int k2 = 0; k2 < M; ++k2) {
      C2<T> t[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) t[n1] = y[n1 * M + k2];
      rdft_base<P, INV, T>(t);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) v[M * k1 + k2] = t[k1];
    }
  }
}

// Supported register sizes and runtime dispatch on v[0..R)
constexpr int RMAX = 20;
__host__ __device__ constexpr bool reg_size_ok(int r) {
  return r == 1 || r == 2 || r == 3 || r == 4 || r == 5 || r == 6 || r == 7 || r == 8 || r == 9 || r == 10 ||
         r == 12 || r == 14 || r == 15 || r == 16 || r == 18 || r == 20;
}

template <bool INV, typename T>
__device__ __forceinline__ void rdft_dispatch(int R, C2<T>* v) {
  switch (R) {
    case 2: rdft<2, INV, T>(v); break;
    case 3: rdft<3, INV, T>(v); break;
    case 4: rdft<4, INV, T>(v); break;
    case 5: rdft<5, INV, T>(v); break;
    case 6: rdft<6, INV, T>(v); break;
    case 7: rdft<7, INV, T>(v); break;
    case 8: rdft<8, INV, T>(v); break;
    case 9: rdft<9, INV, T>(v); break;
    case 10: rdft<10, INV, T>(v); break;
    case 12: rdft<12, INV, T>(v); break;
    case 14: rdft<14, INV, T>(v); break;
    case 15: rdft<15, INV, T>(v); break;
    case 16: rdft<16, INV, T>(v); break;
    case 18: rdft<18, INV, T>(v); break;
    case 20: rdft<20, INV, T>(v); break;
    default: break;  // R == 1
  }
}

}  // namespace clat
