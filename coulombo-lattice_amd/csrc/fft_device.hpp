// Device-side building blocks of the engine's own mixed-radix FFT:
// complex arithmetic, radix 2/3/4/5/7/8 butterflies and one LDS-resident
// Stockham autosort stage with register staging.
//
// A tile of `nl` lines of length L lives in LDS in one of two layouts:
//   line-major  (LM = true):  element (line, i) at line*L + i
//   index-major (LM = false): element (line, i) at i*nl + line
// Butterflies are assigned to threads so that consecutive lanes touch
// consecutive LDS words in either layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clat {

constexpr int NT = 256;  // threads per block of every FFT kernel

template <typename T> struct C2T;
template <> struct C2T<float> { using type = float2; };
template <> struct C2T<double> { using type = double2; };
template <typename T> using C2 = typename C2T<T>::type;

template <typename T>
__device__ __forceinline__ C2<T> mk(T x, T y) { C2<T> r; r.x = x; r.y = y; return r; }
template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ C2<T> cmulc(C2<T> a, C2<T> b) {
  return mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ C2<T> mul_mi(C2<T> a) {
  return INV ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

// Exact integer division by a runtime divisor d with e*d < 2^32:
// magic = ceil(2^32 / d) (0 encodes d == 1).
__device__ __forceinline__ uint32_t fdiv(uint32_t e, uint32_t magic) {
  return magic ? __umulhi(e, magic) : e;
}

// ---- butterflies (forward kernel e^{-2 pi i q r / R}; INV conjugates) ----

template <typename T, bool INV>
__device__ __forceinline__ void bfly2(C2<T>* v) {
  C2<T> a = v[0], b = v[1];
  v[0] = cadd<T>(a, b);
  v[1] = csub<T>(a, b);
}

template <typename T, bool INV>
__device__ __forceinline__ void bfly4(C2<T>* v) {
  C2<T> s02 = cadd<T>(v[0], v[2]), d02 = csub<T>(v[0], v[2]);
  C2<T> s13 = cadd<T>(v[1], v[3]), d13 = mul_mi<T, INV>(csub<T>(v[1], v[3]));
  v[0] = cadd<T>(s02, s13);
  v[2] = csub<T>(s02, s13);
  v[1] = cadd<T>(d02, d13);
  v[3] = csub<T>(d02, d13);
}

template <typename T, bool INV>
__device__ __forceinline__ void bfly3(C2<T>* v) {
  const T s3 = (T)0.86602540378443864676;  // sin(2 pi / 3)
  C2<T> t = cadd<T>(v[1], v[2]);
  C2<T> m = csub<T>(v[0], cscale<T>(t, (T)0.5));
  C2<T> s = mul_mi<T, INV>(cscale<T>(csub<T>(v[1], v[2]), s3));
  v[0] = cadd<T>(v[0], t);
  v[1] = cadd<T>(m, s);
  v[2] = csub<T>(m, s);
}

template <typename T, bool INV>
__device__ __forceinline__ void bfly5(C2<T>* v) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  C2<T> t1 = cadd<T>(v[1], v[4]), t2 = cadd<T>(v[2], v[3]);
  C2<T> d1 = csub<T>(v[1], v[4]), d2 = csub<T>(v[2], v[3]);
  C2<T> a0 = v[0];
  C2<T> m1 = mk<T>(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
  C2<T> m2 = mk<T>(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
  C2<T> n1 = mul_mi<T, INV>(mk<T>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
  C2<T> n2 = mul_mi<T, INV>(mk<T>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
  v[0] = mk<T>(a0.x + t1.x + t2.x, a0.y + t1.y + t2.y);
  v[1] = cadd<T>(m1, n1);
  v[4] = csub<T>(m1, n1);
  v[2] = cadd<T>(m2, n2);
  v[3] = csub<T>(m2, n2);
}

template <typename T, bool INV>
__device__ __forceinline__ void bfly7(C2<T>* v) {
  // cos(2 pi k / 7), sin(2 pi k / 7), k = 1..3
  const T c[3] = {(T)0.62348980185873353053, (T)-0.22252093395631440429,
                  (T)-0.90096886790241912624};
  const T s[3] = {(T)0.78183148246802980871, (T)0.97492791218182360702,
                  (T)0.43388373911755812048};
  C2<T> t[3], d[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    t[r] = cadd<T>(v[1 + r], v[6 - r]);
    d[r] = csub<T>(v[1 + r], v[6 - r]);
  }
  C2<T> a0 = v[0];
  v[0] = mk<T>(a0.x + t[0].x + t[1].x + t[2].x, a0.y + t[0].y + t[1].y + t[2].y);
#pragma unroll
  for (int q = 1; q <= 3; ++q) {
    C2<T> m = a0, n = mk<T>((T)0, (T)0);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int idx = ((r + 1) * q) % 7;  // angle index in [0, 7)
      const T cc = idx <= 3 ? c[(idx == 0 ? 1 : idx) - 1] : c[7 - idx - 1];
      const T ss = idx <= 3 ? s[(idx == 0 ? 1 : idx) - 1] : -s[7 - idx - 1];
      m.x += cc * t[r].x; m.y += cc * t[r].y;
      n.x += ss * d[r].x; n.y += ss * d[r].y;
    }
    n = mul_mi<T, INV>(n);
    v[q] = cadd<T>(m, n);
    v[7 - q] = csub<T>(m, n);
  }
}

template <typename T, bool INV>
__device__ __forceinline__ void bfly8(C2<T>* v) {
  const T h = (T)0.70710678118654752440;
  C2<T> b[4], c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    b[k] = cadd<T>(v[k], v[k + 4]);
    c[k] = csub<T>(v[k], v[k + 4]);
  }
  // c[k] *= W8^k (forward W8 = e^{-i pi/4}); inverse conjugates
  c[1] = INV ? mk<T>(h * (c[1].x - c[1].y), h * (c[1].x + c[1].y))
             : mk<T>(h * (c[1].x + c[1].y), h * (c[1].y - c[1].x));
  c[2] = mul_mi<T, INV>(c[2]);
  c[3] = INV ? mk<T>(-h * (c[3].x + c[3].y), h * (c[3].x - c[3].y))
             : mk<T>(h * (c[3].y - c[3].x), -h * (c[3].x + c[3].y));
  bfly4<T, INV>(b);
  bfly4<T, INV>(c);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    v[2 * m] = b[m];
    v[2 * m + 1] = c[m];
  }
}

template <int R, typename T, bool INV>
__device__ __forceinline__ void bfly(C2<T>* v) {
  if constexpr (R == 2) bfly2<T, INV>(v);
  else if constexpr (R == 3) bfly3<T, INV>(v);
  else if constexpr (R == 4) bfly4<T, INV>(v);
  else if constexpr (R == 5) bfly5<T, INV>(v);
  else if constexpr (R == 7) bfly7<T, INV>(v);
  else bfly8<T, INV>(v);
}

// ---- plan of one axis -----------------------------------------------------

constexpr int MAX_STAGES = 14;

struct AxisPlan {
  int L;                       // transform length
  int nst;                     // number of stages
  int radix[MAX_STAGES];
  int ns[MAX_STAGES];          // product of the radices before the stage
  uint32_t ns_magic[MAX_STAGES];
  uint32_t lr_magic[MAX_STAGES];  // divisor L / radix
  uint32_t l_magic;            // divisor L
  uint32_t l2_magic;           // divisor 2L
};

// One Stockham stage of radix R over nl lines held in LDS.  Twiddles
// W_L^m = e^{-2 pi i m / L} come from the LDS table tw[0..L).
template <int R, bool INV, int EMAX, bool LM, typename T>
__device__ __forceinline__ void fft_stage(C2<T>* __restrict__ buf,
                                          const C2<T>* __restrict__ tw,
                                          int nl, int lgnl, int L, int Ns,
                                          uint32_t ns_magic, uint32_t lr_magic) {
  constexpr int MB = (EMAX + R - 1) / R;
  const int LR = L / R;
  const int nb = nl * LR;
  const int tws = LR / Ns;  // = L / (Ns * R)
  C2<T> v[MB][R];
  int jv[MB], lv[MB];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = (int)threadIdx.x + m * NT;
    if (b < nb) {
      int line, j;
      if (LM) {
        line = (int)fdiv((uint32_t)b, lr_magic);
        j = b - line * LR;
      } else {
        line = b & (nl - 1);
        j = b >> lgnl;
      }
      const int k = j - (int)fdiv((uint32_t)j, ns_magic) * Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = j + r * LR;
        v[m][r] = buf[LM ? line * L + i : i * nl + line];
      }
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const C2<T> w = tw[r * k * tws];
          v[m][r] = INV ? cmulc<T>(v[m][r], w) : cmul<T>(v[m][r], w);
        }
      }
      bfly<R, T, INV>(v[m]);
      jv[m] = j;
      lv[m] = line;
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int b = (int)threadIdx.x + m * NT;
    if (b < nb) {
      const int j = jv[m], line = lv[m];
      const int q = (int)fdiv((uint32_t)j, ns_magic);
      const int d = q * Ns * R + (j - q * Ns);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = d + r * Ns;
        buf[LM ? line * L + i : i * nl + line] = v[m][r];
      }
    }
  }
  __syncthreads();
}

// Full in-place LDS FFT of nl lines (nl a power of two = 1 << lgnl).
// The caller must __syncthreads() after filling buf and tw.
template <bool INV, int EMAX, bool LM, typename T>
__device__ void lds_fft(C2<T>* __restrict__ buf, const C2<T>* __restrict__ tw,
                        const AxisPlan& p, int nl, int lgnl) {
  for (int s = 0; s < p.nst; ++s) {
    const int Ns = p.ns[s];
    switch (p.radix[s]) {
      case 2: fft_stage<2, INV, EMAX, LM, T>(buf, tw, nl, lgnl, p.L, Ns, p.ns_magic[s], p.lr_magic[s]); break;
      case 3: fft_stage<3, INV, EMAX, LM, T>(buf, tw, nl, lgnl, p.L, Ns, p.ns_magic[s], p.lr_magic[s]); break;
      case 4: fft_stage<4, INV, EMAX, LM, T>(buf, tw, nl, lgnl, p.L, Ns, p.ns_magic[s], p.lr_magic[s]); break;
      case 5: fft_stage<5, INV, EMAX, LM, T>(buf, tw, nl, lgnl, p.L, Ns, p.ns_magic[s], p.lr_magic[s]); break;
      case 7: fft_stage<7, INV, EMAX, LM, T>(buf, tw, nl, lgnl, p.L, Ns, p.ns_magic[s], p.lr_magic[s]); break;
      default: fft_stage<8, INV, EMAX, LM, T>(buf, tw, nl, lgnl, p.L, Ns, p.ns_magic[s], p.lr_magic[s]); break;
    }
  }
}

}  // namespace clat
