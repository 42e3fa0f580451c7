// Two-level line transforms: L = R1 * R2 with both factors register-resident
// DFTs and a single LDS exchange per transform (the pass kernels of prepare()).
//
//   n = n1 + R1 n2, k = R2 k1 + k2
//   forward:  X[R2 k1 + k2] = sum_n1 W_R1^{n1 k1} W_L^{n1 k2} sum_n2 x[n1 + R1 n2] W_R2^{n2 k2}
//   inverse:  x[n1 + R1 n2] = sum_k2 W_R2^{-n2 k2} W_L^{-n1 k2} sum_k1 X[R2 k1 + k2] W_R1^{-n1 k1}
//
// Work items: phase P1/P3 = (line c, n1) holding R2 values; phase P2/P0 =
// (line c, sub-line s, k2) holding R1 values.  Sub-lines (nsub = 2) are the
// two half-shift parities of the shared schedule; nsub = 1 is one round.
//
// Column mode (LINE = false): a block owns C consecutive columns (the
// contiguous index) of one outer index o = blockIdx.y; lanes run along c.
// Line mode (LINE = true): a block owns C contiguous lines; lanes run along
// the line.  Line mode can read the pair density straight from the atom sites
// (fused scatter, forward) and write the potential straight to them (fused
// gather, inverse).
#pragma once

#include "regfft.hpp"

namespace clat {

enum PassMode { PM_FWD = 0, PM_CONV = 1, PM_INV = 2, PM_CONV2 = 3 };  // PM_CONV2: coupled two-coset conv (column kernels only)

struct P2Args {
  int L, R1, R2, R1p;
  uint32_t mR1, mR2, mL;
  int nsub;          // 1 or 2
  int C, lgC;        // lines / columns per block (power of two)
  int W;             // number of columns (column mode) or lines (line mode)
  long long in_os, in_is, out_os, out_is;
  // rounds schedule phase of an x pass: (kx ? h[n] : 1) * hy[y]^ky * hz[z]^kz, line = y + Yl*z
  int rphase, kx, ky, kz, Yl;
  // kernel-table lookup of the conv pass: Ĝ stored as [r(fy)][r(fz)][fx],
  // fx over the whole 2X range (x reflection folded in), pitch X2, rows GZ1
  int zpar, gxm, gxa, gym, gya, X2, Y2, Z2, GZ1;
  int acc;           // inverse line pass: add to the output instead of writing it
  // atom sites of line mode: CSR over lines, internal linear cell of each site
  const int32_t* site_ptr;
  const int32_t* site_cell;
  const double2* site_in;   // forward: pair density at the sites (replaces `in`)
  double2* site_out;        // inverse: potential at the sites (replaces `out`)
  // rows without sites (column kernel only; the general kernel rejects it)
  const uint32_t* rowmask;
  int mask_os;
  // batched prepares (column kernel only): count and element strides
  int nb;
  long long in_bs, out_bs;
  // column kernels only: element stride between blocks of C columns (0: C, row-major plane)
  long long in_cbs, out_cbs;
  int outer_x;  // column kernels only: outer index on blockIdx.x
};

// register DFT of size R: compile-time when RT > 0, else dispatched on R
template <int RT, bool INV, typename T>
__device__ __forceinline__ void rdft_r(int R, C2<T>* v) {
  if constexpr (RT > 0) rdft<RT, INV, T>(v);
  else rdft_dispatch<INV, T>(R, v);
}

template <typename TA, typename TB>
__device__ __forceinline__ C2<TB> cvt(C2<TA> a) {
  return mk<TB>((TB)a.x, (TB)a.y);
}

// R1T, R2T: compile-time radices (0 = read a.R1 / a.R2 at run time and
// dispatch the register DFTs with a switch; only for uncommon lengths, the
// switch makes the kernel several times larger).
template <typename TI, typename TC, typename TO, int MODE, bool LINE, int R1T, int R2T>
__global__ __launch_bounds__(256) void k_pass2(const C2<TI>* in, C2<TO>* out, P2Args a,
                                               const C2<TC>* __restrict__ tw, const C2<TC>* __restrict__ h,
                                               const C2<TC>* __restrict__ hy, const C2<TC>* __restrict__ hz,
                                               const TC* __restrict__ G) {
  using CT = C2<TC>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CT* lds = reinterpret_cast<CT*>(smem_raw);
  const int NTH = blockDim.x;
  const int tid = threadIdx.x;
  constexpr int M1 = R1T ? R1T : RMAX;  // register array extents
  constexpr int M2 = R2T ? R2T : RMAX;
  const int L = a.L, nsub = a.nsub, C = a.C, lgC = a.lgC;
  const int R1 = R1T ? R1T : a.R1;
  const int R2 = R2T ? R2T : a.R2;
  const int base = blockIdx.x * C;  // first column / line of the block
  const int nC = min(C, a.W - base);
  const int o = LINE ? 0 : blockIdx.y;
  const CT zero = mk<TC>((TC)0, (TC)0);

  auto lidx = [&](int c, int s, int k2, int n1) -> int {
    return LINE ? ((c * nsub + s) * R2 + k2) * a.R1p + n1 : ((((s * R2 + k2) * R1 + n1)) << lgC) + c;
  };
  auto gin = [&](int c, int i) -> size_t {
    return LINE ? (size_t)(base + c) * a.in_os + (size_t)i * a.in_is
                : (size_t)o * a.in_os + (size_t)i * a.in_is + base + c;
  };
  auto gout = [&](int c, int i) -> size_t {
    return LINE ? (size_t)(base + c) * a.out_os + (size_t)i * a.out_is
                : (size_t)o * a.out_os + (size_t)i * a.out_is + base + c;
  };
  // rounds-schedule line factor hy[y]^ky hz[z]^kz of x-line base + c
  auto line_phase = [&](int c) -> CT {
    CT p = mk<TC>((TC)1, (TC)0);
    if (LINE && (a.ky | a.kz)) {
      const int l = base + min(c, nC - 1);
      const int z = l / a.Yl, y = l - z * a.Yl;
      if (a.ky) p = cmul<TC>(p, hy[y]);
      if (a.kz) p = cmul<TC>(p, hz[z]);
    }
    return p;
  };
  const int nP13 = C * R1;
  const int nP02 = C * nsub * R2;

  // ---------------------------------------------------------------- P1
  if (MODE != PM_INV) {
    const bool from_sites = LINE && a.site_in != nullptr;
    if (from_sites) {
      // fused scatter: zero C lines in LDS, drop the site values in
      for (int e = tid; e < C * L; e += NTH) lds[e] = zero;
      __syncthreads();
      const int s0 = a.site_ptr[base], s1 = a.site_ptr[base + nC];
      for (int k = s0 + tid; k < s1; k += NTH) {
        const uint32_t cell = (uint32_t)a.site_cell[k];
        const int l = (int)fdiv(cell, a.mL);
        const int x = (int)cell - l * L;
        const double2 v = a.site_in[k];
        lds[(l - base) * L + x] = mk<TC>((TC)v.x, (TC)v.y);
      }
      __syncthreads();
    }
    for (int w0 = 0; w0 < nP13; w0 += NTH) {
      const int w = w0 + tid;
      const bool has = w < nP13;
      int c = 0, n1 = 0;
      if (has) {
        if (LINE) { c = R1T ? w / (R1T ? R1T : 1) : (int)fdiv((uint32_t)w, a.mR1); n1 = w - c * R1; }
        else { c = w & (C - 1); n1 = w >> lgC; }
      }
      const bool ok = has && c < nC;
      CT x[M2];
#pragma unroll
      for (int n2 = 0; n2 < M2; ++n2) {
        if (n2 < R2) {
          const int n = n1 + R1 * n2;
          if (from_sites) x[n2] = has ? lds[c * L + n] : zero;
          else x[n2] = ok ? cvt<TI, TC>(in[gin(c, n)]) : zero;
        }
      }
      if (from_sites) __syncthreads();  // the line buffer and the exchange share LDS
      if (has) {
        const CT lp = (LINE && a.rphase) ? line_phase(c) : mk<TC>((TC)1, (TC)0);
        for (int s = 0; s < nsub; ++s) {
          CT v[M2];
#pragma unroll
          for (int n2 = 0; n2 < M2; ++n2) {
            if (n2 < R2) {
              const int n = n1 + R1 * n2;
              CT t = x[n2];
              if (nsub == 2) {
                if (s == 1) t = cmul<TC>(t, h[n]);
              } else if (LINE && a.rphase) {
                CT ph = a.kx ? cmul<TC>(h[n], lp) : lp;
                t = cmul<TC>(t, ph);
              }
              v[n2] = t;
            }
          }
          rdft_r<R2T, false, TC>(R2, v);
#pragma unroll
          for (int k2 = 0; k2 < M2; ++k2) {
            if (k2 < R2) {
              CT t = v[k2];
              if (n1 * k2) t = cmul<TC>(t, tw[n1 * k2]);
              lds[lidx(c, s, k2, n1)] = t;
            }
          }
        }
      }
    }
    __syncthreads();
  }

  // ---------------------------------------------------------------- P2 / P0
  {
    int fy = 0, ry = 0;
    if (MODE == PM_CONV) {
      fy = o * a.gym + a.gya;
      ry = min(fy, a.Y2 - fy);
    }
    for (int w = tid; w < nP02; w += NTH) {
      int c, s, k2;
      if (LINE) {
        s = w & (nsub - 1);
        const int t = w >> (nsub - 1);
        c = R2T ? t / (R2T ? R2T : 1) : (int)fdiv((uint32_t)t, a.mR2);
        k2 = t - c * R2;
      } else {
        c = w & (C - 1);
        const int r = w >> lgC;
        s = r & (nsub - 1);
        k2 = r >> (nsub - 1);
      }
      const bool ok = c < nC;
      CT v[M1];
      if (MODE == PM_INV) {
#pragma unroll
        for (int k1 = 0; k1 < M1; ++k1)
          if (k1 < R1) v[k1] = ok ? cvt<TI, TC>(in[gin(c, (R2 * k1 + k2) * nsub + s)]) : zero;
      } else {
#pragma unroll
        for (int n1 = 0; n1 < M1; ++n1)
          if (n1 < R1) v[n1] = lds[lidx(c, s, k2, n1)];
        rdft_r<R1T, false, TC>(R1, v);  // v[k1] = X[R2 k1 + k2]
      }
      if (MODE == PM_FWD) {
        if (ok) {
#pragma unroll
          for (int k1 = 0; k1 < M1; ++k1)
            if (k1 < R1) out[gout(c, (R2 * k1 + k2) * nsub + s)] = cvt<TC, TO>(v[k1]);
        }
        continue;
      }
      if (MODE == PM_CONV) {
        const int fx = (base + min(c, nC - 1)) * a.gxm + a.gxa;
        const int zp = (nsub == 2) ? s : a.zpar;
        const TC* Grow = G + (size_t)ry * a.GZ1 * a.X2 + fx;
#pragma unroll
        for (int k1 = 0; k1 < M1; ++k1) {
          if (k1 < R1) {
            const int fz = 2 * (R2 * k1 + k2) + zp;
            const int rz = min(fz, a.Z2 - fz);
            const TC g = Grow[(size_t)rz * a.X2];
            v[k1] = cscale<TC>(v[k1], g);
          }
        }
      }
      rdft_r<R1T, true, TC>(R1, v);  // v[n1] = sum_k1 X W_R1^{-n1 k1}
#pragma unroll
      for (int n1 = 0; n1 < M1; ++n1) {
        if (n1 < R1) {
          CT t = v[n1];
          if (n1 * k2) t = cmulc<TC>(t, tw[n1 * k2]);
          lds[lidx(c, s, k2, n1)] = t;
        }
      }
    }
  }
  if (MODE == PM_FWD) return;
  __syncthreads();

  // ---------------------------------------------------------------- P3
  const bool to_sites = LINE && a.site_out != nullptr;
  for (int w0 = 0; w0 < nP13; w0 += NTH) {
    const int w = w0 + tid;
    const bool has = w < nP13;
    int c = 0, n1 = 0;
    if (has) {
      if (LINE) { c = R1T ? w / (R1T ? R1T : 1) : (int)fdiv((uint32_t)w, a.mR1); n1 = w - c * R1; }
      else { c = w & (C - 1); n1 = w >> lgC; }
    }
    const bool ok = has && c < nC;
    CT acc[M2];
#pragma unroll
    for (int n2 = 0; n2 < M2; ++n2) acc[n2] = zero;
    if (has) {
      const CT lp = (LINE && a.rphase) ? line_phase(c) : mk<TC>((TC)1, (TC)0);
      for (int s = 0; s < nsub; ++s) {
        CT v[M2];
#pragma unroll
        for (int k2 = 0; k2 < M2; ++k2)
          if (k2 < R2) v[k2] = lds[lidx(c, s, k2, n1)];
        rdft_r<R2T, true, TC>(R2, v);  // v[n2] = x_s[n1 + R1 n2]
#pragma unroll
        for (int n2 = 0; n2 < M2; ++n2) {
          if (n2 < R2) {
            const int n = n1 + R1 * n2;
            CT t = v[n2];
            if (nsub == 2) {
              if (s == 1) t = cmulc<TC>(t, h[n]);
            } else if (LINE && a.rphase) {
              CT ph = a.kx ? cmul<TC>(h[n], lp) : lp;
              t = cmulc<TC>(t, ph);
            }
            acc[n2] = cadd<TC>(acc[n2], t);
          }
        }
      }
    }
    if (to_sites) {
      // fused gather: the line goes back to LDS, the sites pick their value
      __syncthreads();
      if (has) {
#pragma unroll
        for (int n2 = 0; n2 < M2; ++n2)
          if (n2 < R2) lds[c * L + n1 + R1 * n2] = acc[n2];
      }
      __syncthreads();
      const int s0 = a.site_ptr[base], s1 = a.site_ptr[base + nC];
      for (int k = s0 + tid; k < s1; k += NTH) {
        const uint32_t cell = (uint32_t)a.site_cell[k];
        const int l = (int)fdiv(cell, a.mL);
        const int x = (int)cell - l * L;
        const CT v = lds[(l - base) * L + x];
        double2 r = make_double2((double)v.x, (double)v.y);
        if (a.acc) {
          const double2 p = a.site_out[k];
          r.x += p.x;
          r.y += p.y;
        }
        a.site_out[k] = r;
      }
    } else if (ok) {
#pragma unroll
      for (int n2 = 0; n2 < M2; ++n2) {
        if (n2 < R2) {
          const size_t g = gout(c, n1 + R1 * n2);
          C2<TO> r = cvt<TC, TO>(acc[n2]);
          if (a.acc) r = cadd<TO>(out[g], r);
          out[g] = r;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launch descriptor and the switch over the (types, mode, layout) triples the
// engine uses; instantiated once per compile-time radix pair.
// ---------------------------------------------------------------------------
enum P2Combo { CB_DDD = 0, CB_FFF, CB_DFF, CB_FFD, CB_DDF, CB_FDF, CB_FDD };

struct P2Launch {
  dim3 grid, block;
  size_t lds;
  hipStream_t stream;
  const void* in;
  void* out;
  P2Args a;
  const void *tw, *h, *hy, *hz, *G;
};

template <typename TI, typename TC, typename TO, int MODE, bool LINE, int R1T, int R2T>
hipError_t p2_go(const P2Launch& L) {
  hipLaunchKernelGGL((k_pass2<TI, TC, TO, MODE, LINE, R1T, R2T>), L.grid, L.block, L.lds, L.stream,
                     (const C2<TI>*)L.in, (C2<TO>*)L.out, L.a, (const C2<TC>*)L.tw, (const C2<TC>*)L.h,
                     (const C2<TC>*)L.hy, (const C2<TC>*)L.hz, (const TC*)L.G);
  return hipGetLastError();
}

// returns hipErrorInvalidValue for a triple that is not instantiated
template <int R1T, int R2T>
hipError_t p2_launch(int combo, int mode, bool line, const P2Launch& L) {
  using d = double;
  using f = float;
  switch (combo) {
    case CB_DDD:
      if (mode == PM_FWD) return line ? p2_go<d, d, d, PM_FWD, true, R1T, R2T>(L) : p2_go<d, d, d, PM_FWD, false, R1T, R2T>(L);
      if (mode == PM_CONV && !line) return p2_go<d, d, d, PM_CONV, false, R1T, R2T>(L);
      if (mode == PM_INV) return line ? p2_go<d, d, d, PM_INV, true, R1T, R2T>(L) : p2_go<d, d, d, PM_INV, false, R1T, R2T>(L);
      break;
    case CB_FFF:
      if (mode == PM_FWD) return line ? p2_go<f, f, f, PM_FWD, true, R1T, R2T>(L) : p2_go<f, f, f, PM_FWD, false, R1T, R2T>(L);
      if (mode == PM_CONV && !line) return p2_go<f, f, f, PM_CONV, false, R1T, R2T>(L);
      if (mode == PM_INV) return line ? p2_go<f, f, f, PM_INV, true, R1T, R2T>(L) : p2_go<f, f, f, PM_INV, false, R1T, R2T>(L);
      break;
    case CB_DFF:
      if (mode == PM_FWD && line) return p2_go<d, f, f, PM_FWD, true, R1T, R2T>(L);
      break;
    case CB_FFD:
      if (mode == PM_INV && line) return p2_go<f, f, d, PM_INV, true, R1T, R2T>(L);
      break;
    case CB_DDF:
      if (mode == PM_FWD && line) return p2_go<d, d, f, PM_FWD, true, R1T, R2T>(L);
      break;
    case CB_FDF:
      if (mode == PM_FWD && !line) return p2_go<f, d, f, PM_FWD, false, R1T, R2T>(L);
      if (mode == PM_CONV && !line) return p2_go<f, d, f, PM_CONV, false, R1T, R2T>(L);
      break;
    case CB_FDD:
      if (mode == PM_INV) return line ? p2_go<f, d, d, PM_INV, true, R1T, R2T>(L) : p2_go<f, d, d, PM_INV, false, R1T, R2T>(L);
      break;
    default:
      break;
  }
  return hipErrorInvalidValue;
}

// compile-time radix pairs with their own code objects (pass2_pair.hip)
#define CLAT_P2_PAIRS(X)                                                                 \
  X(16, 16) X(16, 8) X(8, 4) X(8, 8) X(9, 5) X(12, 8) X(12, 10) X(16, 10) X(16, 12) X(20, 20) \
  X(6, 6) X(8, 5) X(8, 6) X(10, 5) X(9, 6) X(10, 6) X(9, 8) X(15, 5) \
  X(10, 8) X(9, 9) X(10, 9) X(10, 10) X(12, 9) X(15, 9) X(12, 12) X(15, 10) \
  X(18, 9) X(15, 12) X(20, 10) X(18, 12) X(15, 15) X(16, 15) X(18, 15) X(18, 16) \
  X(20, 15) X(20, 16) X(18, 18) X(20, 18)
#define CLAT_P2_DECL(a, b) hipError_t clat_p2_##a##x##b(int combo, int mode, bool line, const P2Launch& L);
CLAT_P2_PAIRS(CLAT_P2_DECL)

}  // namespace clat
