// x passes of the atom path (shared schedule) with every size fixed at
// compile time: radix pair R1 x R2, both half-shift parities (NSUB = 2) and C
// lines per block.  Only x-lines that hold at least one atom site are visited:
// a block takes C consecutive entries of the compact list of occupied lines.
//
//   forward: sites -> LDS line buffer -> two-level DFT of both parities ->
//            A[line][(R2 k1 + k2) NSUB + s]
//   inverse: D[line][...] -> two-level inverse DFT, parities combined ->
//            LDS line buffer -> potential at the sites
//
// Index map (L = R1 R2): n = n1 + R1 n2, k = R2 k1 + k2; the half-shift phase
// h[n] = W_2L^n is split into W_2L^{n1 (2 k2 + s)} (table T2L of 2L roots)
// and a compile-time rotation by W_2R2^n2, as in the column passes.
//
// Half-length lines (HL, table PH): a line of 2L points whose sites all have
// x = 2j + p, p = (y + z) mod 2 (the fcc sublattice of the grid), is carried
// as the L points j.  Its 2L frequencies below 2L of the 4L-point transform
// are F(f) = W_4L^{f p} sum_j u_j W_2L^{f j}: the L-point line's transform
// times a phase (table PH of 4L roots, W_4L^f), applied to the odd lines on
// the way out of the forward pass and conjugated on the way into the
// inverse one, whose first L outputs are then the potential at j.
#pragma once

#include "colpass.hpp"

namespace clat {

constexpr int LINE_NB = 16;  // prepares (or coset planes of prepares) one launch can carry

struct LineArgs {
  int n_live;                   // occupied lines
  const int32_t* lines;         // line id of each compact index
  const int32_t* live_ptr;      // CSR of the sites over compact lines (n_live + 1)
  const int32_t* site_ccell;    // compact cell of each site: compact line * L + x (half-length
                                // lines: compact line * 2L + x, whose half is compact line * L + j)
  // batched prepares (blockIdx.y = b < LINE_NB): site arrays and dense-side offset of each
  const double2* site_in[LINE_NB];  // forward: pair density at the sites
  double2* site_out[LINE_NB];       // inverse: potential at the sites
  // two real densities in one prepare (null when unused): the line carries
  // site_in[b].x + i site_in2[b].x, and since the kernel is real the potential
  // splits back into site_out[b] = (Re, 0) and site_out2[b] = (Im, 0)
  const double2* site_in2[LINE_NB];
  double2* site_out2[LINE_NB];
  // dense side (A forward, D inverse; line id = z Yl + y, 2X frequencies f per
  // line): element z lz + y ly + f in plain rows (lgsw = 0), or in strips of
  // sw = 2^lgsw frequencies, z lz + y ly + (f / sw) strip + f % sw
  // (byte offsets within one prepare's dense side stay below COL_RANGE)
  int Yl, lgsw;
  int lz, ly, strip;
  // frequencies carried on the dense side: f < fmax (the rest are not
  // stored by the forward pass and read as zero by the inverse one)
  int fmax;
  long long in_bs, out_bs;          // element stride between batched prepares on the dense side
};

// element offset of (line, first frequency) and of frequency f within the line
__device__ __forceinline__ int line_base(const LineArgs& a, int line) {
  const int z = line / a.Yl;
  return z * a.lz + (line - z * a.Yl) * a.ly;
}
// (y + z) mod 2 of a line: the x parity of its sites on half-length lines
__device__ __forceinline__ bool line_odd(const LineArgs& a, int line) {
  const int z = line / a.Yl;
  return ((line - z * a.Yl + z) & 1) != 0;
}
__device__ __forceinline__ int freq_off(const LineArgs& a, int f) {
  return a.lgsw ? (f >> a.lgsw) * a.strip + (f & ((1 << a.lgsw) - 1)) : f;
}
// offsets of frequencies f0 + k1 P, k1 < R1: linear in k1 when P is a multiple
// of the strip width (or in plain rows), else one split per frequency
template <int R1, int P>
__device__ __forceinline__ void freq_offs(const LineArgs& a, int f0, int* off) {
  if (a.lgsw == 0 || (P & ((1 << a.lgsw) - 1)) == 0) {
    const int base = freq_off(a, f0), step = a.lgsw ? (P >> a.lgsw) * a.strip : P;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) off[k1] = base + k1 * step;
  } else {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) off[k1] = freq_off(a, f0 + k1 * P);
  }
}

template <int C, int R1, int R2, int NSUB>
constexpr int line_threads() {
  constexpr int m = R2 > R1 ? C * NSUB * R2 : C * NSUB * R1;
  return (m + 63) / 64 * 64;
}
template <int R1>
constexpr int line_r1p() {
  return (R1 > 1 && R1 % 2 == 0) ? R1 + 1 : R1;
}
// waves per CU that blocks of C lines of the FP64 inverse pass over
// complex64 input can hold: by their LDS tile (160 KiB per CU) and by
// registers (about 160 VGPRs: 12 waves)
template <int R1, int R2, int NSUB>
constexpr int line_fdd_waves(int C) {
  const int lds = C * NSUB * R2 * line_r1p<R1>() * 16;
  const int wpb = (C * NSUB * (R1 > R2 ? R1 : R2) + 63) / 64;
  const int by_lds = (160 * 1024 / lds) * wpb;
  return by_lds < 12 / wpb * wpb ? by_lds : 12 / wpb * wpb;
}
// lines per block: about 256 threads (NSUB R2 items per line in one phase,
// NSUB R1 in the others), half as many lines when the arithmetic is FP64.
// The FP64 inverse pass over complex64 input halves them again where that
// fits more waves per CU (its FP64 LDS tile bounds the blocks, and its input
// loads alone do not keep the memory busy: the 256-point lines)
template <typename TI, typename TC, int R1, int R2, int NSUB>
constexpr int line_c() {
  constexpr int q0 = (256 / (NSUB * (R1 > R2 ? R1 : R2))) / (sizeof(TC) == 8 ? 2 : 1);
  constexpr int q = q0 >= 16 ? 16 : q0 >= 8 ? 8 : q0 >= 4 ? 4 : q0 >= 2 ? 2 : 1;
  if constexpr (sizeof(TC) == 8 && sizeof(TI) == 4 && q > 1)
    return line_fdd_waves<R1, R2, NSUB>(q / 2) > line_fdd_waves<R1, R2, NSUB>(q) ? q / 2 : q;
  return q;
}
template <typename TC, int R1, int R2, int NSUB, int C>
constexpr size_t line_lds() {
  return (size_t)C * NSUB * R2 * line_r1p<R1>() * sizeof(C2<TC>);
}

// LV: 0 plain lines; 1 half-length lines (HL); 2 the forward's stored
// frequencies f < L only (the half box on full lines, R1 even: outputs
// k1 >= R1 / 2 of the R1-point DFTs are never formed, the rest bit-identical)
template <typename TI, typename TC, typename TO, int MODE, int R1, int R2, int NSUB, int C, int LV>
// the complex64 forward held to six waves per SIMD (80 VGPRs) where that
// spills nothing (R1 <= 16, not on half-length lines): the 16 x 8 x forward
// of the 1M-atom grid from 89 VGPRs, 0.285-0.291 -> 0.271-0.274 ms per step
__attribute__((amdgpu_waves_per_eu((MODE == PM_FWD && sizeof(TC) == 4 && R1 <= 16 && LV != 1) ? 6 : 1)))
__global__ __launch_bounds__((line_threads<C, R1, R2, NSUB>())) void k_linepass(const C2<TI>* in, C2<TO>* out,
                                                                                LineArgs a,
                                                                                const C2<TC>* __restrict__ T2L,
                                                                                const C2<TC>* __restrict__ PH) {
  using CT = C2<TC>;
  constexpr bool HL = LV == 1;
  constexpr int K1N = (LV == 2 && MODE == PM_FWD && R1 % 2 == 0) ? R1 / 2 : R1;  // forward outputs formed
  constexpr int L = R1 * R2, R1P = line_r1p<R1>();
  constexpr int NTH = line_threads<C, R1, R2, NSUB>();
  constexpr uint32_t SC = sizeof(CT);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CT* lds = reinterpret_cast<CT*>(smem_raw);
  const int tid = threadIdx.x;
  const int b0 = blockIdx.x * C;
  const int nC = min(C, a.n_live - b0);
  const CT zero = mk<TC>((TC)0, (TC)0);
  const int s0 = a.live_ptr[b0], s1 = a.live_ptr[b0 + nC];
  const int cell0 = b0 * L;
  auto xi = [](int c, int s, int k2, int n1) { return ((c * NSUB + s) * R2 + k2) * R1P + n1; };


  // (c, s, n1) items of the phases holding R2 values: one parity each
  const int n1 = tid % R1, s13 = (tid / R1) % NSUB, c13 = tid / (R1 * NSUB);
  const bool ok13 = tid < C * NSUB * R1 && c13 < nC;
  // (c, s, k2) items of the phase holding R1 values
  const int s2 = tid & (NSUB - 1), k2 = (tid / NSUB) % R2, c2 = tid / (NSUB * R2);
  const bool ok2 = tid < C * NSUB * R2 && c2 < nC;

  if constexpr (MODE == PM_FWD) {
    // FP32: pass twiddles W_2L^m, m < 2L, in LDS (the scatter's barriers
    // cover the fill).  FP64: none in LDS (it would cost a block per CU);
    // each thread forms its R2 twiddles from two table entries, below.
    constexpr bool TWL = sizeof(TC) == 4;
    __shared__ CT twl[TWL ? 2 * L : 1];
    if constexpr (TWL)
      for (int i = tid; i < 2 * L; i += NTH) twl[i] = T2L[i];
    auto TW = [&](uint32_t byte_off) {
      return *reinterpret_cast<const CT*>(reinterpret_cast<const char*>(twl) + byte_off);
    };
    // ---------------------------------------------------------- scatter
    for (int e = tid; e < C * L; e += NTH) lds[e] = zero;
    __syncthreads();
    const double2* in2 = a.site_in2[blockIdx.y];
    for (int k = s0 + tid; k < s1; k += NTH) {
      double2 v = ldg_stream(a.site_in[blockIdx.y] + k);
      if (in2) v.y = ldg_stream(in2 + k).x;
      lds[(a.site_ccell[k] >> (HL ? 1 : 0)) - cell0] = mk<TC>((TC)v.x, (TC)v.y);
    }
    __syncthreads();
    CT v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) v[n2] = ok13 ? lds[c13 * L + n1 + R1 * n2] : zero;
    __syncthreads();  // the line buffer and the exchange share LDS
    // ---------------------------------------------------------- P1
    if (ok13) {
      if (s13 == 1) rot_all<R2, false, TC>(v);
      rdft<R2, false, TC>(v);
      if constexpr (TWL) {
        const uint32_t tstep = (uint32_t)(2 * n1) * SC, tbase = (uint32_t)(s13 * n1) * SC;
#pragma unroll
        for (int q = 0; q < R2; ++q) {
          CT t = v[q];
          if (q > 0 || s13 == 1) t = cmul<TC>(t, TW(tbase + (uint32_t)q * tstep));
          lds[xi(c13, s13, q, n1)] = t;
        }
      } else {
        // W_2L^{n1 (2q + s13)} = W^{s13 n1} (W^{2 n1})^q, the powers by
        // repeated products (at most R2 - 1 deep: ~1e-15 relative in FP64)
        const CT d = T2L[2 * n1];
        CT cur = T2L[s13 * n1];
#pragma unroll
        for (int q = 0; q < R2; ++q) {
          CT t = v[q];
          if (q > 0 || s13 == 1) t = cmul<TC>(t, cur);
          lds[xi(c13, s13, q, n1)] = t;
          cur = cmul<TC>(cur, d);
        }
      }
    }
    __syncthreads();
    // ---------------------------------------------------------- P2 -> A
    if (ok2) {
      CT w[R1];
#pragma unroll
      for (int m = 0; m < R1; ++m) w[m] = lds[xi(c2, s2, k2, m)];
      rdft<R1, false, TC>(w);
      // block-uniform base of the prepare, 32-bit byte offsets per lane
      C2<TO>* ob = out + (size_t)blockIdx.y * a.out_bs;
      const int line = a.lines[b0 + c2];
      const int lb = line_base(a, line);
      int off[R1];
      freq_offs<R1, R2 * NSUB>(a, k2 * NSUB + s2, off);
      const int f0 = k2 * NSUB + s2;
      if constexpr (HL) {
        if (line_odd(a, line)) {
#pragma unroll
          for (int k1 = 0; k1 < R1; ++k1) w[k1] = cmul<TC>(w[k1], PH[f0 + k1 * R2 * NSUB]);
        }
      }
#pragma unroll
      for (int k1 = 0; k1 < K1N; ++k1) {
        const uint32_t o = f0 + k1 * R2 * NSUB < a.fmax ? (uint32_t)(lb + off[k1]) * (uint32_t)sizeof(C2<TO>) : COL_OOB;
        stg_buf<ST_XFWD>(ob, o, cvt<TC, TO>(w[k1]));
      }
    }
  } else {
    // ---------------------------------------------------------- P2 <- D
    // (the twiddles come from global memory here: a barrier before this phase
    // costs more than their loads)
    if (ok2) {
      CT w[R1];
      const C2<TI>* ib = in + (size_t)blockIdx.y * a.in_bs;
      const int line = a.lines[b0 + c2];
      const int lb = line_base(a, line);
      int off[R1];
      freq_offs<R1, R2 * NSUB>(a, k2 * NSUB + s2, off);
      const int f0 = k2 * NSUB + s2;
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const uint32_t o = f0 + k1 * R2 * NSUB < a.fmax ? (uint32_t)(lb + off[k1]) * (uint32_t)sizeof(C2<TI>) : COL_OOB;
        w[k1] = cvt<TI, TC>(ldg_stream_off(ib, o));
      }
      if constexpr (HL) {
        if (line_odd(a, line)) {
#pragma unroll
          for (int k1 = 0; k1 < R1; ++k1) w[k1] = cmulc<TC>(w[k1], PH[f0 + k1 * R2 * NSUB]);
        }
      }
      rdft<R1, true, TC>(w);
      const uint32_t tstep = (uint32_t)(2 * k2 + s2) * SC;
      // twiddles W_2L^{m q}, q = 2 k2 + s2.  FP64: the powers of two m come
      // from the table and every other m is the product of its highest power
      // of two and the rest (at most three products deep), so a thread loads
      // 4 sixteen-byte twiddles instead of R1 - 1, which through the L1 would
      // outweigh the pass's own data
      CT tw[R1];
#pragma unroll
      for (int m = 1; m < R1; ++m) {
        if (sizeof(TC) == 4 || (m & (m - 1)) == 0) {
          tw[m] = ldg_off(T2L, (uint32_t)m * tstep);
        } else {
          const int hi = 1 << (31 - __builtin_clz((unsigned)m));
          tw[m] = cmul<TC>(tw[hi], tw[m - hi]);
        }
      }
#pragma unroll
      for (int m = 0; m < R1; ++m) {
        CT t = w[m];
        if (m > 0) t = cmulc<TC>(t, tw[m]);
        lds[xi(c2, s2, k2, m)] = t;
      }
    }
    __syncthreads();
    // ---------------------------------------------------------- P3: one parity per item
    CT v[R2];
    if (ok13) {
#pragma unroll
      for (int q = 0; q < R2; ++q) v[q] = lds[xi(c13, s13, q, n1)];
      rdft<R2, true, TC>(v);
      if (s13 == 1) rot_all<R2, true, TC>(v);
    }
    __syncthreads();
    // both parities of each line side by side; the gather adds them
    if (ok13) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) lds[(c13 * NSUB + s13) * L + n1 + R1 * n2] = v[n2];
    }
    __syncthreads();
    // ---------------------------------------------------------- gather
    double2* out2 = a.site_out2[blockIdx.y];
    for (int k = s0 + tid; k < s1; k += NTH) {
      const int cc = (a.site_ccell[k] >> (HL ? 1 : 0)) - cell0;
      const int c = cc / L, x = cc - c * L;
      const CT u = lds[(c * NSUB) * L + x];
      double2 r = make_double2((double)u.x, (double)u.y);
      if (NSUB == 2) {
        const CT u1 = lds[(c * NSUB + 1) * L + x];
        r.x += (double)u1.x;
        r.y += (double)u1.y;
      }
      if (out2) {
        a.site_out[blockIdx.y][k] = make_double2(r.x, 0.0);
        out2[k] = make_double2(r.y, 0.0);
      } else {
        a.site_out[blockIdx.y][k] = r;
      }
    }
  }
}

struct LineLaunch {
  unsigned blocks, nb;
  hipStream_t stream;
  const void* in;
  void* out;
  LineArgs a;
  const void* T2L;
  const void* PH;  // half-length lines: W_4L^f, f < 2L (else null)
  bool hl;         // half-length lines
  bool ho;         // forward on full lines storing f < L only (fmax = L)
};

template <typename TI, typename TC, typename TO, int MODE, int R1, int R2>
hipError_t line_go(const LineLaunch& L) {
  constexpr int NSUB = 2;
  constexpr int C = line_c<TI, TC, R1, R2, NSUB>();
  constexpr int nth = line_threads<C, R1, R2, NSUB>();
  constexpr size_t lds = line_lds<TC, R1, R2, NSUB, C>();
  static_assert(lds <= 65536, "line pass tile exceeds 64 KiB of LDS");
  if (L.hl)
    hipLaunchKernelGGL((k_linepass<TI, TC, TO, MODE, R1, R2, NSUB, C, 1>), dim3(L.blocks, L.nb), dim3(nth), lds,
                       L.stream, (const C2<TI>*)L.in, (C2<TO>*)L.out, L.a, (const C2<TC>*)L.T2L,
                       (const C2<TC>*)L.PH);
  else if (MODE == PM_FWD && R1 % 2 == 0 && L.ho)
    hipLaunchKernelGGL((k_linepass<TI, TC, TO, MODE, R1, R2, NSUB, C, MODE == PM_FWD ? 2 : 0>), dim3(L.blocks, L.nb),
                       dim3(nth), lds, L.stream, (const C2<TI>*)L.in, (C2<TO>*)L.out, L.a, (const C2<TC>*)L.T2L,
                       (const C2<TC>*)L.PH);
  else
    hipLaunchKernelGGL((k_linepass<TI, TC, TO, MODE, R1, R2, NSUB, C, 0>), dim3(L.blocks, L.nb), dim3(nth), lds,
                       L.stream, (const C2<TI>*)L.in, (C2<TO>*)L.out, L.a, (const C2<TC>*)L.T2L,
                       (const C2<TC>*)L.PH);
  return hipGetLastError();
}

// lines per block of a pair and type combination, for the host's grid size
template <int R1, int R2>
int line_lines(int combo) {
  using d = double;
  using f = float;
  switch (combo) {
    case CB_DDD: case CB_DDF: return line_c<d, d, R1, R2, 2>();
    case CB_DFF: return line_c<d, f, R1, R2, 2>();
    case CB_FFD: return line_c<f, f, R1, R2, 2>();
    case CB_FDD: return line_c<f, d, R1, R2, 2>();
    default: return 0;
  }
}

// (types, mode) pairs of the atom-path x passes; hipErrorInvalidValue otherwise
template <int R1, int R2>
hipError_t line_launch(int combo, int mode, const LineLaunch& L) {
  using d = double;
  using f = float;
  if (mode == PM_FWD) {
    switch (combo) {
      case CB_DDD: return line_go<d, d, d, PM_FWD, R1, R2>(L);
      case CB_DFF: return line_go<d, f, f, PM_FWD, R1, R2>(L);
      case CB_DDF: return line_go<d, d, f, PM_FWD, R1, R2>(L);
      default: break;
    }
  } else if (mode == PM_INV) {
    switch (combo) {
      case CB_DDD: return line_go<d, d, d, PM_INV, R1, R2>(L);
      case CB_FFD: return line_go<f, f, d, PM_INV, R1, R2>(L);
      case CB_FDD: return line_go<f, d, d, PM_INV, R1, R2>(L);
      default: break;
    }
  }
  return hipErrorInvalidValue;
}

#define CLAT_L2_DECL(a, b)                                                       \
  hipError_t clat_l2_##a##x##b(int combo, int mode, const LineLaunch& L);        \
  int clat_l2_lines_##a##x##b(int combo);
CLAT_P2_PAIRS(CLAT_L2_DECL)

}  // namespace clat
