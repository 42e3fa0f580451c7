// HIP kernels of the engine.  Every kernel here is launched from engine.hip.
//
// Pass kernels of one prepare() (CoulombCalculator.cpp:105-163).  With the
// shared schedule the eight half-shift rounds are the eight parity classes
// of the 2X x 2Y x 2Z frequency grid; each pass turns a line of length L into
// two parity sub-lines (forward) or folds them back (inverse):
//   forward, parity p: FFT_L(x(n) e^{-i pi p n / L})  -> frequency 2k+p
//   inverse:           y(n) = IFFT_L(Y_0)(n) + e^{+i pi n / L} IFFT_L(Y_1)(n)
// With the rounds schedule each pass carries one sub-line and the phase of the
// round is applied in the first x pass and removed in the last one, as the
// reference's copy/phase sweeps do (CoulombCalculator.cpp:113-126,148-161).
#pragma once

#include "fft_device.hpp"

namespace clat {

// ---------------------------------------------------------------------------
// x passes: contiguous lines, line-major LDS tile of C lines.
// ---------------------------------------------------------------------------

// Forward x pass.  F: nlines lines of L; out: nlines lines of nsub*L.
// nsub == 2: out[l][2k+p] = FFT(x e^{-i pi p n/L})[k].
// nsub == 1: out[l][k] = FFT(x * ph)[k], ph = (kx ? hx[n] : 1) * hy[y]^ky * hz[z]^kz
//            with l = y + Y*z (the conj phase of round (kx,ky,kz)).
template <typename T, int EMAX>
__global__ __launch_bounds__(NT) void k_xfwd(const C2<T>* __restrict__ F, C2<T>* __restrict__ out,
                                             AxisPlan p, const C2<T>* __restrict__ tw,
                                             const C2<T>* __restrict__ hx, int nlines, int C, int lgC,
                                             int nsub, int kx, int ky, int kz,
                                             const C2<T>* __restrict__ hy, const C2<T>* __restrict__ hz,
                                             int Y) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int L = p.L;
  const int l0 = blockIdx.x * C;
  const int nC = min(C, nlines - l0);
  const int tot = C * L;
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int c = (int)fdiv((uint32_t)e, p.l_magic);
    const int i = e - c * L;
    C2<T> x = mk<T>((T)0, (T)0);
    if (c < nC) x = F[(size_t)(l0 + c) * L + i];
    if (nsub == 2) {
      buf[(2 * c) * L + i] = x;
      buf[(2 * c + 1) * L + i] = cmul<T>(x, hx[i]);
    } else {
      C2<T> ph = kx ? hx[i] : mk<T>((T)1, (T)0);
      if (ky | kz) {
        const int l = l0 + c;
        const int z = (c < nC) ? l / Y : 0;
        const int y = (c < nC) ? l - z * Y : 0;
        if (ky) ph = cmul<T>(ph, hy[y]);
        if (kz) ph = cmul<T>(ph, hz[z]);
      }
      buf[c * L + i] = cmul<T>(x, ph);
    }
  }
  __syncthreads();
  const int nl = nsub * C;
  lds_fft<false, EMAX, true, T>(buf, tw, p, nl, lgC + nsub - 1);
  if (nsub == 2) {
    const int tot2 = 2 * tot;
    for (int e = threadIdx.x; e < tot2; e += NT) {
      const int c = (int)fdiv((uint32_t)e, p.l2_magic);
      const int f = e - c * 2 * L;
      if (c < nC) out[(size_t)(l0 + c) * 2 * L + f] = buf[(2 * c + (f & 1)) * L + (f >> 1)];
    }
  } else {
    for (int e = threadIdx.x; e < tot; e += NT) {
      const int c = (int)fdiv((uint32_t)e, p.l_magic);
      if (c < nC) out[(size_t)l0 * L + e] = buf[e];
    }
  }
}

// Inverse x pass.  in: nlines lines of nsub*L; V: nlines lines of L.
// nsub == 2: V[l][n] = IFFT(Y_even)(n) + e^{i pi n/L} IFFT(Y_odd)(n).
// nsub == 1: V[l][n] (+)= IFFT(Y)(n) * conj(ph) (accumulate when acc != 0).
template <typename T, int EMAX>
__global__ __launch_bounds__(NT) void k_xinv(const C2<T>* __restrict__ in, C2<T>* __restrict__ V,
                                             AxisPlan p, const C2<T>* __restrict__ tw,
                                             const C2<T>* __restrict__ hx, int nlines, int C, int lgC,
                                             int nsub, int kx, int ky, int kz,
                                             const C2<T>* __restrict__ hy, const C2<T>* __restrict__ hz,
                                             int Y, int acc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int L = p.L;
  const int l0 = blockIdx.x * C;
  const int nC = min(C, nlines - l0);
  const int tot = C * L;
  if (nsub == 2) {
    const int tot2 = 2 * tot;
    for (int e = threadIdx.x; e < tot2; e += NT) {
      const int c = (int)fdiv((uint32_t)e, p.l2_magic);
      const int f = e - c * 2 * L;
      C2<T> x = mk<T>((T)0, (T)0);
      if (c < nC) x = in[(size_t)(l0 + c) * 2 * L + f];
      buf[(2 * c + (f & 1)) * L + (f >> 1)] = x;
    }
  } else {
    for (int e = threadIdx.x; e < tot; e += NT) {
      const int c = (int)fdiv((uint32_t)e, p.l_magic);
      buf[e] = (c < nC) ? in[(size_t)l0 * L + e] : mk<T>((T)0, (T)0);
    }
  }
  __syncthreads();
  lds_fft<true, EMAX, true, T>(buf, tw, p, nsub * C, lgC + nsub - 1);
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int c = (int)fdiv((uint32_t)e, p.l_magic);
    const int n = e - c * L;
    if (c >= nC) continue;
    C2<T> v;
    if (nsub == 2) {
      v = cadd<T>(buf[(2 * c) * L + n], cmulc<T>(buf[(2 * c + 1) * L + n], hx[n]));
    } else {
      C2<T> ph = kx ? hx[n] : mk<T>((T)1, (T)0);
      if (ky | kz) {
        const int l = l0 + c;
        const int z = l / Y;
        const int y = l - z * Y;
        if (ky) ph = cmul<T>(ph, hy[y]);
        if (kz) ph = cmul<T>(ph, hz[z]);
      }
      v = cmulc<T>(buf[c * L + n], ph);
    }
    const size_t g = (size_t)l0 * L + e;
    if (acc) v = cadd<T>(V[g], v);
    V[g] = v;
  }
}

// ---------------------------------------------------------------------------
// Column passes (y and z axes): a block holds C consecutive columns (the
// contiguous fx index) of one outer index o; element (c, i) of the line is at
// base + o*os + i*is + col0 + c.  Index-major LDS tile.
// ---------------------------------------------------------------------------

enum ColMode { COL_FWD = 0, COL_CONV = 1, COL_INV = 2 };

struct ColArgs {
  int W;          // number of columns
  int C, lgC;     // columns per block
  long long in_os, in_is, out_os, out_is;
  int nsub;       // 2: shared schedule (parity sub-lines), 1: one round
  // COL_CONV only: kernel table lookup
  int zpar;       // z parity when nsub == 1
  int gxm, gxa;   // fx = (col)*gxm + gxa
  int gym, gya;   // fy = o*gym + gya
  int X2, Y2, Z2; // 2*Xi, 2*Yi, 2*Zi (reflection r(f) = min(f, 2N - f))
  int GX1, GY1;   // table pitch: (Xi+1), (Yi+1)
};

template <typename T, int EMAX, int MODE>
__global__ __launch_bounds__(NT) void k_col(const C2<T>* in, C2<T>* out, AxisPlan p,
                                            const C2<T>* __restrict__ tw,
                                            const C2<T>* __restrict__ h, ColArgs a,
                                            const T* __restrict__ G) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int L = p.L;
  const int C = a.C, lgC = a.lgC;
  const int col0 = blockIdx.x * C;
  const int o = blockIdx.y;
  const int nC = min(C, a.W - col0);
  const int nsub = a.nsub;
  const int nl = nsub * C;
  const int lgnl = lgC + nsub - 1;
  const C2<T>* src = in + (size_t)o * a.in_os + col0;
  C2<T>* dst = out + (size_t)o * a.out_os + col0;

  if (MODE == COL_INV) {
    const int tot = nsub * L * C;
    for (int e = threadIdx.x; e < tot; e += NT) {
      const int c = e & (C - 1);
      const int r = e >> lgC;
      const int s = r & (nsub - 1);
      const int k = r >> (nsub - 1);
      buf[k * nl + s * C + c] = (c < nC) ? src[(size_t)r * a.in_is + c] : mk<T>((T)0, (T)0);
    }
  } else {
    const int tot = L * C;
    for (int e = threadIdx.x; e < tot; e += NT) {
      const int c = e & (C - 1);
      const int i = e >> lgC;
      C2<T> x = (c < nC) ? src[(size_t)i * a.in_is + c] : mk<T>((T)0, (T)0);
      if (nsub == 2) {
        buf[i * nl + c] = x;
        buf[i * nl + C + c] = cmul<T>(x, h[i]);
      } else {
        buf[i * nl + c] = x;
      }
    }
  }
  __syncthreads();

  if (MODE != COL_INV) lds_fft<false, EMAX, false, T>(buf, tw, p, nl, lgnl);

  if (MODE == COL_FWD) {
    const int tot = nsub * L * C;
    for (int e = threadIdx.x; e < tot; e += NT) {
      const int c = e & (C - 1);
      const int r = e >> lgC;
      const int s = r & (nsub - 1);
      const int k = r >> (nsub - 1);
      if (c < nC) dst[(size_t)r * a.out_is + c] = buf[k * nl + s * C + c];
    }
    return;
  }

  if (MODE == COL_CONV) {
    int fy = o * a.gym + a.gya;
    const int ry = min(fy, a.Y2 - fy);
    const int tot = nsub * L * C;
    for (int e = threadIdx.x; e < tot; e += NT) {
      const int c = e & (C - 1);
      const int r = e >> lgC;         // r = k*nsub + s
      const int s = r & (nsub - 1);
      const int k = r >> (nsub - 1);
      const int fz = 2 * k + (nsub == 2 ? s : a.zpar);
      const int rz = min(fz, a.Z2 - fz);
      int fx = (col0 + min(c, nC - 1)) * a.gxm + a.gxa;
      const int rx = min(fx, a.X2 - fx);
      const T g = G[((size_t)rz * a.GY1 + ry) * a.GX1 + rx];
      const int idx = k * nl + s * C + c;
      buf[idx] = cscale<T>(buf[idx], g);
    }
    __syncthreads();
  }

  lds_fft<true, EMAX, false, T>(buf, tw, p, nl, lgnl);
  {
    const int tot = L * C;
    for (int e = threadIdx.x; e < tot; e += NT) {
      const int c = e & (C - 1);
      const int n = e >> lgC;
      if (c >= nC) continue;
      C2<T> v = buf[n * nl + c];
      if (nsub == 2) v = cadd<T>(v, cmulc<T>(buf[n * nl + C + c], h[n]));
      dst[(size_t)n * a.out_is + c] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Interaction table (FP64) -- InteractionBase::map (Interaction.hpp:32-51)
// ---------------------------------------------------------------------------

struct GArgs {
  int X1, Y1, Z1;      // table extent (Xi+1, Yi+1, Zi+1)
  int ux, uy, uz;      // extent of the user-visible octant (X+1, ...); beyond it G = 0
  double sx, sy, sz;   // grid steps
  double onsite;
  double dielectric;
  int tf;              // Thomas-Fermi-Resta model
  double rTF, qTF;
};

__device__ __forceinline__ double eps_of_r(const GArgs& g, double r) {
  // InteractionThomasFermi::dielectric (Interaction.cpp:34-39)
  if (g.tf && r < g.rTF) return g.dielectric * (g.qTF * g.rTF / (sinh(g.qTF * (g.rTF - r)) + g.qTF * r));
  return g.dielectric;
}

// G at fine offset (ix, iy, iz) of the octant (zero past the user's table)
__device__ __forceinline__ double gval(const GArgs& g, int ix, int iy, int iz) {
  double v = 0.0;
  if (ix < g.ux && iy < g.uy && iz < g.uz) {
    const double z = iz * g.sz, z2 = z * z;
    const double y = iy * g.sy, y2z2 = z2 + y * y;
    const double x = ix * g.sx, r2 = y2z2 + x * x, r = sqrt(r2);
    v = 14.39963737103201 / (eps_of_r(g, r) * r);
    if (ix == 0 && iy == 0 && iz == 0) v = g.onsite;
  }
  return v;
}

__global__ void k_gtable(double* __restrict__ G, GArgs g) {
  const size_t n = (size_t)g.X1 * g.Y1 * g.Z1;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(t % g.X1);
    const size_t q = t / g.X1;
    const int iy = (int)(q % g.Y1);
    const int iz = (int)(q / g.Y1);
    G[t] = gval(g, ix, iy, iz);
  }
}

// Pads a user table of (ux)(uy)(uz) into the (X1)(Y1)(Z1) engine table.
__global__ void k_gpad(const double* __restrict__ src, double* __restrict__ G, GArgs g) {
  const size_t n = (size_t)g.X1 * g.Y1 * g.Z1;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(t % g.X1);
    const size_t q = t / g.X1;
    const int iy = (int)(q % g.Y1);
    const int iz = (int)(q / g.Y1);
    G[t] = (ix < g.ux && iy < g.uy && iz < g.uz) ? src[((size_t)iz * g.uy + iy) * g.ux + ix] : 0.0;
  }
}

// DCT-I (FFTW REDFT00, unnormalised) of every line of the table along one
// axis, through the 2M-point FFT of the even extension (M = extent - 1):
//   Y_k = g_0 + (-1)^k g_M + 2 sum_{n=1}^{M-1} g_n cos(pi n k / M),  k = 0..M.
// Line element i of line (o, c) is at G[o*os + i*is + c*cs]; a block holds
// C lines.  The tile is index-major when cs == 1 (consecutive columns) and
// line-major when is == 1.  The result is scaled by `scale`.
template <int EMAX, bool LM>
__device__ __forceinline__ void dct1_block(int bid, double2* buf, double* __restrict__ G, const AxisPlan& p,
                                           const double2* __restrict__ tw, int C, int lgC, long long os,
                                           long long is, long long cs, int ncols, double scale) {
  // Two real columns per complex line: the even extensions of columns c and
  // c + C are real and even, so their transforms are real and come back as
  // the real and imaginary parts of one complex FFT.
  const int L2 = p.L;      // 2M
  const int M = L2 / 2;
  // block -> (outer o, first column c0); 2C columns per block
  const int ncb = (ncols + 2 * C - 1) / (2 * C);
  const int o = bid / ncb;
  const int c0 = (bid - o * ncb) * 2 * C;
  const int nC = min(2 * C, ncols - c0);
  double* base = G + (size_t)o * os + (size_t)c0 * cs;
  const int tot = C * L2;
  // loads in batches of UB per thread, all issued before the first LDS write
  constexpr int UB = 8;
  for (int e0 = threadIdx.x; e0 < tot; e0 += NT * UB) {
    double v0[UB], v1[UB];
    int dst[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const int e = e0 + u * NT;
      v0[u] = 0.0;
      v1[u] = 0.0;
      dst[u] = -1;
      if (e < tot) {
        int c, m;
        if (LM) { c = (int)fdiv((uint32_t)e, p.l_magic); m = e - c * L2; }
        else { c = e & (C - 1); m = e >> lgC; }
        const int i = m <= M ? m : L2 - m;
        if (c < nC) v0[u] = base[(size_t)i * is + (size_t)c * cs];
        if (c + C < nC) v1[u] = base[(size_t)i * is + (size_t)(c + C) * cs];
        dst[u] = LM ? c * L2 + m : m * C + c;
      }
    }
#pragma unroll
    for (int u = 0; u < UB; ++u)
      if (dst[u] >= 0) buf[dst[u]] = make_double2(v0[u], v1[u]);
  }
  __syncthreads();
  lds_fft<false, EMAX, LM, double>(buf, tw, p, C, lgC);
  const int tot1 = C * (M + 1);
  for (int e = threadIdx.x; e < tot1; e += NT) {
    int c, k;
    if (LM) { c = e / (M + 1); k = e - c * (M + 1); }
    else { c = e & (C - 1); k = e >> lgC; }
    const double2 r = buf[LM ? c * L2 + k : k * C + c];
    if (c < nC) base[(size_t)k * is + (size_t)c * cs] = r.x * scale;
    if (c + C < nC) base[(size_t)k * is + (size_t)(c + C) * cs] = r.y * scale;
  }
}

template <int EMAX, bool LM>
__global__ __launch_bounds__(NT) void k_dct1(double* __restrict__ G, AxisPlan p,
                                             const double2* __restrict__ tw, int nlines, int C, int lgC,
                                             long long os, long long is, long long cs, int ncols,
                                             double scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  dct1_block<EMAX, LM>((int)blockIdx.x, reinterpret_cast<double2*>(smem_raw), G, p, tw, C, lgC, os, is, cs, ncols,
                       scale);
  (void)nlines;
}

// DCT-II (FFTW REDFT10, unnormalised) of every line along one axis,
//   Y_k = 2 sum_{n=0}^{M-1} x_n cos(pi (2n + 1) k / (2M)),  k = 0..M (Y_M = 0),
// through the 2M-point FFT of the zero-padded line:
//   Y_k = 2 Re(e^{-i pi k / (2M)} X_k).
// Two real columns c and c + C share one complex line (the real and
// imaginary parts); their transforms are separated from X_k and X_{2M-k}.
// Line layout and launch shape as k_dct1 (run by k_dct12); a line holds M inputs and M + 1
// outputs in place (the table's pitch has room for the extra one).
template <int EMAX, bool LM>
__device__ __forceinline__ void dct2_block(int bid, double2* buf, double* __restrict__ G, const AxisPlan& p,
                                           const double2* __restrict__ tw, int C, int lgC, long long os,
                                           long long is, long long cs, int ncols, double scale) {
  const int L2 = p.L;  // 2M
  const int M = L2 / 2;
  const int ncb = (ncols + 2 * C - 1) / (2 * C);
  const int o = bid / ncb;
  const int c0 = (bid - o * ncb) * 2 * C;
  const int nC = min(2 * C, ncols - c0);
  double* base = G + (size_t)o * os + (size_t)c0 * cs;
  const int tot = C * L2;
  for (int e = threadIdx.x; e < tot; e += NT) {
    int c, m;
    if (LM) { c = (int)fdiv((uint32_t)e, p.l_magic); m = e - c * L2; }
    else { c = e & (C - 1); m = e >> lgC; }
    double v0 = 0.0, v1 = 0.0;
    if (m < M) {
      if (c < nC) v0 = base[(size_t)m * is + (size_t)c * cs];
      if (c + C < nC) v1 = base[(size_t)m * is + (size_t)(c + C) * cs];
    }
    buf[LM ? c * L2 + m : m * C + c] = make_double2(v0, v1);
  }
  __syncthreads();
  lds_fft<false, EMAX, LM, double>(buf, tw, p, C, lgC);
  const int tot1 = C * (M + 1);
  for (int e = threadIdx.x; e < tot1; e += NT) {
    int c, k;
    if (LM) { c = e / (M + 1); k = e - c * (M + 1); }
    else { c = e & (C - 1); k = e >> lgC; }
    const int kr = k == 0 ? 0 : L2 - k;
    const double2 x = buf[LM ? c * L2 + k : k * C + c];
    const double2 r = buf[LM ? c * L2 + kr : kr * C + c];
    double sn, cn;
    sincospi((double)k / (double)L2, &sn, &cn);
    const double ya = cn * (x.x + r.x) + sn * (x.y - r.y);
    const double yb = cn * (x.y + r.y) - sn * (x.x - r.x);
    if (c < nC) base[(size_t)k * is + (size_t)c * cs] = ya * scale;
    if (c + C < nC) base[(size_t)k * is + (size_t)(c + C) * cs] = yb * scale;
  }
}

// The two-coset tables' transforms along one axis in one launch: DCT-I of
// the lines of G0 in blocks [0, nblk), DCT-II of the lines of Gt (same
// shape) in blocks [nblk, 2 nblk).  Each alone is a launch of about one wave
// per SIMD on the 1M-atom workload's coarse octant.
template <int EMAX, bool LM>
__global__ __launch_bounds__(NT) void k_dct12(double* __restrict__ G0, double* __restrict__ Gt, AxisPlan p,
                                              const double2* __restrict__ tw, int nblk, int C, int lgC,
                                              long long os, long long is, long long cs, int ncols, double scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* buf = reinterpret_cast<double2*>(smem_raw);
  const int b = (int)blockIdx.x;
  if (b < nblk) dct1_block<EMAX, LM>(b, buf, G0, p, tw, C, lgC, os, is, cs, ncols, scale);
  else dct2_block<EMAX, LM>(b - nblk, buf, Gt, p, tw, C, lgC, os, is, cs, ncols, scale);
}

// Real-space samples of the two-coset kernels from the fine real-space octant
// G ((Zi+1)(Yi+1)(Xi+1), G at fine offset (ix, iy, iz)): G0 = G(2d) on the
// coarse octant d in [0, Xc]^3, Gt = G(2d + (1,1,1)) on d in [0, Xc)^3 in a
// table of the same shape (zero where d reaches Xc on some axis).
__global__ void k_coset_samples(const double* __restrict__ G, int Xi, int Yi, int Zi, double* __restrict__ G0,
                                double* __restrict__ Gt) {
  const int Xc = Xi / 2, Yc = Yi / 2, Zc = Zi / 2;
  const size_t X1 = (size_t)Xi + 1, Y1 = (size_t)Yi + 1;
  const size_t n = (size_t)(Xc + 1) * (Yc + 1) * (Zc + 1);
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int dx = (int)(t % (Xc + 1));
    const size_t u = t / (Xc + 1);
    const int dy = (int)(u % (Yc + 1));
    const int dz = (int)(u / (Yc + 1));
    G0[t] = G[((size_t)(2 * dz) * Y1 + 2 * dy) * X1 + 2 * dx];
    Gt[t] = (dx < Xc && dy < Yc && dz < Zc) ? G[((size_t)(2 * dz + 1) * Y1 + 2 * dy + 1) * X1 + 2 * dx + 1] : 0.0;
  }
}

// The same samples straight from the interaction's arguments (no fine table):
// G0 = G(2d), Gt = G(2d + (1,1,1)), bit-identical to k_gtable's entries.
__global__ void k_coset_gtable(GArgs g, int Xi, int Yi, int Zi, double* __restrict__ G0, double* __restrict__ Gt) {
  const int Xc = Xi / 2, Yc = Yi / 2, Zc = Zi / 2;
  const size_t n = (size_t)(Xc + 1) * (Yc + 1) * (Zc + 1);
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int dx = (int)((uint32_t)t % (uint32_t)(Xc + 1));  // 32-bit: fewer than 2^32 entries
    const uint32_t u = (uint32_t)t / (uint32_t)(Xc + 1);
    const int dy = (int)(u % (uint32_t)(Yc + 1));
    const int dz = (int)(u / (uint32_t)(Yc + 1));
    G0[t] = gval(g, 2 * dx, 2 * dy, 2 * dz);
    Gt[t] = (dx < Xc && dy < Yc && dz < Zc) ? gval(g, 2 * dx + 1, 2 * dy + 1, 2 * dz + 1) : 0.0;
  }
}

// Half-box form of the two-coset tables, for coarse sites that all lie on
// mx + my + mz even (the fcc sublattice of the coarse grid, as zincblende's
// are): a density there has a transform periodic under the half-box shift
// s = (Xc, Yc, Zc), so only the frequencies kx < Xc are carried, against the
// kernel summed over the shift, K(k) + K(k + s).  Columns kx in [Xc, W) (W
// rounded up to the strip width) hold zero.
__device__ __forceinline__ int refl(int k, int n2) {  // k mod n2 for k in [0, 2 n2), reflected into [0, n2/2]
  if (k >= n2) k -= n2;
  return k <= n2 / 2 ? k : n2 - k;
}

// within-coset table [ry][column block][rz][c], ry in [0, Yc], rz in [0, Zc]
// (the summed kernel stays even in y and z), from the octant G0
template <typename T>
__global__ void k_gconv_tiled_half(const double* __restrict__ G0, int Xc, int Yc, int Zc, int C, int ncb,
                                   T* __restrict__ out) {
  const size_t n = (size_t)(Yc + 1) * ncb * (Zc + 1) * C;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    // 32-bit index arithmetic: the table has fewer than 2^32 entries
    const uint32_t c = (uint32_t)t % (uint32_t)C;
    uint32_t u = (uint32_t)t / (uint32_t)C;
    const int rz = (int)(u % (uint32_t)(Zc + 1));
    u /= (uint32_t)(Zc + 1);
    const int blk = (int)(u % (uint32_t)ncb);
    const int ry = (int)(u / (uint32_t)ncb);
    const int kx = blk * C + (int)c;
    double v = 0.0;
    if (kx < Xc) {
      const size_t X1 = (size_t)Xc + 1, Y1 = (size_t)Yc + 1;
      v = G0[((size_t)rz * Y1 + ry) * X1 + refl(kx, 2 * Xc)] +
          G0[((size_t)refl(rz + Zc, 2 * Zc) * Y1 + refl(ry + Yc, 2 * Yc)) * X1 + refl(kx + Xc, 2 * Xc)];
    }
    out[t] = (T)v;
  }
}

// cross-coset table [fy][column block][fz][c], fy in [0, 2Yc), fz in [0, 2Zc)
// (k_colconv2), Gt(k) + Gt(k + s) with Gt(k) = e^{i pi (kx/2Xc +
// ky/2Yc + kz/2Zc)} sign(k) St[r(kz)][r(ky)][r(kx)] on k in the coarse box
// ph: the per-axis phases e^{i pi k / (2 Xc)}, k < 2 Xc, then the same for y
// and z (2 Xc + 2 Yc + 2 Zc entries)
__device__ __forceinline__ double2 coarse_gt_at(const double* __restrict__ St, const double2* __restrict__ ph, int Xc,
                                               int Yc, int Zc, int kx, int ky, int kz) {
  const int rx = refl(kx, 2 * Xc), ry = refl(ky, 2 * Yc), rz = refl(kz, 2 * Zc);
  const int neg = (kx > Xc) + (ky > Yc) + (kz > Zc);
  const double st = ((neg & 1) ? -1.0 : 1.0) * St[((size_t)rz * (Yc + 1) + ry) * (Xc + 1) + rx];
  const double2 p = cmul<double>(cmul<double>(ph[kx], ph[2 * Xc + ky]), ph[2 * Xc + 2 * Yc + kz]);
  return make_double2(st * p.x, st * p.y);
}

template <typename T>
__global__ void k_coarse_q_half(const double* __restrict__ St, const double2* __restrict__ ph, int Xc, int Yc, int Zc,
                                int C, int ncb, C2<T>* __restrict__ Q) {
  const int Z2 = 2 * Zc, Y2 = 2 * Yc;
  const size_t n = (size_t)Y2 * ncb * Z2 * C;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)t % (uint32_t)C;  // 32-bit: fewer than 2^32 entries
    uint32_t u = (uint32_t)t / (uint32_t)C;
    const int fz = (int)(u % (uint32_t)Z2);
    u /= (uint32_t)Z2;
    const int blk = (int)(u % (uint32_t)ncb);
    const int fy = (int)(u / (uint32_t)ncb);
    const int kx = blk * C + (int)c;
    C2<T> v = mk<T>((T)0, (T)0);
    if (kx < Xc) {
      const double2 a = coarse_gt_at(St, ph, Xc, Yc, Zc, kx, fy, fz);
      const double2 b = coarse_gt_at(St, ph, Xc, Yc, Zc, kx + Xc, fy + Yc < Y2 ? fy + Yc : fy - Yc,
                                     fz + Zc < Z2 ? fz + Zc : fz - Zc);
      v = mk<T>((T)(a.x + b.x), (T)(a.y + b.y));
    }
    Q[t] = v;
  }
}

// The cross-coset table of k_colconv2 from the octant St (Zc+1)(Yc+1)(Xc+1)
// of DCT-II(Gt) / (8 Xc Yc Zc), for all 2 Xc x frequencies: Gt(fx, fy, fz) as
// in coarse_gt_at, tiled [fy][column block][fz][c] (fy in [0, 2Yc), fz in
// [0, 2Zc), fx = block * C + c < 2 Xc).
template <typename T>
__global__ void k_coarse_q(const double* __restrict__ St, const double2* __restrict__ ph, int Xc, int Yc, int Zc, int C,
                           int ncb, C2<T>* __restrict__ Q) {
  const int W = 2 * Xc, Z2 = 2 * Zc, Y2 = 2 * Yc;
  const size_t n = (size_t)Y2 * ncb * Z2 * C;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % C);
    size_t u = t / C;
    const int fz = (int)(u % Z2);
    u /= Z2;
    const int blk = (int)(u % ncb);
    const int fy = (int)(u / ncb);
    const int fx = blk * C + c;
    C2<T> v = mk<T>((T)0, (T)0);
    if (fx < W) {
      const double2 a = coarse_gt_at(St, ph, Xc, Yc, Zc, fx, fy, fz);
      v = mk<T>((T)a.x, (T)a.y);
    }
    Q[t] = v;
  }
}

// ---------------------------------------------------------------------------
// Small elementwise kernels
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_zero(C2<T>* __restrict__ p, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    p[t] = mk<T>((T)0, (T)0);
}

// Ĝ for the two-level conv pass in the old row layout:
// out[(ry*Z1 + rz)*2X + fx] = Ghat[rz][ry][r(fx)]
template <typename T>
__global__ void k_gconv(const double* __restrict__ G, T* __restrict__ out, int X, int Y1, int Z1) {
  const size_t n = (size_t)2 * X * Y1 * Z1;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int fx = (int)(t % (2 * X));
    const size_t q = t / (2 * X);
    const int rz = (int)(q % Z1);
    const int ry = (int)(q / Z1);
    const int rx = min(fx, 2 * X - fx);
    out[t] = (T)G[((size_t)rz * Y1 + ry) * (X + 1) + rx];
  }
}

// Ĝ tiled for the compile-time column kernel: [ry][p][blk][rz][c] with
// column col = blk*C + c of parity p, fx = col (shared schedule, npar = 1)
// or 2 col + p (rounds schedule, npar = 2); columns past W hold zero.
template <typename T>
__global__ void k_gconv_tiled(const double* __restrict__ G, T* __restrict__ out, int X, int Y1, int Z1, int C,
                              int ncb, int npar, int W) {
  const size_t n = (size_t)Y1 * npar * ncb * Z1 * C;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % C);
    size_t q = t / C;
    const int rz = (int)(q % Z1);
    q /= Z1;
    const int blk = (int)(q % ncb);
    q /= ncb;
    const int p = (int)(q % npar);
    const int ry = (int)(q / npar);
    const int col = blk * C + c;
    T v = (T)0;
    if (col < W) {
      const int fx = npar == 1 ? col : 2 * col + p;
      const int rx = min(fx, 2 * X - fx);
      v = (T)G[((size_t)rz * Y1 + ry) * (X + 1) + rx];
    }
    out[t] = v;
  }
}

// Kernel tables of the two-coset path from the fine table Gs (the octant
// (Zi+1)(Yi+1)(Xi+1) of DCT-I(G) / (8 Xi Yi Zi), Xi Yi Zi even).  With the
// fine grid split into the cosets 2m + p of the coarse grid (Xc = Xi / 2 per
// axis), the coarse transform of G(2d + p) over the coarse box of Xi points
// per axis is
//   Gp(q) = W_2Xi^(-p.q) sum_{s in {0,1}^3} (-1)^(p.s) Gs(q + s Xi)
// (the fine box split into its eight half-box shifts), already scaled for the
// coarse inverse.  Gs(k) for k in [Xi, 2Xi) is the octant entry 2Xi - k, so a
// coarse frequency q reads the octant at q and Xi - q on every axis.
__device__ __forceinline__ double coarse_sum(const double* __restrict__ Gs, int Xi, int Yi, int Zi, int qx, int qy,
                                             int qz, bool odd) {
  const size_t X1 = (size_t)Xi + 1, Y1 = (size_t)Yi + 1;
  const int xs[2] = {qx, Xi - qx}, ys[2] = {qy, Yi - qy}, zs[2] = {qz, Zi - qz};
  double t = 0.0;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const double v = Gs[((size_t)zs[c] * Y1 + ys[b]) * X1 + xs[a]];
        t += (odd && ((a + b + c) & 1)) ? -v : v;
      }
  return t;
}

// p = 0: the kernel within a coset, real and even about Xc on every axis;
// its octant (Zc+1)(Yc+1)(Xc+1) for the coarse engine's conv table
__global__ void k_coarse_g0(const double* __restrict__ Gs, int Xi, int Yi, int Zi, double* __restrict__ G0) {
  const int Xc = Xi / 2, Yc = Yi / 2, Zc = Zi / 2;
  const size_t n = (size_t)(Xc + 1) * (Yc + 1) * (Zc + 1);
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int qx = (int)(t % (Xc + 1));
    const size_t u = t / (Xc + 1);
    const int qy = (int)(u % (Yc + 1));
    const int qz = (int)(u / (Yc + 1));
    G0[t] = coarse_sum(Gs, Xi, Yi, Zi, qx, qy, qz, false);
  }
}

// p = (1,1,1): the kernel from coset 0 to coset 1,
//   Gt(q) = e^{i pi (qx/Xi + qy/Yi + qz/Zi)} St(q),  St = the alternating sum,
// St odd about Xc on every axis; its octant (Zc+1)(Yc+1)(Xc+1) (k_coarse_q
// tiles it for the coupled conv)
__global__ void k_coarse_st(const double* __restrict__ Gs, int Xi, int Yi, int Zi, double* __restrict__ St) {
  const int Xc = Xi / 2, Yc = Yi / 2, Zc = Zi / 2;
  const size_t n = (size_t)(Xc + 1) * (Yc + 1) * (Zc + 1);
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int qx = (int)(t % (Xc + 1));
    const size_t u = t / (Xc + 1);
    const int qy = (int)(u % (Yc + 1));
    const int qz = (int)(u / (Yc + 1));
    St[t] = coarse_sum(Gs, Xi, Yi, Zi, qx, qy, qz, true);
  }
}

__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    b[t] = (float)a[t];
}

// Engine grid (Xi,Yi,Zi) <-> user grid (X,Y,Z) copies with precision change.
template <typename T>
__global__ void k_grid_in(const double2* __restrict__ src, C2<T>* __restrict__ dst, int X, int Y, int Z,
                          int Xi, int Yi) {
  const size_t n = (size_t)X * Y * Z;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(t % X);
    const size_t q = t / X;
    const int iy = (int)(q % Y);
    const int iz = (int)(q / Y);
    const double2 v = src[t];
    dst[((size_t)iz * Yi + iy) * Xi + ix] = mk<T>((T)v.x, (T)v.y);
  }
}

template <typename T>
__global__ void k_grid_out(const C2<T>* __restrict__ src, double2* __restrict__ dst, int X, int Y, int Z,
                           int Xi, int Yi) {
  const size_t n = (size_t)X * Y * Z;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(t % X);
    const size_t q = t / X;
    const int iy = (int)(q % Y);
    const int iz = (int)(q / Y);
    const C2<T> v = src[((size_t)iz * Yi + iy) * Xi + ix];
    dst[t] = make_double2((double)v.x, (double)v.y);
  }
}

// F[cell[k]] = rho[k]  (Product::map, Product.cpp:16-30; the grid is zeroed first)
template <typename T>
__global__ void k_scatter(C2<T>* __restrict__ F, const int32_t* __restrict__ cell,
                          const double2* __restrict__ rho, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = rho[t];
    F[cell[t]] = mk<T>((T)v.x, (T)v.y);
  }
}

// out[k] = V[cell[k]] in FP64
template <typename T>
__global__ void k_gather(const C2<T>* __restrict__ V, const int32_t* __restrict__ cell,
                         double2* __restrict__ out, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const C2<T> v = V[cell[t]];
    out[t] = make_double2((double)v.x, (double)v.y);
  }
}

// ---------------------------------------------------------------------------
// Pair densities: rho_P[k] = sum_o conj(c_L[k,o]) c_R[k,o] for every product
// P = L(L+1)/2 + R, R <= L (FunctionCollection::createProducts,
// FunctionCollection.cpp:91-105; ProductFromTightBinding::map, Product.cpp:16-30).
// coeffs: [state][site][orbital] (site order, gathered at upload).
// A block stages S sites x all states in LDS as [state][orbital][site] (pitch
// S + 1: conflict-free 16-byte accesses), G = blockDim / S thread groups
// share the products of each site.
// ---------------------------------------------------------------------------
// read-once double2 load with the nontemporal hint
__device__ __forceinline__ double2 nt_load(const double2* p) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  const d2v u = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
  return make_double2(u.x, u.y);
}

__device__ __forceinline__ void pair_of(int p, int& L, int& R) {
  L = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
  while (L * (L + 1) / 2 > p) --L;
  while ((L + 1) * (L + 2) / 2 <= p) ++L;
  R = p - L * (L + 1) / 2;
}

// FLAT: the staging loads run over (state, element) on all threads, UB per
// thread in flight before their LDS writes (many states, few sites per block:
// 32 states x 4 sites, 1.57 against 1.77 ms); else over the elements of one
// state with the state loop unrolled by four (8 states x 8 sites: 0.59
// against 0.61 ms)
template <bool FLAT>
__global__ __launch_bounds__(NT) void k_products(const double2* __restrict__ coeffs, int64_t n_sites,
                                                 int n_states, int n_orb, int S, double2* __restrict__ rho) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2* cs = reinterpret_cast<double2*>(smem_raw);
  const int SP = S + 1;
  const int64_t k0 = (int64_t)blockIdx.x * S;
  const int nA = (int)min((int64_t)S, n_sites - k0);
  const int per = nA * n_orb;  // contiguous elements of one state for this block
  if constexpr (FLAT) {
    const int tot = n_states * per;
    constexpr int UB = 8;
    for (int e0 = threadIdx.x; e0 < tot; e0 += NT * UB) {
      double2 v[UB];
      int dst[UB];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        const int e = e0 + u * NT;
        dst[u] = -1;
        if (e < tot) {
          const int s = e / per, w = e - s * per;
          const int a = w / n_orb, o = w - a * n_orb;
          v[u] = nt_load(coeffs + ((size_t)s * n_sites + k0) * n_orb + w);
          dst[u] = (s * n_orb + o) * SP + a;
        }
      }
#pragma unroll
      for (int u = 0; u < UB; ++u)
        if (dst[u] >= 0) cs[dst[u]] = v[u];
    }
  } else {
    for (int w = threadIdx.x; w < per; w += NT) {
      const int a = w / n_orb, o = w - a * n_orb;
      const double2* src = coeffs + (size_t)k0 * n_orb + w;
      double2* dst = cs + (size_t)o * SP + a;
      int s = 0;
      for (; s + 4 <= n_states; s += 4) {
        double2 v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = nt_load(src + (size_t)(s + j) * n_sites * n_orb);
#pragma unroll
        for (int j = 0; j < 4; ++j) dst[(size_t)(s + j) * n_orb * SP] = v[j];
      }
      for (; s < n_states; ++s) dst[(size_t)s * n_orb * SP] = nt_load(src + (size_t)s * n_sites * n_orb);
    }
  }
  __syncthreads();
  const int a = threadIdx.x % S, g = threadIdx.x / S, G = NT / S;
  if (a >= nA) return;
  const int P = n_states * (n_states + 1) / 2;
  for (int p = g; p < P; p += G) {
    int L, R;
    pair_of(p, L, R);
    const double2* l = cs + (size_t)L * n_orb * SP + a;
    const double2* r = cs + (size_t)R * n_orb * SP + a;
    double re = 0.0, im = 0.0;
    for (int o = 0; o < n_orb; ++o) {
      const double2 x = l[(size_t)o * SP], y = r[(size_t)o * SP];
      re += x.x * y.x + x.y * y.y;
      im += x.x * y.y - x.y * y.x;
    }
    // a diagonal density |c_L|^2 is real: its imaginary part is zero, not
    // the rounding residue of the contracted products
    rho[(size_t)p * n_sites + k0 + a] = make_double2(re, L == R ? 0.0 : im);
  }
}

// The self density of one state at the sites, sum_o |c_f(site, o)|^2: the
// product a potential of that state needs (FunctionCollection::
// createSelfProducts, FunctionCollection.cpp:108-120, + Product::map,
// Product.cpp:16-30), without forming the
// other n(n+1)/2 - 1 pair densities.  Real: the imaginary part is zero.
__global__ void k_self_density(const double2* __restrict__ coeffs, int64_t n_sites, int n_orb, int f,
                               double2* __restrict__ rho) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n_sites; t += (int64_t)gridDim.x * blockDim.x) {
    const double2* c = coeffs + ((size_t)f * n_sites + t) * n_orb;
    double re = 0.0;
    for (int o = 0; o < n_orb; ++o) {
      const double2 x = c[o];
      re += x.x * x.x + x.y * x.y;
    }
    rho[t] = make_double2(re, 0.0);
  }
}

// out[(k*n_orb + o)] = in[(atom[k]*n_orb + o)]: one state's coefficients in site order
__global__ void k_coeffs_to_sites(const double2* __restrict__ in, const int32_t* __restrict__ atom,
                                  int64_t n_sites, int n_orb, double2* __restrict__ out) {
  const int64_t n = n_sites * n_orb;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = t / n_orb;
    const int32_t at = atom[k];  // -1: a two-coset site with no atom in that coset
    out[t] = at < 0 ? make_double2(0.0, 0.0) : in[(int64_t)at * n_orb + (t - k * n_orb)];
  }
}

// Same without LDS staging, for state counts whose staging would not fit.
__global__ void k_products_direct(const double2* __restrict__ coeffs, int64_t n_sites, int n_states, int n_orb,
                                  double2* __restrict__ rho) {
  const int P = n_states * (n_states + 1) / 2;
  const int64_t tot = (int64_t)P * n_sites;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int pidx = (int)(t / n_sites);
    const int64_t k = t - (int64_t)pidx * n_sites;
    int fL, fR;
    pair_of(pidx, fL, fR);
    const double2* l = coeffs + ((size_t)fL * n_sites + k) * n_orb;
    const double2* r = coeffs + ((size_t)fR * n_sites + k) * n_orb;
    double re = 0.0, im = 0.0;
    for (int o = 0; o < n_orb; ++o) {
      const double2 x = l[o], y = r[o];
      re += x.x * y.x + x.y * y.y;
      im += x.x * y.y - x.y * y.x;
    }
    rho[t] = make_double2(re, fL == fR ? 0.0 : im);
  }
}

// ---------------------------------------------------------------------------
// Overlaps (CoulombCalculator::calculate, CoulombCalculator.cpp:165-172) at
// atom sites.  For every (left slot l, right product r) four real sums over
// the sites are formed,
//   A = sum Vr*pr, B = sum Vi*pi, C = sum Vr*pi, D = sum Vi*pr,
// from which  sum V*rho = (A-B) + i(C+D)  and  sum V*conj(rho) = (A+B) + i(D-C).
// Stage 1: a T x T (slot, product) tile per block, PT x PT pairs per thread,
// over one chunk of sites; partial sums to part[slice][l][r][4].  Two shapes:
// TX = 16, PT = 3 (48 x 48, one group of 256 threads), and TX = 9, PT = 4
// (36 x 36 without idle pairs for up to 36 slots and products, as eight
// states give; three groups of 81 threads, each summing every third staged
// site into its own slice, so slice = chunk * G + group).
// ---------------------------------------------------------------------------
constexpr int OV_S = 16;   // sites staged per step (48 x 48 tiles)

// s += a*b with Kahan compensation: c carries the negated low part lost so
// far, so the compensated sum is s - c
__device__ __forceinline__ void kahan_fma(double& s, double& c, double a, double b) {
  const double y = fma(a, b, -c);
  const double t = s + y;
  c = (t - s) - y;
  s = t;
}

// CMP: every one of the 4 PT^2 sums is Kahan-compensated site by site (a
// second register per sum), so no sum over the sites is left uncompensated
template <int TX, int PT, int G, int OVS, bool CMP = false>
__global__ __launch_bounds__(TX * TX * G) void k_overlap_partial(const double2* __restrict__ Vat,
                                                                 const double2* __restrict__ rho,
                                                                 int n_left, int n_right, int64_t n_sites,
                                                                 int64_t chunk, double* __restrict__ part) {
  static_assert(OVS % G == 0, "each group takes the same number of staged sites");
  constexpr int T = PT * TX, NG = TX * TX, NTH = NG * G;
  __shared__ double2 sv[T][OVS + 1];
  __shared__ double2 sr[T][OVS + 1];
  const int g = threadIdx.x / NG, t = threadIdx.x - g * NG;
  const int tx = t % TX, ty = t / TX;
  const int l0 = blockIdx.x * T, r0 = blockIdx.y * T;
  const int64_t s0 = (int64_t)blockIdx.z * chunk;
  const int64_t s1 = min(n_sites, s0 + chunk);
  double acc[PT][PT][4];
  double cmp[CMP ? PT : 1][CMP ? PT : 1][4];
#pragma unroll
  for (int i = 0; i < PT; ++i)
#pragma unroll
    for (int j = 0; j < PT; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[i][j][q] = 0.0;
        if constexpr (CMP) cmp[i][j][q] = 0.0;
      }
  constexpr int NE = (T * OVS + NTH - 1) / NTH;  // staged elements per thread and table
  // software pipeline: the next batch of sites is loaded into registers while
  // the current one is summed from LDS
  double2 bv[NE], br[NE];
  auto fetch = [&](int64_t sb) {
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * NTH;
      const int row = e / OVS, col = e - row * OVS;
      const int64_t site = sb + col;
      const bool ok = e < T * OVS && site < s1;
      bv[j] = (ok && l0 + row < n_left) ? nt_load(Vat + (size_t)(l0 + row) * n_sites + site) : make_double2(0.0, 0.0);
      br[j] = (ok && r0 + row < n_right) ? nt_load(rho + (size_t)(r0 + row) * n_sites + site) : make_double2(0.0, 0.0);
    }
  };
  if (s0 < s1) fetch(s0);
  for (int64_t sb = s0; sb < s1; sb += OVS) {
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * NTH;
      if (e < T * OVS) {
        const int row = e / OVS, col = e - row * OVS;
        sv[row][col] = bv[j];
        sr[row][col] = br[j];
      }
    }
    __syncthreads();
    if (sb + OVS < s1) fetch(sb + OVS);
#pragma unroll 2
    for (int s = g; s < OVS; s += G) {
      double2 v[PT], r[PT];
#pragma unroll
      for (int i = 0; i < PT; ++i) { v[i] = sv[ty + TX * i][s]; r[i] = sr[tx + TX * i][s]; }
#pragma unroll
      for (int i = 0; i < PT; ++i)
#pragma unroll
        for (int j = 0; j < PT; ++j) {
          if constexpr (CMP) {
            kahan_fma(acc[i][j][0], cmp[i][j][0], v[i].x, r[j].x);
            kahan_fma(acc[i][j][1], cmp[i][j][1], v[i].y, r[j].y);
            kahan_fma(acc[i][j][2], cmp[i][j][2], v[i].x, r[j].y);
            kahan_fma(acc[i][j][3], cmp[i][j][3], v[i].y, r[j].x);
          } else {
            acc[i][j][0] = fma(v[i].x, r[j].x, acc[i][j][0]);
            acc[i][j][1] = fma(v[i].y, r[j].y, acc[i][j][1]);
            acc[i][j][2] = fma(v[i].x, r[j].y, acc[i][j][2]);
            acc[i][j][3] = fma(v[i].y, r[j].x, acc[i][j][3]);
          }
        }
    }
    __syncthreads();
  }
  const size_t slice = (size_t)blockIdx.z * G + g;
#pragma unroll
  for (int i = 0; i < PT; ++i)
#pragma unroll
    for (int j = 0; j < PT; ++j) {
      const int l = l0 + ty + TX * i, r = r0 + tx + TX * j;
      if (l < n_left && r < n_right) {
        double* d = part + ((slice * n_left + l) * n_right + r) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if constexpr (CMP) d[q] = acc[i][j][q] - cmp[i][j][q];
          else d[q] = acc[i][j][q];
        }
      }
    }
}

// Stage 1 on FP64 MFMA, for many slots and products: the four sums of a
// 64 x 64 (slot, product) tile are four real matrix products over the sites,
//   A = Vr Pr^T, B = Vi Pi^T, C = Vr Pi^T, D = Vi Pr^T,
// each a 4 x 4 grid of 16 x 16 v_mfma_f64_16x16x4f64 tiles.  Wave w owns
// slots 16w..16w+15 against all 64 products: 16 independent accumulator
// chains (4 product tiles x 4 sums).  Operand lanes: A[i][k] in lane
// i + 16k (slot i, site k), B[k][j] in lane j + 16k (site k, product j);
// result D[i][j] in lane j + 16 (i % 4), element i / 4.
// Same chunking and partial layout as k_overlap_partial.
// tri > 0: every product is a left slot (n_left == n_right, slot index =
// product index) and the table is symmetric: sum V_a rho_b = sum rho_a V_b
// (the kernel G is real and even), so only the tri (tri + 1) / 2 tiles on or
// above the diagonal are formed, blockIdx.x walking them row by row; the
// host reads an entry below the diagonal from its mirror with C and D
// exchanged.
constexpr int OVM_T = 64;   // tile edge (slots and products)
constexpr int OVM_S = 16;   // sites staged per step
constexpr int OVM_MIN = 96; // both sides at least this many: the MFMA tiles are used

__global__ __launch_bounds__(256) void k_overlap_partial_mfma(const double2* __restrict__ Vat,
                                                              const double2* __restrict__ rho,
                                                              int n_left, int n_right, int64_t n_sites,
                                                              int64_t chunk, double* __restrict__ part, int tri) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int P = OVM_S + 1;  // LDS pitch in double2
  __shared__ double2 sv[OVM_T * P];
  __shared__ double2 sr[OVM_T * P];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int tl = (int)blockIdx.x, tr = (int)blockIdx.y;
  if (tri > 0) {
    // upper-triangle tile number -> (row tl, column tr >= tl)
    int t = (int)blockIdx.x;
    tl = 0;
    while (t >= tri - tl) {
      t -= tri - tl;
      ++tl;
    }
    tr = tl + t;
  }
  const int l0 = tl * OVM_T, r0 = tr * OVM_T;
  const int64_t s0 = (int64_t)blockIdx.z * chunk;
  const int64_t s1 = min(n_sites, s0 + chunk);
  d4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[t][q] = d4{0.0, 0.0, 0.0, 0.0};
  // staging: 64 rows x 16 sites per table, 4 elements per thread and table;
  // 16 consecutive threads read one row's 256 contiguous bytes
  constexpr int NE = OVM_T * OVM_S / 256;
  double2 bv[NE], br[NE];
  auto fetch = [&](int64_t sb) {
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * 256;
      const int row = e >> 4, col = e & 15;
      const int64_t site = sb + col;
      const bool ok = site < s1;
      bv[j] = (ok && l0 + row < n_left) ? nt_load(Vat + (size_t)(l0 + row) * n_sites + site) : make_double2(0.0, 0.0);
      br[j] = (ok && r0 + row < n_right) ? nt_load(rho + (size_t)(r0 + row) * n_sites + site) : make_double2(0.0, 0.0);
    }
  };
  const int fi = lane & 15, fk = lane >> 4;  // operand row/column and site within a 4-site step
  if (s0 < s1) fetch(s0);
  for (int64_t sb = s0; sb < s1; sb += OVM_S) {
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * 256;
      const int row = e >> 4, col = e & 15;
      sv[row * P + col] = bv[j];
      sr[row * P + col] = br[j];
    }
    __syncthreads();
    if (sb + OVM_S < s1) fetch(sb + OVM_S);
#pragma unroll
    for (int k = 0; k < OVM_S; k += 4) {
      const double2 a = sv[(16 * w + fi) * P + k + fk];
      double2 b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) b[t] = sr[(16 * t + fi) * P + k + fk];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b[t].x, acc[t][0], 0, 0, 0);
        acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b[t].y, acc[t][1], 0, 0, 0);
        acc[t][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b[t].y, acc[t][2], 0, 0, 0);
        acc[t][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b[t].x, acc[t][3], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int r = r0 + 16 * t + fi;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int l = l0 + 16 * w + 4 * v + fk;
      if (l < n_left && r < n_right) {
        double2* d = reinterpret_cast<double2*>(part + (((size_t)blockIdx.z * n_left + l) * n_right + r) * 4);
        d[0] = make_double2(acc[t][0][v], acc[t][1][v]);
        d[1] = make_double2(acc[t][2][v], acc[t][3][v]);
      }
    }
  }
}

// Stage 2: compensated sum of the chunk partials.  A block owns 64 outputs;
// its 16 row groups of lanes each run a Neumaier sum over a stride of the
// chunks, then the 16 (sum, compensation) pairs are merged in LDS.
__device__ __forceinline__ void neumaier_add(double& s, double& c, double x) {
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
  s = t;
}

__global__ __launch_bounds__(1024) void k_overlap_reduce(const double* __restrict__ part, int64_t n_chunks,
                                                         int64_t n_out, double* __restrict__ out) {
  __shared__ double ss[16][64], sc[16][64];
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int64_t o = (int64_t)blockIdx.x * 64 + lane;
  double s = 0.0, c = 0.0;
  if (o < n_out)
    for (int64_t ch = grp; ch < n_chunks; ch += 16) neumaier_add(s, c, part[(size_t)ch * n_out + o]);
  ss[grp][lane] = s;
  sc[grp][lane] = c;
  __syncthreads();
  if (grp == 0 && o < n_out) {
    double S = 0.0, Cc = 0.0;
    for (int g = 0; g < 16; ++g) {
      neumaier_add(S, Cc, ss[g][lane]);
      Cc += sc[g][lane];
    }
    out[o] = S + Cc;
  }
}

// Full-grid bilinear dot sum_n V(n) F(n) (the reference's calculate()).
// Per-thread FP64 sums, wave and block reduction, partials per block.
template <typename T>
__global__ __launch_bounds__(NT) void k_grid_dot(const C2<T>* __restrict__ V, const double2* __restrict__ F,
                                                 int X, int Y, int Z, int Xi, int Yi,
                                                 double* __restrict__ part) {
  const size_t n = (size_t)X * Y * Z;
  double re = 0.0, im = 0.0;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const double2 f = F[t];
    if (f.x == 0.0 && f.y == 0.0) continue;
    const int ix = (int)(t % X);
    const size_t q = t / X;
    const int iy = (int)(q % Y);
    const int iz = (int)(q / Y);
    const C2<T> v = V[((size_t)iz * Yi + iy) * Xi + ix];
    re += (double)v.x * f.x - (double)v.y * f.y;
    im += (double)v.x * f.y + (double)v.y * f.x;
  }
  for (int off = 32; off > 0; off >>= 1) {
    re += __shfl_xor(re, off, 64);
    im += __shfl_xor(im, off, 64);
  }
  __shared__ double sre[NT / 64], sim[NT / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sre[w] = re; sim[w] = im; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < NT / 64; ++i) { a += sre[i]; b += sim[i]; }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

}  // namespace clat
