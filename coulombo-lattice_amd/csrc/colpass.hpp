// Column passes (y forward, z convolution, y inverse) with every size fixed at
// compile time: radix pair R1 x R2, sub-line count NSUB and the number of
// columns per block C.  Compared with the general two-level kernel this one
//   * keeps one sub-line in LDS at a time (the parities run one after the
//     other), halving the LDS per block;
//   * addresses LDS with compile-time offsets from one per-thread base, and
//     global memory through buffer resources: a per-block base in scalar
//     registers and 32-bit lane offsets (masked rows take an out-of-range
//     offset instead of a branch);
//   * folds the half-shift phase h[n] = W_2L^n = W_2L^n1 W_2R2^n2 into the
//     pass twiddle (W_2L^{n1 (2 k2 + s)}, one table of 2L roots) and a
//     compile-time pre/post rotation by W_2R2^n2.
//
// Index map (L = R1 R2):  n = n1 + R1 n2,  k = R2 k1 + k2.
// Thread t of the block owns column c = t % C and row r = t / C; r is n1 in
// the phases that hold R2 values and k2 in the phase that holds R1 values.
#pragma once

#include "pass2.hpp"

namespace clat {

struct ColArgs2 {
  int W;                          // columns in the plane
  // outer index o (after the conv reordering) -> element offset
  //   (o / od) * os_hi + (o % od) * os_lo
  int od;
  long long in_os_hi, in_os_lo, out_os_hi, out_os_lo;
  long long in_bs, out_bs;        // element stride between batched prepares (blockIdx.z, or
                                  // blockIdx.x / nxb when folded)
  long long in_cbs, out_cbs;      // element stride between blocks of C columns (blockIdx.x): C for
                                  // row-major planes, C x rows for planes tiled in C-column strips
  // rows: n-domain row n1 + R1 n2 -> byte offset n2 * is_hi + n1 * is_lo;
  // k-domain rows j (y passes only) -> j * is_lo
  unsigned in_is_lo, in_is_hi, out_is_lo, out_is_hi;
  // conv pass: fy = o * gym + gya; the kernel table is tiled per block of C
  // columns as [ry][parity p][column block][rz][C]; pair = 1 orders the
  // blocks so that fy and 2Y - fy (same table tile) run next to each other
  int gym, gya, zpar, Y2, GZ1, ncb, npar, par, pair;
  // y passes of the atom path: word rowmask[o * mask_os + n1], bit n2 clear =
  // n-domain row n1 + R1 n2 is an x-line without sites, never read (forward)
  // or written (inverse)
  const uint32_t* rowmask;
  int mask_os;
  // launch order: 0 = blockIdx.x is the column block and blockIdx.y the outer
  // index, 1 = the reverse (outer index fastest, so the blocks in flight share
  // column blocks)
  int outer_x;
  // > 0: the batched prepares are folded into blockIdx.x = prepare * nxb +
  // column block (nxb column blocks), instead of blockIdx.z
  int nxb;
  // coupled conv of the two-coset path: element offset of the second coset's
  // plane from the first (input and output)
  long long cset;
};

template <typename T>
__device__ __forceinline__ T ldg_off(const T* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <typename T>
__device__ __forceinline__ void stg_off(T* base, uint32_t byte_off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off) = v;
}
// nontemporal store through a 64-bit address (the y forward's output)
template <typename T>
__device__ __forceinline__ void stg_nt(T* base, uint32_t byte_off, T v) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  static_assert(sizeof(T) == 8 || sizeof(T) == 16, "complex64 or complex128 element");
  T* p = reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off);
  if constexpr (sizeof(T) == 8) {
    unsigned long long u;
    __builtin_memcpy(&u, &v, 8);
    __builtin_nontemporal_store(u, reinterpret_cast<unsigned long long*>(p));
  } else {
    d2v u;
    __builtin_memcpy(&u, &v, 16);
    __builtin_nontemporal_store(u, reinterpret_cast<d2v*>(p));
  }
}
// Cache policy of each pass's stores (buffer-instruction policy bits: 0
// default, 2 nontemporal).  Measured on the 1M-atom step, same box, two or
// three alternations each.  FP32X: the x forward (A), y forward (B) and z conv
// (B in place) storing nontemporal and the y inverse (D) with the default
// policy gives 9.75-9.78 ms per step, everything default 10.05-10.09; with
// the y inverse nontemporal too the x inverse reading D slows from 0.60 to
// 0.69 ms (D of four prepares, 268 MB, is partly still in the 256 MB
// last-level cache when it starts).  The y forward's nontemporal stores take
// a 64-bit address: as buffer stores with the same policy the pass ran
// 2.01-2.08 ms per step against 1.89-1.90.  FP64 (complex128): the same
// policies except a default-policy y forward give 18.14-18.18 ms against
// 18.36-18.43 all default; a nontemporal complex128 y forward took that pass
// from 3.7 to 4.2-4.4 ms, and the x forward with the default policy slowed
// the y forward reading A from 3.73 to 3.98 ms.
constexpr int ST_CONV = 2, ST_YINV = 0, ST_XFWD = 2;
// streaming load (read once by this pass): nontemporal hint, so the line does
// not displace reused data in the caches; measured 16-26 % faster y passes
template <typename T>
__device__ __forceinline__ T ldg_stream(const T* p) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  static_assert(sizeof(T) == 8 || sizeof(T) == 16, "complex64 or complex128 element");
  T r;
  if constexpr (sizeof(T) == 8) {
    const unsigned long long u = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(p));
    __builtin_memcpy(&r, &u, 8);
  } else {
    const d2v u = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    __builtin_memcpy(&r, &u, 16);
  }
  return r;
}
// Column-pass memory access from a block-uniform base: a buffer resource in
// scalar registers with the byte offset as the lane offset, so the lanes need
// no 64-bit address arithmetic (a nontemporal global load takes a full
// per-lane address).  The resource covers COL_RANGE bytes; a lane that must
// not touch memory (a row of an empty x-line, a column past the plane) passes
// COL_OOB, which is out of range: the load returns zero and the store is
// dropped by the hardware, with no branch around the instruction.  Offsets of
// real elements stay below COL_RANGE (the host checks rows x stride + 4 KiB
// < 2^32).
// a block-uniform product (row index x row stride) kept in a scalar register:
// the compiler would otherwise fold it with the lane offset into a 64-bit
// integer multiply-add per row
__device__ __forceinline__ uint32_t suni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

constexpr uint32_t COL_RANGE = 0xFFFFFF00u;
constexpr uint32_t COL_OOB = 0xFFFFFFF0u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t col_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)COL_RANGE, 0x00020000);
}
// streaming load (nontemporal cache policy)
template <typename T>
__device__ __forceinline__ T ldg_stream_off(const T* base, uint32_t byte_off) {
  static_assert(sizeof(T) == 8 || sizeof(T) == 16, "complex64 or complex128 element");
  const __amdgpu_buffer_rsrc_t rs = col_rsrc(base);
  constexpr int NT_POLICY = 2;  // nt
  T r;
  if constexpr (sizeof(T) == 8) {
    typedef int i2 __attribute__((ext_vector_type(2)));
    const i2 u = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)byte_off, 0, NT_POLICY);
    __builtin_memcpy(&r, &u, 8);
  } else {
    typedef int i4 __attribute__((ext_vector_type(4)));
    const i4 u = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)byte_off, 0, NT_POLICY);
    __builtin_memcpy(&r, &u, 16);
  }
  return r;
}
template <int POL = 0, typename T>
__device__ __forceinline__ void stg_buf(T* base, uint32_t byte_off, T v) {
  static_assert(sizeof(T) == 8 || sizeof(T) == 16, "complex64 or complex128 element");
  const __amdgpu_buffer_rsrc_t rs = col_rsrc(base);
  if constexpr (sizeof(T) == 8) {
    typedef int i2 __attribute__((ext_vector_type(2)));
    i2 u;
    __builtin_memcpy(&u, &v, 8);
    __builtin_amdgcn_raw_buffer_store_b64(u, rs, (int)byte_off, 0, POL);
  } else {
    typedef int i4 __attribute__((ext_vector_type(4)));
    i4 u;
    __builtin_memcpy(&u, &v, 16);
    __builtin_amdgcn_raw_buffer_store_b128(u, rs, (int)byte_off, 0, POL);
  }
}

// v * W_R^{e} (forward sign) or v * W_R^{-e} (CONJ), W_R = exp(-2 pi i / R), e compile-time
template <int R, int E, bool CONJ, typename T>
__device__ __forceinline__ C2<T> rot_const(C2<T> v) {
  constexpr int e = ((E % R) + R) % R;
  if constexpr (e == 0) {
    return v;
  } else if constexpr (2 * e == R) {
    return mk<T>(-v.x, -v.y);
  } else if constexpr (4 * e == R) {
    // W^{R/4} = -i
    return CONJ ? mk<T>(-v.y, v.x) : mk<T>(v.y, -v.x);
  } else if constexpr (4 * e == 3 * R) {
    return CONJ ? mk<T>(v.y, -v.x) : mk<T>(-v.y, v.x);
  } else {
    constexpr TwTable<R> tw{};
    const T c = (T)tw.c[e];
    const T s = CONJ ? (T)tw.s[e] : (T)(-tw.s[e]);
    return mk<T>(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}

template <int R, bool CONJ, typename T, int I = 0>
__device__ __forceinline__ void rot_all(C2<T>* v) {
  // v[i] *= W_{2R}^{+-i}
  if constexpr (I < R) {
    v[I] = rot_const<2 * R, I, CONJ, T>(v[I]);
    rot_all<R, CONJ, T, I + 1>(v);
  }
}

// C: columns per block (power of two); NTH = C * R1 rounded up to a wave
template <int C, int R1>
constexpr int col_threads() {
  return (C * R1 + 63) / 64 * 64;
}

template <typename TI, typename TC, typename TO, int MODE, int R1, int R2, int NSUB, int C>
__global__ __launch_bounds__((col_threads<C, R1>())) void k_colpass(const C2<TI>* in,
                                                                  C2<TO>* out, ColArgs2 a,
                                                                  const C2<TC>* __restrict__ T2L,
                                                                  const TC* __restrict__ G) {
  using CT = C2<TC>;
  constexpr int L = R1 * R2;
  constexpr uint32_t SI = sizeof(C2<TI>), SO = sizeof(C2<TO>), SC = sizeof(CT);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CT* lds = reinterpret_cast<CT*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid & (C - 1);
  const int r = tid / C;
  // column block and batched prepare
  int cbk = a.outer_x ? (int)blockIdx.y : (int)blockIdx.x;
  int bz = (int)blockIdx.z;
  if (a.nxb > 0) {
    bz = cbk / a.nxb;
    cbk -= bz * a.nxb;
  }
  const int col = cbk * C + c;
  const bool okc = col < a.W;
  const CT zero = mk<TC>((TC)0, (TC)0);
  // outer index; the conv pass pairs fy with its reflection 2Y - fy
  int o = a.outer_x ? (int)blockIdx.x : (int)blockIdx.y;
  if (MODE == PM_CONV && a.pair) {
    const int j = o, last = (int)(a.outer_x ? gridDim.x : gridDim.y) - 1;
    const int t = (j + 1) >> 1;
    o = (j == 0 || j == last) ? (j == 0 ? 0 : a.Y2 >> 1) : ((j & 1) ? t : a.Y2 - t);
  }
  const int oh = o / a.od, ol = o - oh * a.od;
  // per-block bases; every later offset is 32-bit
  const C2<TI>* bin = in + (size_t)bz * a.in_bs + (size_t)oh * a.in_os_hi + (size_t)ol * a.in_os_lo +
                      (size_t)cbk * a.in_cbs;
  C2<TO>* bout = out + (size_t)bz * a.out_bs + (size_t)oh * a.out_os_hi + (size_t)ol * a.out_os_lo +
                 (size_t)cbk * a.out_cbs;
  const uint32_t cin = (uint32_t)c * SI, cout = (uint32_t)c * SO;

  // pass twiddles W_2L^m, m < 2L, go to LDS: read there by every phase, they
  // cost no vector-memory instructions next to the column traffic
  constexpr int NTH = col_threads<C, R1>();
  constexpr int NTW = (2 * L + NTH - 1) / NTH;
  __shared__ CT twl[2 * L];
  CT twv[NTW];
#pragma unroll
  for (int j = 0; j < NTW; ++j) twv[j] = T2L[min(tid + j * NTH, 2 * L - 1)];  // unconditional: stays in registers

  // ------------------------------------------------------------------ load
  // FWD / CONV: thread (c, n1 = r) loads x[n1 + R1 n2], n2 < R2
  CT x[MODE == PM_INV ? 1 : R2];
  if constexpr (MODE != PM_INV) {
    const bool ok = okc && r < R1;
    const uint32_t rbase = (uint32_t)r * a.in_is_lo + cin;
    const uint32_t live = (MODE == PM_FWD && a.rowmask && ok) ? a.rowmask[o * a.mask_os + r] : ~0u;
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const uint32_t off = (ok && ((live >> n2) & 1u)) ? rbase + suni((uint32_t)n2 * a.in_is_hi) : COL_OOB;
      x[n2] = cvt<TI, TC>(ldg_stream_off(bin, off));
    }
  }
  // INV: thread (c, k2 = r) loads X_s[R2 k1 + k2] of every sub-line up front
  CT xi[MODE == PM_INV ? NSUB * R1 : 1];
  if constexpr (MODE == PM_INV) {
    const bool ok = okc && r < R2;
    // row (R2 k1 + r) NSUB + s: one per-thread base and block-uniform steps
    // (a multiply per row would cost a 64-bit integer multiply-add each)
    const uint32_t rb = (uint32_t)(r * NSUB) * a.in_is_lo + cin;
    const uint32_t k1s = (uint32_t)(R2 * NSUB) * a.in_is_lo;
#pragma unroll
    for (int s = 0; s < NSUB; ++s)
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const uint32_t off = ok ? rb + suni((uint32_t)k1 * k1s + (uint32_t)s * a.in_is_lo) : COL_OOB;
        xi[s * R1 + k1] = cvt<TI, TC>(ldg_stream_off(bin, off));
      }
  }

#pragma unroll
  for (int j = 0; j < NTW; ++j)
    if (tid + j * NTH < 2 * L) twl[tid + j * NTH] = twv[j];
  __syncthreads();
  auto TW = [&](uint32_t byte_off) {
    return *reinterpret_cast<const CT*>(reinterpret_cast<const char*>(twl) + byte_off);
  };

  CT acc[MODE == PM_FWD ? 1 : R2];
  if constexpr (MODE != PM_FWD) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) acc[n2] = zero;
  }

  // conv pass: the block's kernel-table tile, C values per rz row.  The tile
  // pointer stays block-uniform and the column goes into the 32-bit offset,
  // so the loads take a scalar base (no per-thread 64-bit address arithmetic)
  const TC* Gb = nullptr;
  uint32_t gcol = 0;
  if constexpr (MODE == PM_CONV) {
    const int fy = o * a.gym + a.gya;
    const int ry = min(fy, a.Y2 - fy);
    Gb = G + ((size_t)(ry * a.npar + a.par) * a.ncb + cbk) * a.GZ1 * C;
    gcol = (uint32_t)c * (uint32_t)sizeof(TC);
  }

#pragma unroll
  for (int s = 0; s < NSUB; ++s) {
    // -------------------------------------------------------------- P1
    if constexpr (MODE != PM_INV) {
      if (r < R1) {
        CT v[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) v[n2] = x[n2];
        if (s == 1) rot_all<R2, false, TC>(v);
        rdft<R2, false, TC>(v);
        const uint32_t tstep = (uint32_t)(2 * r) * SC;
        const uint32_t tbase = (uint32_t)(s * r) * SC;
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
          CT t = v[k2];
          if (k2 > 0 || s == 1) t = cmul<TC>(t, TW(tbase + (uint32_t)k2 * tstep));
          lds[(k2 * R1 + r) * C + c] = t;
        }
      }
      __syncthreads();
    }
    // -------------------------------------------------------------- P2
    if (r < R2) {
      CT v[R1];
      if constexpr (MODE == PM_INV) {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) v[k1] = xi[s * R1 + k1];
      } else {
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) v[n1] = lds[(r * R1 + n1) * C + c];
        rdft<R1, false, TC>(v);  // v[k1] = X_s[R2 k1 + k2]
      }
      if constexpr (MODE == PM_FWD) {
        if (okc) {
          const uint32_t rb = (uint32_t)(r * NSUB + s) * a.out_is_lo + cout;
          const uint32_t k1s = (uint32_t)(R2 * NSUB) * a.out_is_lo;
#pragma unroll
          for (int k1 = 0; k1 < R1; ++k1) {
            if constexpr (sizeof(C2<TO>) == 8) stg_nt(bout, rb + suni((uint32_t)k1 * k1s), cvt<TC, TO>(v[k1]));
            else stg_off(bout, rb + suni((uint32_t)k1 * k1s), cvt<TC, TO>(v[k1]));
          }
        }
      } else {
        if constexpr (MODE == PM_CONV) {
          // fz = 2 (R2 k1 + k2) + parity, rz = min(fz, 2L - fz)
          const int q = 2 * r + (NSUB == 2 ? s : a.zpar);
          const uint32_t lo = (uint32_t)q * (C * sizeof(TC));
#pragma unroll
          for (int k1 = 0; k1 < R1; ++k1) {
            // resolved per unrolled k1: rows below L, above L, or straddling it
            constexpr uint32_t RB = C * sizeof(TC);
            uint32_t off;
            if (2 * R2 * (k1 + 1) - 1 <= L) off = lo + (uint32_t)(2 * R2 * k1) * RB;
            else if (2 * R2 * k1 >= L) off = (uint32_t)(2 * L - 2 * R2 * k1) * RB - lo;
            else {
              const int fz = 2 * R2 * k1 + q;
              off = (uint32_t)min(fz, 2 * L - fz) * RB;
            }
            v[k1] = cscale<TC>(v[k1], ldg_off(Gb, off + gcol));
          }
        }
        rdft<R1, true, TC>(v);  // v[n1] = sum_k1 X W_R1^{-n1 k1}
        const uint32_t tstep = (uint32_t)(2 * r + s) * SC;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
          CT t = v[n1];
          if (n1 > 0) t = cmulc<TC>(t, TW((uint32_t)n1 * tstep));
          lds[(r * R1 + n1) * C + c] = t;
        }
      }
    }
    if constexpr (MODE != PM_FWD) {
      __syncthreads();
      // ------------------------------------------------------------ P3
      if (r < R1) {
        CT v[R2];
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[k2] = lds[(k2 * R1 + r) * C + c];
        rdft<R2, true, TC>(v);  // v[n2] = x_s[n1 + R1 n2] / W_2L^{n1 s}
        if (s == 1) rot_all<R2, true, TC>(v);
        // the first sub-line initialises the sums (adding it to zero would
        // cost R2 complex additions the compiler may not drop)
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) acc[n2] = s == 0 ? v[n2] : cadd<TC>(acc[n2], v[n2]);
      }
    }
    if (s + 1 < NSUB) __syncthreads();
  }

  // ------------------------------------------------------------------ store
  if constexpr (MODE != PM_FWD) {
    const bool ok = okc && r < R1;
    const uint32_t rbase = (uint32_t)r * a.out_is_lo + cout;
    const uint32_t live = (MODE == PM_INV && a.rowmask && ok) ? a.rowmask[o * a.mask_os + r] : ~0u;
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const uint32_t off = (ok && ((live >> n2) & 1u)) ? rbase + suni((uint32_t)n2 * a.out_is_hi) : COL_OOB;
      stg_buf<MODE == PM_INV ? ST_YINV : ST_CONV>(bout, off, cvt<TC, TO>(acc[n2]));
    }
  }
}

// Coupled z conv of the two-coset path.  The sites of the grid lie on two
// cosets of the coarse lattice 2Z^3 (x = y = z = 0 and = 1 mod 2); each coset
// is a plane of the coarse grid, and V on coset i is the sum over cosets j of
// the open-boundary convolution of rho_j with G sampled at offsets
// 2 d + t_i - t_j.  Per frequency the two transformed columns are mixed by
//   [V0; V1] = [g0, conj(gt); gt, g0] [R0; R1]
// with g0 = the kernel within a coset (real, the conv-table tile of the
// single-plane pass, even in fy and fz) and gt = the kernel from coset 0 to
// coset 1 (complex, a tile per fy: Q[fy][column block][fz][c]).  Same
// blocks, index maps, strips and batching as the conv mode of k_colpass; the
// second coset's plane is a.cset elements after the first, in and out.
template <typename TI, typename TC, typename TO, int R1, int R2, int C>
__global__ __launch_bounds__((col_threads<C, R1>())) void k_colconv2(const C2<TI>* in, C2<TO>* out, ColArgs2 a,
                                                                   const C2<TC>* __restrict__ T2L,
                                                                   const TC* __restrict__ G,
                                                                   const C2<TC>* __restrict__ Q) {
  using CT = C2<TC>;
  constexpr int L = R1 * R2, NSUB = 2;
  constexpr uint32_t SI = sizeof(C2<TI>), SO = sizeof(C2<TO>), SC = sizeof(CT);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CT* lds0 = reinterpret_cast<CT*>(smem_raw);
  CT* lds1 = lds0 + L * C;
  const int tid = threadIdx.x;
  const int c = tid & (C - 1);
  const int r = tid / C;
  int cbk = (int)blockIdx.x;
  int bz = (int)blockIdx.z;
  if (a.nxb > 0) {
    bz = cbk / a.nxb;
    cbk -= bz * a.nxb;
  }
  const int col = cbk * C + c;
  const bool okc = col < a.W;
  // fy and its reflection 2Y - fy in adjacent blocks (same table tiles)
  int o = (int)blockIdx.y;
  if (a.pair) {
    const int j = o, last = (int)gridDim.y - 1;
    const int t = (j + 1) >> 1;
    o = (j == 0 || j == last) ? (j == 0 ? 0 : a.Y2 >> 1) : ((j & 1) ? t : a.Y2 - t);
  }
  const C2<TI>* bin0 = in + (size_t)bz * a.in_bs + (size_t)o * a.in_os_hi + (size_t)cbk * a.in_cbs;
  const C2<TI>* bin1 = bin0 + a.cset;
  C2<TO>* bout0 = out + (size_t)bz * a.out_bs + (size_t)o * a.out_os_hi + (size_t)cbk * a.out_cbs;
  C2<TO>* bout1 = bout0 + a.cset;
  const uint32_t cin = (uint32_t)c * SI, cout = (uint32_t)c * SO;

  constexpr int NTH = col_threads<C, R1>();
  constexpr int NTW = (2 * L + NTH - 1) / NTH;
  __shared__ CT twl[2 * L];
  CT twv[NTW];
#pragma unroll
  for (int j = 0; j < NTW; ++j) twv[j] = T2L[min(tid + j * NTH, 2 * L - 1)];

  // thread (c, n1 = r) loads x_i[n1 + R1 n2], n2 < R2, of both cosets
  CT x0[R2], x1[R2];
  {
    const bool ok = okc && r < R1;
    const uint32_t rbase = (uint32_t)r * a.in_is_lo + cin;
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const uint32_t off = ok ? rbase + suni((uint32_t)n2 * a.in_is_hi) : COL_OOB;
      x0[n2] = cvt<TI, TC>(ldg_stream_off(bin0, off));
      x1[n2] = cvt<TI, TC>(ldg_stream_off(bin1, off));
    }
  }
#pragma unroll
  for (int j = 0; j < NTW; ++j)
    if (tid + j * NTH < 2 * L) twl[tid + j * NTH] = twv[j];
  __syncthreads();
  auto TW = [&](uint32_t byte_off) {
    return *reinterpret_cast<const CT*>(reinterpret_cast<const char*>(twl) + byte_off);
  };

  CT acc0[R2], acc1[R2];
  // table tiles of the block (block-uniform bases, the column in the offset)
  const int fy = o * a.gym + a.gya;
  const int ry = min(fy, a.Y2 - fy);
  const TC* Gb = G + ((size_t)(ry * a.npar + a.par) * a.ncb + cbk) * a.GZ1 * C;
  const CT* Qb = Q + ((size_t)fy * a.ncb + cbk) * (size_t)(2 * L) * C;
  const uint32_t gcol = (uint32_t)c * (uint32_t)sizeof(TC), qcol = (uint32_t)c * SC;

#pragma unroll
  for (int s = 0; s < NSUB; ++s) {
    // -------------------------------------------------------------- P1
    if (r < R1) {
      CT v[R2], w[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        v[n2] = x0[n2];
        w[n2] = x1[n2];
      }
      if (s == 1) {
        rot_all<R2, false, TC>(v);
        rot_all<R2, false, TC>(w);
      }
      rdft<R2, false, TC>(v);
      rdft<R2, false, TC>(w);
      const uint32_t tstep = (uint32_t)(2 * r) * SC;
      const uint32_t tbase = (uint32_t)(s * r) * SC;
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        CT t0 = v[k2], t1 = w[k2];
        if (k2 > 0 || s == 1) {
          const CT tw = TW(tbase + (uint32_t)k2 * tstep);
          t0 = cmul<TC>(t0, tw);
          t1 = cmul<TC>(t1, tw);
        }
        lds0[(k2 * R1 + r) * C + c] = t0;
        lds1[(k2 * R1 + r) * C + c] = t1;
      }
    }
    __syncthreads();
    // -------------------------------------------------------------- P2
    if (r < R2) {
      // the thread's table values first: their loads are in flight during
      // the LDS reads and the forward transforms (issued one pair at a time
      // next to their use, each waited for in turn)
      // fz = 2 (R2 k1 + k2) + s; g0 row rz = min(fz, 2L - fz), Q row fz
      const int q = 2 * r + s;
      const uint32_t lo = (uint32_t)q * (C * sizeof(TC));
      TC gv[R1];
      CT qv[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        constexpr uint32_t RB = C * sizeof(TC);
        uint32_t off;
        if (2 * R2 * (k1 + 1) - 1 <= L) off = lo + (uint32_t)(2 * R2 * k1) * RB;
        else if (2 * R2 * k1 >= L) off = (uint32_t)(2 * L - 2 * R2 * k1) * RB - lo;
        else {
          const int fz = 2 * R2 * k1 + q;
          off = (uint32_t)min(fz, 2 * L - fz) * RB;
        }
        gv[k1] = ldg_off(Gb, off + gcol);
        qv[k1] = ldg_off(Qb, (uint32_t)(2 * R2 * k1 + q) * (C * SC) + qcol);
      }
      CT v[R1], w[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) {
        v[n1] = lds0[(r * R1 + n1) * C + c];
        w[n1] = lds1[(r * R1 + n1) * C + c];
      }
      rdft<R1, false, TC>(v);  // v[k1] = R0_s[R2 k1 + k2]
      rdft<R1, false, TC>(w);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const TC g = gv[k1];
        const CT gt = qv[k1];
        const CT a0 = v[k1], a1 = w[k1];
        v[k1] = mk<TC>(g * a0.x + gt.x * a1.x + gt.y * a1.y, g * a0.y + gt.x * a1.y - gt.y * a1.x);
        w[k1] = mk<TC>(gt.x * a0.x - gt.y * a0.y + g * a1.x, gt.x * a0.y + gt.y * a0.x + g * a1.y);
      }
      rdft<R1, true, TC>(v);
      rdft<R1, true, TC>(w);
      const uint32_t tstep = (uint32_t)(2 * r + s) * SC;
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) {
        CT t0 = v[n1], t1 = w[n1];
        if (n1 > 0) {
          const CT tw = TW((uint32_t)n1 * tstep);
          t0 = cmulc<TC>(t0, tw);
          t1 = cmulc<TC>(t1, tw);
        }
        lds0[(r * R1 + n1) * C + c] = t0;
        lds1[(r * R1 + n1) * C + c] = t1;
      }
    }
    __syncthreads();
    // -------------------------------------------------------------- P3
    if (r < R1) {
      CT v[R2], w[R2];
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        v[k2] = lds0[(k2 * R1 + r) * C + c];
        w[k2] = lds1[(k2 * R1 + r) * C + c];
      }
      rdft<R2, true, TC>(v);
      rdft<R2, true, TC>(w);
      if (s == 1) {
        rot_all<R2, true, TC>(v);
        rot_all<R2, true, TC>(w);
      }
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        acc0[n2] = s == 0 ? v[n2] : cadd<TC>(acc0[n2], v[n2]);
        acc1[n2] = s == 0 ? w[n2] : cadd<TC>(acc1[n2], w[n2]);
      }
    }
    if (s + 1 < NSUB) __syncthreads();
  }

  const bool ok = okc && r < R1;
  const uint32_t rbase = (uint32_t)r * a.out_is_lo + cout;
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const uint32_t off = ok ? rbase + suni((uint32_t)n2 * a.out_is_hi) : COL_OOB;
    stg_buf<ST_CONV>(bout0, off, cvt<TC, TO>(acc0[n2]));
    stg_buf<ST_CONV>(bout1, off, cvt<TC, TO>(acc1[n2]));
  }
}

// columns per block for a pass of length L: 16 (128-byte rows of complex64);
// complex128 passes of up to 256 points take 16 as well (256-byte rows: the
// FP64 y passes of the 1M-atom grid run at 5.4-6.2 TB/s against 4.4-5.2 with
// 8 columns, the whole FP64 step 18.4 against 19.4 ms), longer ones 8 (their
// LDS tile would hold a block per CU)
__host__ __device__ constexpr int col_c_bytes(size_t c2bytes, int L) {
  return c2bytes == 8 ? 16 : (L <= 256 ? 16 : 8);
}
template <typename TC>
constexpr int col_c(int L) {
  return col_c_bytes(sizeof(C2<TC>), L);
}

struct ColLaunch {
  dim3 grid;
  hipStream_t stream;
  const void* in;
  void* out;
  ColArgs2 a;
  const void* T2L;
  const void* G;
  const void* Q;   // coupled conv: the cross-coset table
};

template <int R1, int R2, typename TC>
constexpr size_t col_lds() {
  return (size_t)R1 * R2 * col_c<TC>(R1 * R2) * sizeof(C2<TC>);
}

template <typename TI, typename TC, typename TO, int MODE, int R1, int R2, int NSUB>
hipError_t col_go(const ColLaunch& L) {
  constexpr int C = col_c<TC>(R1 * R2);
  constexpr int nth = col_threads<C, R1>();
  constexpr size_t lds = col_lds<R1, R2, TC>();
  hipLaunchKernelGGL((k_colpass<TI, TC, TO, MODE, R1, R2, NSUB, C>), L.grid, dim3(nth), lds, L.stream, (const C2<TI>*)L.in, (C2<TO>*)L.out, L.a,
                     (const C2<TC>*)L.T2L, (const TC*)L.G);
  return hipGetLastError();
}

// LDS of the coupled conv: one exchange tile per coset
template <int R1, int R2, typename TC>
constexpr size_t conv2_lds() {
  return 2 * col_lds<R1, R2, TC>();
}

template <typename TI, typename TC, typename TO, int R1, int R2>
hipError_t conv2_go(const ColLaunch& L) {
  constexpr int C = col_c<TC>(R1 * R2);
  constexpr int nth = col_threads<C, R1>();
  constexpr size_t lds = conv2_lds<R1, R2, TC>();
  if constexpr (lds + 2 * R1 * R2 * sizeof(C2<TC>) > 65536) {
    return hipErrorInvalidValue;  // the host keeps such grids on the single-plane path
  } else {
    hipLaunchKernelGGL((k_colconv2<TI, TC, TO, R1, R2, C>), L.grid, dim3(nth), lds, L.stream, (const C2<TI>*)L.in,
                       (C2<TO>*)L.out, L.a, (const C2<TC>*)L.T2L, (const TC*)L.G, (const C2<TC>*)L.Q);
    return hipGetLastError();
  }
}

template <typename TI, typename TC, typename TO, int MODE, int R1, int R2>
hipError_t col_go_n(int nsub, const ColLaunch& L) {
  return nsub == 2 ? col_go<TI, TC, TO, MODE, R1, R2, 2>(L) : col_go<TI, TC, TO, MODE, R1, R2, 1>(L);
}

// (types, mode) pairs of the column passes; hipErrorInvalidValue otherwise
template <int R1, int R2>
hipError_t col_launch(int combo, int mode, int nsub, const ColLaunch& L) {
  using d = double;
  using f = float;
  switch (combo) {
    case CB_DDD:
      if (mode == PM_FWD) return col_go_n<d, d, d, PM_FWD, R1, R2>(nsub, L);
      if (mode == PM_CONV) return col_go_n<d, d, d, PM_CONV, R1, R2>(nsub, L);
      if (mode == PM_CONV2) return conv2_go<d, d, d, R1, R2>(L);
      return col_go_n<d, d, d, PM_INV, R1, R2>(nsub, L);
    case CB_FFF:
      if (mode == PM_FWD) return col_go_n<f, f, f, PM_FWD, R1, R2>(nsub, L);
      if (mode == PM_CONV) return col_go_n<f, f, f, PM_CONV, R1, R2>(nsub, L);
      if (mode == PM_CONV2) return conv2_go<f, f, f, R1, R2>(L);
      return col_go_n<f, f, f, PM_INV, R1, R2>(nsub, L);
    case CB_FDF:
      if (mode == PM_FWD) return col_go_n<f, d, f, PM_FWD, R1, R2>(nsub, L);
      if (mode == PM_CONV) return col_go_n<f, d, f, PM_CONV, R1, R2>(nsub, L);
      if (mode == PM_CONV2) return conv2_go<f, d, f, R1, R2>(L);
      break;
    case CB_FDD:
      if (mode == PM_INV) return col_go_n<f, d, d, PM_INV, R1, R2>(nsub, L);
      break;
    default:
      break;
  }
  return hipErrorInvalidValue;
}

#define CLAT_C2_DECL(a, b) hipError_t clat_c2_##a##x##b(int combo, int mode, int nsub, const ColLaunch& L);
CLAT_P2_PAIRS(CLAT_C2_DECL)

}  // namespace clat
