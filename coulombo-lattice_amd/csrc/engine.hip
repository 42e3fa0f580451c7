// The C-ABI engine: device buffers, FFT plans, launch sequences of
// prepare()/calculate() and the batched integral path.  Built for gfx950.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/clat.h"
#include "host_logic.hpp"
#include "kernels.hpp"

using namespace clat;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace {

thread_local std::string g_err;

int set_err(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

struct HipError {
  int code;
  std::string msg;
};

#define HIPX(call)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw HipError{e_ == hipErrorOutOfMemory ? CLAT_ENOMEM : CLAT_EDEVICE,             \
                     std::string(#call) + ": " + hipGetErrorString(e_)};                 \
  } while (0)

#define LAUNCH_CHECK() HIPX(hipGetLastError())

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return CLAT_OK;
  } catch (const HipError& e) {
    return set_err(e.code, "%s", e.msg.c_str());
  } catch (const host::Error& e) {
    return set_err(e.code, "%s", e.what());
  } catch (const std::bad_alloc&) {
    return set_err(CLAT_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return set_err(CLAT_EINVAL, "%s", e.what());
  }
}

// ---------------------------------------------------------------------------
// plans and tables
// ---------------------------------------------------------------------------
uint32_t magic_of(uint32_t d) {
  return d <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + d - 1) / d);
}

bool make_plan(int L, AxisPlan& p) {
  std::memset(&p, 0, sizeof p);
  p.L = L;
  int rem = L;
  std::vector<int> r;
  while (rem % 8 == 0) { r.push_back(8); rem /= 8; }
  while (rem % 4 == 0) { r.push_back(4); rem /= 4; }
  while (rem % 2 == 0) { r.push_back(2); rem /= 2; }
  while (rem % 5 == 0) { r.push_back(5); rem /= 5; }
  while (rem % 3 == 0) { r.push_back(3); rem /= 3; }
  while (rem % 7 == 0) { r.push_back(7); rem /= 7; }
  if (rem != 1 || (int)r.size() > MAX_STAGES) return false;
  int ns = 1;
  p.nst = (int)r.size();
  for (int s = 0; s < p.nst; ++s) {
    p.radix[s] = r[s];
    p.ns[s] = ns;
    p.ns_magic[s] = magic_of((uint32_t)ns);
    p.lr_magic[s] = magic_of((uint32_t)(L / r[s]));
    ns *= r[s];
  }
  p.l_magic = magic_of((uint32_t)L);
  p.l2_magic = magic_of((uint32_t)(2 * L));
  return true;
}

// e^{-2 pi i m / (den)} for m in [0, n), from long double trigonometry
std::vector<double> unit_roots(int n, int den) {
  std::vector<double> v(2 * (size_t)n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; ++m) {
    const long double a = -2.0L * pi * (long double)m / (long double)den;
    v[2 * m] = (double)cosl(a);
    v[2 * m + 1] = (double)sinl(a);
  }
  return v;
}

int floor_pow2(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}
int log2i(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

enum Kind { K_XFWD = 0, K_YFWD, K_ZCONV, K_YINV, K_XINV, K_SCATTER, K_GATHER, K_PRODUCTS,
            K_OVERLAP, K_DCT, K_COUNT };
const char* kKindNames[K_COUNT] = {"xfwd", "yfwd", "zconv", "yinv", "xinv", "scatter",
                                   "gather", "products", "overlap", "dct1"};

unsigned grid1d(size_t n, int bs = 256) {
  size_t b = (n + bs - 1) / bs;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (unsigned)b;
}

}  // namespace

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
struct clat_engine {
  int X = 0, Y = 0, Z = 0;     // user grid
  int Xi = 0, Yi = 0, Zi = 0;  // engine grid (2*3*5*7-smooth)
  int prec = 0, sched = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  AxisPlan px{}, py{}, pz{};   // line transforms
  AxisPlan dx{}, dy{}, dz{};   // DCT-I via 2M-point transforms
  size_t csz = 16;             // bytes per complex of the pass precision
  int tile_budget = 0;         // max elements of one LDS tile
  // device tables
  void *tw[3] = {nullptr, nullptr, nullptr}, *half[3] = {nullptr, nullptr, nullptr};
  double2* dtw[3] = {nullptr, nullptr, nullptr};
  double* G64 = nullptr;  // Ghat, FP64
  float* G32 = nullptr;   // Ghat, FP32 copy for the complex64 passes
  bool initialized = false;
  // grids
  void *F = nullptr, *V = nullptr, *A = nullptr, *B = nullptr;
  double2* Fin = nullptr;  // user-layout FP64 copy of the input (calculate())
  bool input_set = false, prepared = false;
  double* dpart = nullptr;
  size_t dpart_cap = 0;
  // atoms and states
  int64_t n_atoms = 0;
  int n_orb = 0;
  int32_t* d_cell = nullptr;       // all atoms
  int64_t n_sites = 0;             // effective atoms (last atom of each cell)
  int32_t* d_site_atom = nullptr;
  int32_t* d_site_cell = nullptr;
  int n_states = 0, cap_states = 0;
  double2* coeffs = nullptr;
  double2* rho = nullptr;
  int rho_states = -1;             // states the pair densities were built for
  double2* Vat = nullptr;
  size_t vat_cap = 0;
  double* part = nullptr;
  size_t part_cap = 0;
  double* ovl = nullptr;
  size_t ovl_cap = 0;
  int64_t device_bytes = 0;
  // timing
  hipEvent_t stage_ev[6] = {};
  double stage_ms[5] = {0, 0, 0, 0, 0};
  bool ktiming = false;
  std::vector<hipEvent_t> kev;
  std::vector<std::pair<int, double>> kpending;  // kind, bytes per pending launch
  size_t kused = 0;
  int64_t kcount[K_COUNT] = {};
  double kms[K_COUNT] = {};
  double kbytes[K_COUNT] = {};

  void* dalloc(size_t bytes) {
    void* p = nullptr;
    HIPX(hipMalloc(&p, bytes ? bytes : 16));
    device_bytes += (int64_t)bytes;
    return p;
  }
  void dfree(void*& p, size_t bytes) {
    if (p) {
      (void)hipFree(p);
      device_bytes -= (int64_t)bytes;
      p = nullptr;
    }
  }
  size_t N() const { return (size_t)Xi * Yi * Zi; }
  size_t NG() const { return (size_t)(Xi + 1) * (Yi + 1) * (Zi + 1); }
  size_t NU() const { return (size_t)X * Y * Z; }
  size_t a_elems() const { return sched == CLAT_SCHED_SHARED ? 2 * N() : N(); }
  size_t b_elems() const { return sched == CLAT_SCHED_SHARED ? 4 * N() : 0; }

  // kernel timing around one launch
  void kbegin() {
    if (!ktiming) return;
    ensure_events(kused + 2);
    HIPX(hipEventRecord(kev[kused], stream));
  }
  void kend(int kind, double bytes) {
    if (!ktiming) return;
    HIPX(hipEventRecord(kev[kused + 1], stream));
    kpending.push_back({kind, bytes});
    kused += 2;
  }
  void ensure_events(size_t n) {
    while (kev.size() < n) {
      hipEvent_t ev;
      HIPX(hipEventCreate(&ev));
      kev.push_back(ev);
    }
  }
  void kcollect() {
    if (!ktiming) return;
    HIPX(hipStreamSynchronize(stream));
    for (size_t i = 0; i < kpending.size(); ++i) {
      float ms = 0.f;
      HIPX(hipEventElapsedTime(&ms, kev[2 * i], kev[2 * i + 1]));
      const int k = kpending[i].first;
      kcount[k] += 1;
      kms[k] += ms;
      kbytes[k] += kpending[i].second;
    }
    kpending.clear();
    kused = 0;
  }
};

namespace {

int tile_lines(const clat_engine* e, int nsub, int L, int cap = 64) {
  const int c = floor_pow2(std::max(1, e->tile_budget / (nsub * L)));
  return std::min(c, cap);
}

// ---------------------------------------------------------------------------
// prepare(): input F -> potential V
// ---------------------------------------------------------------------------
template <typename T, int EMAX>
void run_prepare(clat_engine* e) {
  using CT = C2<T>;
  const int Xi = e->Xi, Yi = e->Yi, Zi = e->Zi;
  const size_t N = e->N();
  const double s = (double)sizeof(CT);
  const CT* twx = (const CT*)e->tw[0];
  const CT* twy = (const CT*)e->tw[1];
  const CT* twz = (const CT*)e->tw[2];
  const CT* hx = (const CT*)e->half[0];
  const CT* hy = (const CT*)e->half[1];
  const CT* hz = (const CT*)e->half[2];
  const T* G = (sizeof(T) == 4) ? (const T*)(const void*)e->G32 : (const T*)(const void*)e->G64;
  CT* F = (CT*)e->F;
  CT* V = (CT*)e->V;
  CT* A = (CT*)e->A;
  CT* B = (CT*)e->B;
  const int nlines = Yi * Zi;

  auto xfwd = [&](int nsub, int kx, int ky, int kz, CT* out) {
    const int C = tile_lines(e, nsub, Xi);
    const size_t lds = (size_t)nsub * C * Xi * sizeof(CT);
    dim3 grid((nlines + C - 1) / C);
    e->kbegin();
    hipLaunchKernelGGL((k_xfwd<T, EMAX>), grid, dim3(NT), lds, e->stream, F, out, e->px, twx, hx, nlines, C,
                       log2i(C), nsub, kx, ky, kz, hy, hz, Yi);
    LAUNCH_CHECK();
    e->kend(K_XFWD, s * (double)N * (1 + nsub));
  };
  auto xinv = [&](int nsub, int kx, int ky, int kz, int acc, const CT* in) {
    const int C = tile_lines(e, nsub, Xi);
    const size_t lds = (size_t)nsub * C * Xi * sizeof(CT);
    dim3 grid((nlines + C - 1) / C);
    e->kbegin();
    hipLaunchKernelGGL((k_xinv<T, EMAX>), grid, dim3(NT), lds, e->stream, in, V, e->px, twx, hx, nlines, C,
                       log2i(C), nsub, kx, ky, kz, hy, hz, Yi, acc);
    LAUNCH_CHECK();
    e->kend(K_XINV, s * (double)N * (1 + nsub + (acc ? 1 : 0)));
  };
  auto col = [&](int mode, const AxisPlan& p, const CT* tw, const CT* h, ColArgs a, int n_outer,
                 const CT* in, CT* out) {
    const int L = p.L;
    a.C = tile_lines(e, a.nsub, L, 64);
    a.C = std::min(a.C, std::max(1, floor_pow2(a.W) * (a.W > floor_pow2(a.W) ? 2 : 1)));
    a.lgC = log2i(a.C);
    const size_t lds = (size_t)a.nsub * a.C * L * sizeof(CT);
    dim3 grid((a.W + a.C - 1) / a.C, n_outer);
    e->kbegin();
    int kind;
    double bytes;
    if (mode == COL_FWD) {
      hipLaunchKernelGGL((k_col<T, EMAX, COL_FWD>), grid, dim3(NT), lds, e->stream, in, out, p, tw, h, a, G);
      kind = K_YFWD;
      bytes = s * (double)N * (a.nsub == 2 ? 6.0 : 2.0);
    } else if (mode == COL_CONV) {
      hipLaunchKernelGGL((k_col<T, EMAX, COL_CONV>), grid, dim3(NT), lds, e->stream, in, out, p, tw, h, a, G);
      kind = K_ZCONV;
      bytes = s * (double)N * (a.nsub == 2 ? 8.0 : 2.0) + (double)sizeof(T) * (double)e->NG();
    } else {
      hipLaunchKernelGGL((k_col<T, EMAX, COL_INV>), grid, dim3(NT), lds, e->stream, in, out, p, tw, h, a, G);
      kind = K_YINV;
      bytes = s * (double)N * (a.nsub == 2 ? 6.0 : 2.0);
    }
    LAUNCH_CHECK();
    e->kend(kind, bytes);
  };

  ColArgs base{};
  base.X2 = 2 * Xi;
  base.Y2 = 2 * Yi;
  base.Z2 = 2 * Zi;
  base.GX1 = Xi + 1;
  base.GY1 = Yi + 1;

  if (e->sched == CLAT_SCHED_SHARED) {
    const long long W = 2LL * Xi;
    xfwd(2, 0, 0, 0, A);
    ColArgs a = base;
    a.W = (int)W;
    a.nsub = 2;
    a.in_os = (long long)Yi * W; a.in_is = W; a.out_os = 2LL * Yi * W; a.out_is = W;
    col(COL_FWD, e->py, twy, hy, a, Zi, A, B);
    ColArgs c = base;
    c.W = (int)W;
    c.nsub = 2;
    c.in_os = c.out_os = W; c.in_is = c.out_is = 2LL * Yi * W;
    c.gxm = 1; c.gxa = 0; c.gym = 1; c.gya = 0;
    col(COL_CONV, e->pz, twz, hz, c, 2 * Yi, B, B);
    ColArgs b = base;
    b.W = (int)W;
    b.nsub = 2;
    b.in_os = 2LL * Yi * W; b.in_is = W; b.out_os = (long long)Yi * W; b.out_is = W;
    col(COL_INV, e->py, twy, hy, b, Zi, B, A);
    xinv(2, 0, 0, 0, 0, A);
  } else {
    const long long W = Xi;
    for (int round = 0; round < 8; ++round) {
      const int kx = round & 1, ky = (round >> 1) & 1, kz = (round >> 2) & 1;
      xfwd(1, kx, ky, kz, A);
      ColArgs a = base;
      a.W = (int)W;
      a.nsub = 1;
      a.in_os = a.out_os = (long long)Yi * W; a.in_is = a.out_is = W;
      col(COL_FWD, e->py, twy, hy, a, Zi, A, A);
      ColArgs c = base;
      c.W = (int)W;
      c.nsub = 1;
      c.in_os = c.out_os = W; c.in_is = c.out_is = (long long)Yi * W;
      c.zpar = kz; c.gxm = 2; c.gxa = kx; c.gym = 2; c.gya = ky;
      col(COL_CONV, e->pz, twz, hz, c, Yi, A, A);
      col(COL_INV, e->py, twy, hy, a, Zi, A, A);
      xinv(1, kx, ky, kz, round > 0 ? 1 : 0, A);
    }
  }
}

void prepare(clat_engine* e) {
  if (e->prec == CLAT_PREC_FP32) run_prepare<float, 32>(e);
  else run_prepare<double, 16>(e);
}

// zero the engine grid F
void zero_F(clat_engine* e) {
  if (e->prec == CLAT_PREC_FP32)
    hipLaunchKernelGGL(k_zero<float>, dim3(grid1d(e->N())), dim3(256), 0, e->stream, (float2*)e->F, e->N());
  else
    hipLaunchKernelGGL(k_zero<double>, dim3(grid1d(e->N())), dim3(256), 0, e->stream, (double2*)e->F, e->N());
  LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// interaction table: G -> DCT-I -> /(8 Xi Yi Zi)
// ---------------------------------------------------------------------------
void dct_axis(clat_engine* e, int axis, double scale) {
  const int X1 = e->Xi + 1, Y1 = e->Yi + 1, Z1 = e->Zi + 1;
  const AxisPlan& p = axis == 0 ? e->dx : (axis == 1 ? e->dy : e->dz);
  const int L2 = p.L;
  const int C = std::min(64, floor_pow2(std::max(1, 4096 / L2)));
  const size_t lds = (size_t)C * L2 * sizeof(double2);
  if (axis == 0) {
    const int ncols = Y1 * Z1;
    dim3 grid((ncols + C - 1) / C);
    hipLaunchKernelGGL((k_dct1<16, true>), grid, dim3(NT), lds, e->stream, e->G64, p, e->dtw[0], ncols, C,
                       log2i(C), 0LL, 1LL, (long long)X1, ncols, scale);
  } else if (axis == 1) {
    const int ncols = X1;
    dim3 grid(((ncols + C - 1) / C) * Z1);
    hipLaunchKernelGGL((k_dct1<16, false>), grid, dim3(NT), lds, e->stream, e->G64, p, e->dtw[1], ncols, C,
                       log2i(C), (long long)X1 * Y1, (long long)X1, 1LL, ncols, scale);
  } else {
    const int ncols = X1 * Y1;
    dim3 grid((ncols + C - 1) / C);
    hipLaunchKernelGGL((k_dct1<16, false>), grid, dim3(NT), lds, e->stream, e->G64, p, e->dtw[2], ncols, C,
                       log2i(C), 0LL, (long long)X1 * Y1, 1LL, ncols, scale);
  }
  LAUNCH_CHECK();
}

void finish_interaction(clat_engine* e) {
  dct_axis(e, 0, 1.0);
  dct_axis(e, 1, 1.0);
  dct_axis(e, 2, 1.0 / (8.0 * (double)e->Xi * (double)e->Yi * (double)e->Zi));
  if (e->prec == CLAT_PREC_FP32) {
    hipLaunchKernelGGL(k_d2f, dim3(grid1d(e->NG())), dim3(256), 0, e->stream, e->G64, e->G32, e->NG());
    LAUNCH_CHECK();
  }
  e->initialized = true;
  e->prepared = false;
}

GArgs base_gargs(const clat_engine* e) {
  GArgs g{};
  g.X1 = e->Xi + 1; g.Y1 = e->Yi + 1; g.Z1 = e->Zi + 1;
  g.ux = e->X + 1; g.uy = e->Y + 1; g.uz = e->Z + 1;
  return g;
}

// G table + DCT-I on the engine stream, timed into stage_ms[0]
void build_table(clat_engine* e, GArgs g) {
  HIPX(hipEventRecord(e->stage_ev[0], e->stream));
  hipLaunchKernelGGL(k_gtable, dim3(grid1d(e->NG())), dim3(256), 0, e->stream, e->G64, g);
  LAUNCH_CHECK();
  finish_interaction(e);
  HIPX(hipEventRecord(e->stage_ev[5], e->stream));
  HIPX(hipStreamSynchronize(e->stream));
  float ms = 0.f;
  HIPX(hipEventElapsedTime(&ms, e->stage_ev[0], e->stage_ev[5]));
  e->stage_ms[0] = ms;
}

void ensure_dpart(clat_engine* e, size_t n) {
  if (e->dpart_cap < n) {
    void* p = e->dpart;
    e->dfree(p, e->dpart_cap * sizeof(double));
    e->dpart = (double*)e->dalloc(n * sizeof(double));
    e->dpart_cap = n;
  }
}

// Neumaier sum on the host
double nsum(const double* v, size_t n, size_t stride) {
  double s = 0.0, c = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double x = v[i * stride];
    const double t = s + x;
    c += (std::fabs(s) >= std::fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  return s + c;
}

// ---------------------------------------------------------------------------
// atom-site path helpers
// ---------------------------------------------------------------------------
void build_products(clat_engine* e) {
  const int P = e->n_states * (e->n_states + 1) / 2;
  if (e->rho_states == e->n_states && e->rho) return;
  void* r = e->rho;
  e->dfree(r, (size_t)(e->rho_states > 0 ? e->rho_states * (e->rho_states + 1) / 2 : 0) * e->n_sites *
                  sizeof(double2));
  e->rho = nullptr;
  e->rho = (double2*)e->dalloc((size_t)P * e->n_sites * sizeof(double2));
  e->rho_states = e->n_states;
  const size_t per_site = (size_t)e->n_states * e->n_orb * sizeof(double2);
  int A = (int)std::min<size_t>(64, 49152 / per_site);
  e->kbegin();
  if (A >= 1) {
    const unsigned nb = (unsigned)((e->n_sites + A - 1) / A);
    hipLaunchKernelGGL(k_products, dim3(nb), dim3(NT), (size_t)A * per_site, e->stream, e->coeffs,
                       e->d_site_atom, e->n_sites, e->n_atoms, e->n_states, e->n_orb, A, e->rho);
  } else {
    hipLaunchKernelGGL(k_products_direct, dim3(grid1d((size_t)P * e->n_sites)), dim3(256), 0, e->stream,
                       e->coeffs, e->d_site_atom, e->n_sites, e->n_atoms, e->n_states, e->n_orb, e->rho);
  }
  LAUNCH_CHECK();
  e->kend(K_PRODUCTS, (double)e->n_sites * (double)per_site + (double)P * e->n_sites * 16.0);
}

// F <- rho of product p at every site (grid zeroed first)
void scatter_product(clat_engine* e, int p) {
  zero_F(e);
  const double2* r = e->rho + (size_t)p * e->n_sites;
  e->kbegin();
  if (e->prec == CLAT_PREC_FP32)
    hipLaunchKernelGGL(k_scatter<float>, dim3(grid1d(e->n_sites)), dim3(256), 0, e->stream, (float2*)e->F,
                       e->d_site_cell, r, e->n_sites);
  else
    hipLaunchKernelGGL(k_scatter<double>, dim3(grid1d(e->n_sites)), dim3(256), 0, e->stream, (double2*)e->F,
                       e->d_site_cell, r, e->n_sites);
  LAUNCH_CHECK();
  e->kend(K_SCATTER, (double)e->n_sites * (16.0 + 4.0 + (double)e->csz));
}

void gather_V(clat_engine* e, const int32_t* cells, int64_t n, double2* out) {
  e->kbegin();
  if (e->prec == CLAT_PREC_FP32)
    hipLaunchKernelGGL(k_gather<float>, dim3(grid1d(n)), dim3(256), 0, e->stream, (const float2*)e->V, cells,
                       out, n);
  else
    hipLaunchKernelGGL(k_gather<double>, dim3(grid1d(n)), dim3(256), 0, e->stream, (const double2*)e->V, cells,
                       out, n);
  LAUNCH_CHECK();
  e->kend(K_GATHER, (double)n * (16.0 + 4.0 + (double)e->csz));
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* clat_last_error(void) { return g_err.c_str(); }

const char* clat_version(void) { return "clat 1.0 (gfx950)"; }

int clat_round_up(int n, int c, const int* primes, int n_primes, int* out) {
  return guarded([&] {
    if (!primes || n_primes <= 0 || !out) throw host::Error(CLAT_EINVAL, "invalid arguments");
    *out = host::round_up(n, c, std::vector<int>(primes, primes + n_primes));
  });
}

int clat_lattice_grid(const double* xyz, int64_t n, int dim_raw[3], int dim_pad[3], double origin[3],
                      double step[3]) {
  return guarded([&] {
    if (!xyz || n <= 0) throw host::Error(CLAT_EINVAL, "no atoms");
    host::Lattice L = host::detect_lattice(xyz, n);
    for (int d = 0; d < 3; ++d) {
      dim_raw[d] = L.raw[d];
      dim_pad[d] = L.pad[d];
      origin[d] = L.origin[d];
      step[d] = L.step[d];
    }
  });
}

int clat_lattice_cells(const double* xyz, int64_t n, const int dim[3], const double origin[3],
                       const double step[3], int32_t* cell) {
  return guarded([&] {
    if (!xyz || !cell || n < 0) throw host::Error(CLAT_EINVAL, "invalid arguments");
    host::cell_indices(xyz, n, dim, origin, step, cell);
  });
}

int clat_pattern_enumerate(const char* spec, int n_states, int n_holes, int32_t* quads, int64_t cap,
                           int64_t* count) {
  return guarded([&] {
    if (!spec || n_states < 0) throw host::Error(CLAT_EINVAL, "invalid arguments");
    host::Pattern pat(spec, n_holes);
    std::vector<int32_t> q = host::enumerate_integrals(pat, n_states);
    const int64_t n = (int64_t)q.size() / 4;
    if (count) *count = n;
    if (quads) std::memcpy(quads, q.data(), (size_t)std::min(n, cap) * 4 * sizeof(int32_t));
  });
}

int clat_vertex_cover(int nv, const int32_t* edges, int64_t ne, int32_t* cover, int* n_cover) {
  return guarded([&] {
    if (nv < 0 || (ne > 0 && !edges) || !cover || !n_cover) throw host::Error(CLAT_EINVAL, "invalid arguments");
    std::set<std::pair<int, int>> es;
    for (int64_t k = 0; k < ne; ++k) {
      const int u = edges[2 * k], v = edges[2 * k + 1];
      if (u < 0 || v < 0 || u >= nv || v >= nv) throw host::Error(CLAT_EINVAL, "edge vertex out of range");
      es.insert({u, v});
    }
    std::vector<int> c = host::vertex_cover(nv, es);
    for (size_t i = 0; i < c.size(); ++i) cover[i] = c[i];
    *n_cover = (int)c.size();
  });
}

int clat_plan(const int32_t* quads, int64_t n, int n_states, int32_t* steps) {
  return guarded([&] {
    if ((n > 0 && !quads) || !steps || n < 0) throw host::Error(CLAT_EINVAL, "invalid arguments");
    std::vector<host::PlanStep> st = host::make_plan(quads, n, n_states);
    for (size_t i = 0; i < st.size(); ++i) {
      int32_t* r = steps + 5 * i;
      r[0] = st[i].id;
      r[1] = st[i].left.product;
      r[2] = st[i].left.conj;
      r[3] = st[i].right.product;
      r[4] = st[i].right.conj;
    }
  });
}

int clat_create(int X, int Y, int Z, int precision, int schedule, clat_engine** out) {
  if (!out) return set_err(CLAT_EINVAL, "out is NULL");
  *out = nullptr;
  clat_engine* e = new (std::nothrow) clat_engine();
  if (!e) return set_err(CLAT_ENOMEM, "host allocation failed");
  int rc = guarded([&] {
    if (X <= 0 || Y <= 0 || Z <= 0) throw host::Error(CLAT_EINVAL, "dimensions must be positive");
    if ((double)X * Y * Z >= 2147483647.0) throw host::Error(CLAT_EINVAL, "dimensions are too large");
    if (precision != CLAT_PREC_FP64 && precision != CLAT_PREC_FP32)
      throw host::Error(CLAT_EINVAL, "unknown precision");
    if (schedule != CLAT_SCHED_SHARED && schedule != CLAT_SCHED_ROUNDS)
      throw host::Error(CLAT_EINVAL, "unknown schedule");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
      throw host::Error(CLAT_EDEVICE, "no HIP device available");
    HIPX(hipGetDevice(&e->device));
    e->X = X; e->Y = Y; e->Z = Z;
    e->Xi = host::round2357(X); e->Yi = host::round2357(Y); e->Zi = host::round2357(Z);
    if ((double)e->Xi * e->Yi * e->Zi * (schedule == CLAT_SCHED_SHARED ? 4.0 : 1.0) >= 4294967295.0)
      throw host::Error(CLAT_EUNSUP, "grid too large for one engine");
    e->prec = precision;
    e->sched = schedule;
    e->csz = precision == CLAT_PREC_FP32 ? 8 : 16;
    e->tile_budget = precision == CLAT_PREC_FP32 ? 32 * NT : 16 * NT;
    if (const char* tb = std::getenv("CLAT_TILE_ELEMS")) {
      const int v = std::atoi(tb);
      if (v >= 64 && v <= e->tile_budget) e->tile_budget = v;
    }
    const int dims[3] = {e->Xi, e->Yi, e->Zi};
    AxisPlan* lp[3] = {&e->px, &e->py, &e->pz};
    AxisPlan* dp[3] = {&e->dx, &e->dy, &e->dz};
    for (int d = 0; d < 3; ++d) {
      if (!make_plan(dims[d], *lp[d]) || !make_plan(2 * dims[d], *dp[d]))
        throw host::Error(CLAT_EUNSUP, "axis length has no radix plan");
      const int nsub = schedule == CLAT_SCHED_SHARED ? 2 : 1;
      if (nsub * dims[d] > e->tile_budget || 2 * dims[d] > 16 * NT)
        throw host::Error(CLAT_EUNSUP, "axis length exceeds the single-tile FFT limit");
    }
    HIPX(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    for (int i = 0; i < 6; ++i) HIPX(hipEventCreate(&e->stage_ev[i]));
    // twiddle and half-shift tables
    for (int d = 0; d < 3; ++d) {
      const int L = dims[d];
      std::vector<double> tw = unit_roots(L, L), hf = unit_roots(L, 2 * L), dt = unit_roots(2 * L, 2 * L);
      if (precision == CLAT_PREC_FP32) {
        std::vector<float> a(tw.begin(), tw.end()), b(hf.begin(), hf.end());
        e->tw[d] = e->dalloc(a.size() * sizeof(float));
        e->half[d] = e->dalloc(b.size() * sizeof(float));
        HIPX(hipMemcpy(e->tw[d], a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPX(hipMemcpy(e->half[d], b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
      } else {
        e->tw[d] = e->dalloc(tw.size() * sizeof(double));
        e->half[d] = e->dalloc(hf.size() * sizeof(double));
        HIPX(hipMemcpy(e->tw[d], tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPX(hipMemcpy(e->half[d], hf.data(), hf.size() * sizeof(double), hipMemcpyHostToDevice));
      }
      e->dtw[d] = (double2*)e->dalloc(dt.size() * sizeof(double));
      HIPX(hipMemcpy(e->dtw[d], dt.data(), dt.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    e->G64 = (double*)e->dalloc(e->NG() * sizeof(double));
    if (precision == CLAT_PREC_FP32) e->G32 = (float*)e->dalloc(e->NG() * sizeof(float));
    e->F = e->dalloc(e->N() * e->csz);
    e->V = e->dalloc(e->N() * e->csz);
    e->A = e->dalloc(e->a_elems() * e->csz);
    if (e->b_elems()) e->B = e->dalloc(e->b_elems() * e->csz);
    zero_F(e);
    HIPX(hipStreamSynchronize(e->stream));
  });
  if (rc != CLAT_OK) {
    clat_destroy(e);
    return rc;
  }
  *out = e;
  return CLAT_OK;
}

int clat_destroy(clat_engine* e) {
  if (!e) return CLAT_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  void* ptrs[] = {e->tw[0], e->tw[1], e->tw[2], e->half[0], e->half[1], e->half[2], e->dtw[0], e->dtw[1],
                  e->dtw[2], e->G64, e->G32, e->F, e->V, e->A, e->B, e->Fin, e->dpart, e->d_cell,
                  e->d_site_atom, e->d_site_cell, e->coeffs, e->rho, e->Vat, e->part, e->ovl};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto ev : e->stage_ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto ev : e->kev) (void)hipEventDestroy(ev);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return CLAT_OK;
}

int clat_set_interaction_simple(clat_engine* e, const double step[3], double onsite, double dielectric) {
  if (!e || !step) return set_err(CLAT_EINVAL, "invalid arguments");
  return guarded([&] {
    GArgs g = base_gargs(e);
    g.sx = step[0]; g.sy = step[1]; g.sz = step[2];
    g.onsite = onsite;
    g.dielectric = dielectric;
    g.tf = 0;
    build_table(e, g);
  });
}

int clat_set_interaction_thomas_fermi(clat_engine* e, const double step[3], double onsite, double dielectric,
                                      double lattice_constant) {
  if (!e || !step) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!(lattice_constant > 0.0)) return set_err(CLAT_EINVAL, "--tf-lattice parameter must be positive");
  if (!(dielectric > 1.0)) return set_err(CLAT_EINVAL, "dielectric constant must be >1 to use Thomas-Fermi model");
  return guarded([&] {
    // Interaction.cpp:20-31: q_TF and the Newton solve of sinh(x) = eps x
    const double eps = 1.0e-12;
    const double qTF = 2.0 / std::sqrt(M_PI) * std::cbrt(96.0 * M_PI * M_PI) / lattice_constant;
    double x = std::sqrt(6.0 * (dielectric - 1.0)), dx;
    do {
      const double Fv = std::sinh(x) - dielectric * x;
      const double dF = std::cosh(x) - dielectric;
      dx = (std::fabs(Fv) < eps) ? 0 : Fv / dF;
      x -= dx;
    } while (dx / qTF > eps);
    GArgs g = base_gargs(e);
    g.sx = step[0]; g.sy = step[1]; g.sz = step[2];
    g.onsite = onsite;
    g.dielectric = dielectric;
    g.tf = 1;
    g.rTF = x / qTF;
    g.qTF = qTF;
    build_table(e, g);
  });
}

int clat_set_interaction_table(clat_engine* e, const double* G) {
  if (!e || !G) return set_err(CLAT_EINVAL, "invalid arguments");
  return guarded([&] {
    const size_t nu = (size_t)(e->X + 1) * (e->Y + 1) * (e->Z + 1);
    double* tmp = (double*)e->dalloc(nu * sizeof(double));
    try {
      HIPX(hipMemcpyAsync(tmp, G, nu * sizeof(double), hipMemcpyHostToDevice, e->stream));
      GArgs g = base_gargs(e);
      hipLaunchKernelGGL(k_gpad, dim3(grid1d(e->NG())), dim3(256), 0, e->stream, tmp, e->G64, g);
      LAUNCH_CHECK();
      finish_interaction(e);
      HIPX(hipStreamSynchronize(e->stream));
    } catch (...) {
      void* t = tmp;
      e->dfree(t, nu * sizeof(double));
      throw;
    }
    void* t = tmp;
    e->dfree(t, nu * sizeof(double));
  });
}

int clat_set_input(clat_engine* e, const double* F) {
  if (!e || !F) return set_err(CLAT_EINVAL, "invalid arguments");
  return guarded([&] {
    if (!e->Fin) e->Fin = (double2*)e->dalloc(e->NU() * sizeof(double2));
    HIPX(hipMemcpyAsync(e->Fin, F, e->NU() * sizeof(double2), hipMemcpyHostToDevice, e->stream));
    zero_F(e);
    if (e->prec == CLAT_PREC_FP32)
      hipLaunchKernelGGL(k_grid_in<float>, dim3(grid1d(e->NU())), dim3(256), 0, e->stream, e->Fin,
                         (float2*)e->F, e->X, e->Y, e->Z, e->Xi, e->Yi);
    else
      hipLaunchKernelGGL(k_grid_in<double>, dim3(grid1d(e->NU())), dim3(256), 0, e->stream, e->Fin,
                         (double2*)e->F, e->X, e->Y, e->Z, e->Xi, e->Yi);
    LAUNCH_CHECK();
    HIPX(hipStreamSynchronize(e->stream));
    e->input_set = true;
  });
}

int clat_prepare(clat_engine* e) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!e->initialized) return set_err(CLAT_ESTATE, "calculator is not initialized");
  return guarded([&] {
    prepare(e);
    e->kcollect();
    HIPX(hipStreamSynchronize(e->stream));
    e->prepared = true;
  });
}

int clat_calculate(clat_engine* e, double result[2]) {
  if (!e || !result) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!e->prepared) return set_err(CLAT_ESTATE, "prepare() has not been called");
  if (!e->input_set) return set_err(CLAT_ESTATE, "input has not been set");
  return guarded([&] {
    const unsigned nb = std::min<unsigned>(grid1d(e->NU()), 2048);
    ensure_dpart(e, 2 * (size_t)nb);
    if (e->prec == CLAT_PREC_FP32)
      hipLaunchKernelGGL(k_grid_dot<float>, dim3(nb), dim3(NT), 0, e->stream, (const float2*)e->V, e->Fin, e->X,
                         e->Y, e->Z, e->Xi, e->Yi, e->dpart);
    else
      hipLaunchKernelGGL(k_grid_dot<double>, dim3(nb), dim3(NT), 0, e->stream, (const double2*)e->V, e->Fin,
                         e->X, e->Y, e->Z, e->Xi, e->Yi, e->dpart);
    LAUNCH_CHECK();
    std::vector<double> h(2 * (size_t)nb);
    HIPX(hipMemcpyAsync(h.data(), e->dpart, h.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPX(hipStreamSynchronize(e->stream));
    result[0] = nsum(h.data(), nb, 2);
    result[1] = nsum(h.data() + 1, nb, 2);
  });
}

int clat_get_potential(clat_engine* e, double* V) {
  if (!e || !V) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!e->prepared) return set_err(CLAT_ESTATE, "prepare() has not been called");
  return guarded([&] {
    double2* tmp = (double2*)e->dalloc(e->NU() * sizeof(double2));
    if (e->prec == CLAT_PREC_FP32)
      hipLaunchKernelGGL(k_grid_out<float>, dim3(grid1d(e->NU())), dim3(256), 0, e->stream, (const float2*)e->V,
                         tmp, e->X, e->Y, e->Z, e->Xi, e->Yi);
    else
      hipLaunchKernelGGL(k_grid_out<double>, dim3(grid1d(e->NU())), dim3(256), 0, e->stream,
                         (const double2*)e->V, tmp, e->X, e->Y, e->Z, e->Xi, e->Yi);
    hipError_t le = hipGetLastError();
    hipError_t ce = hipMemcpyAsync(V, tmp, e->NU() * sizeof(double2), hipMemcpyDeviceToHost, e->stream);
    hipError_t se = hipStreamSynchronize(e->stream);
    void* t = tmp;
    e->dfree(t, e->NU() * sizeof(double2));
    HIPX(le);
    HIPX(ce);
    HIPX(se);
  });
}

int clat_set_atoms(clat_engine* e, const int32_t* cell, int64_t n_atoms, int n_orbitals) {
  if (!e || (!cell && n_atoms > 0) || n_atoms < 0) return set_err(CLAT_EINVAL, "invalid arguments");
  if (n_orbitals <= 0) return set_err(CLAT_EINVAL, "invalid value for orbitals");
  return guarded([&] {
    const int64_t NU = (int64_t)e->NU();
    // effective sites: the last atom of every occupied cell, ordered by cell
    std::vector<int32_t> last((size_t)NU, -1);
    for (int64_t a = 0; a < n_atoms; ++a) {
      if (cell[a] < 0 || cell[a] >= NU) throw host::Error(CLAT_EINVAL, "atom cell index outside the grid");
      last[(size_t)cell[a]] = (int32_t)a;
    }
    std::vector<int32_t> site_atom, site_cell, all_cell((size_t)n_atoms);
    for (int64_t c = 0; c < NU; ++c)
      if (last[(size_t)c] >= 0) {
        site_atom.push_back(last[(size_t)c]);
        const int ix = (int)(c % e->X), iy = (int)((c / e->X) % e->Y), iz = (int)(c / ((int64_t)e->X * e->Y));
        site_cell.push_back((int32_t)(((int64_t)iz * e->Yi + iy) * e->Xi + ix));
      }
    for (int64_t a = 0; a < n_atoms; ++a) {
      const int64_t c = cell[a];
      const int ix = (int)(c % e->X), iy = (int)((c / e->X) % e->Y), iz = (int)(c / ((int64_t)e->X * e->Y));
      all_cell[(size_t)a] = (int32_t)(((int64_t)iz * e->Yi + iy) * e->Xi + ix);
    }
    // drop previous atoms, states and derived buffers
    void* p;
    p = e->d_cell; e->dfree(p, (size_t)e->n_atoms * 4); e->d_cell = nullptr;
    p = e->d_site_atom; e->dfree(p, (size_t)e->n_sites * 4); e->d_site_atom = nullptr;
    p = e->d_site_cell; e->dfree(p, (size_t)e->n_sites * 4); e->d_site_cell = nullptr;
    p = e->coeffs; e->dfree(p, (size_t)e->cap_states * e->n_atoms * e->n_orb * 16); e->coeffs = nullptr;
    p = e->rho;
    e->dfree(p, (size_t)(e->rho_states > 0 ? e->rho_states * (e->rho_states + 1) / 2 : 0) * e->n_sites * 16);
    e->rho = nullptr;
    e->rho_states = -1;
    e->n_states = e->cap_states = 0;
    e->n_atoms = n_atoms;
    e->n_orb = n_orbitals;
    e->n_sites = (int64_t)site_atom.size();
    e->d_cell = (int32_t*)e->dalloc((size_t)n_atoms * 4);
    e->d_site_atom = (int32_t*)e->dalloc((size_t)e->n_sites * 4);
    e->d_site_cell = (int32_t*)e->dalloc((size_t)e->n_sites * 4);
    HIPX(hipMemcpy(e->d_cell, all_cell.data(), (size_t)n_atoms * 4, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(e->d_site_atom, site_atom.data(), (size_t)e->n_sites * 4, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(e->d_site_cell, site_cell.data(), (size_t)e->n_sites * 4, hipMemcpyHostToDevice));
  });
}

int clat_add_state(clat_engine* e, const double* coeffs) {
  if (!e || !coeffs) return set_err(CLAT_EINVAL, "invalid arguments");
  if (e->n_orb <= 0) return set_err(CLAT_ESTATE, "atoms have not been set");
  return guarded([&] {
    const size_t per = (size_t)e->n_atoms * e->n_orb * sizeof(double2);
    if (e->n_states == e->cap_states) {
      const int ncap = std::max(4, 2 * e->cap_states);
      double2* nc = (double2*)e->dalloc(per * ncap);
      if (e->n_states) HIPX(hipMemcpy(nc, e->coeffs, per * e->n_states, hipMemcpyDeviceToDevice));
      void* p = e->coeffs;
      e->dfree(p, per * e->cap_states);
      e->coeffs = nc;
      e->cap_states = ncap;
    }
    HIPX(hipMemcpy((char*)e->coeffs + per * e->n_states, coeffs, per, hipMemcpyHostToDevice));
    e->n_states += 1;
  });
}

int clat_clear_states(clat_engine* e) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  e->n_states = 0;
  return CLAT_OK;
}

int clat_compute_integrals(clat_engine* e, const int32_t* quads, int64_t n, int flags, double* out) {
  if (!e || (n > 0 && (!quads || !out)) || n < 0) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!e->initialized) return set_err(CLAT_ESTATE, "calculator is not initialized");
  if (e->n_states <= 0) return set_err(CLAT_ESTATE, "no states loaded");
  return guarded([&] {
    std::vector<host::PlanStep> steps = host::make_plan(quads, n, e->n_states);
    const int P = e->n_states * (e->n_states + 1) / 2;
    // left products in execution order -> slots
    std::vector<int> slot_of((size_t)P, -1), slots;
    for (const auto& s : steps)
      if (slot_of[(size_t)s.left.product] < 0) {
        slot_of[(size_t)s.left.product] = (int)slots.size();
        slots.push_back(s.left.product);
      }
    const int nL = (int)slots.size();
    HIPX(hipEventRecord(e->stage_ev[1], e->stream));
    build_products(e);
    HIPX(hipEventRecord(e->stage_ev[2], e->stream));
    const size_t vat_need = (size_t)std::max(1, nL) * e->n_sites;
    if (e->vat_cap < vat_need) {
      void* p = e->Vat;
      e->dfree(p, e->vat_cap * sizeof(double2));
      e->Vat = (double2*)e->dalloc(vat_need * sizeof(double2));
      e->vat_cap = vat_need;
    }
    for (int sl = 0; sl < nL; ++sl) {
      scatter_product(e, slots[(size_t)sl]);
      prepare(e);
      gather_V(e, e->d_site_cell, e->n_sites, e->Vat + (size_t)sl * e->n_sites);
    }
    HIPX(hipEventRecord(e->stage_ev[3], e->stream));
    // overlaps of every (slot, product) pair
    const int tl = (nL + OV_T - 1) / OV_T, tr = (P + OV_T - 1) / OV_T;
    int64_t n_chunks = std::max<int64_t>(1, 2048 / std::max(1, tl * tr));
    n_chunks = std::min<int64_t>(n_chunks, (e->n_sites + OV_S - 1) / OV_S);
    n_chunks = std::max<int64_t>(1, n_chunks);
    int64_t chunk = (e->n_sites + n_chunks - 1) / n_chunks;
    chunk = (chunk + OV_S - 1) / OV_S * OV_S;
    n_chunks = std::max<int64_t>(1, (e->n_sites + chunk - 1) / chunk);
    const size_t n_out = (size_t)std::max(1, nL) * P * 4;
    if (e->part_cap < n_out * n_chunks) {
      void* p = e->part;
      e->dfree(p, e->part_cap * sizeof(double));
      e->part = (double*)e->dalloc(n_out * n_chunks * sizeof(double));
      e->part_cap = n_out * n_chunks;
    }
    if (e->ovl_cap < n_out) {
      void* p = e->ovl;
      e->dfree(p, e->ovl_cap * sizeof(double));
      e->ovl = (double*)e->dalloc(n_out * sizeof(double));
      e->ovl_cap = n_out;
    }
    std::vector<double> ov(n_out, 0.0);
    if (nL > 0 && e->n_sites > 0) {
      e->kbegin();
      hipLaunchKernelGGL(k_overlap_partial, dim3(tl, tr, (unsigned)n_chunks), dim3(NT), 0, e->stream, e->Vat,
                         e->rho, nL, P, e->n_sites, chunk, e->part);
      LAUNCH_CHECK();
      hipLaunchKernelGGL(k_overlap_reduce, dim3(grid1d(n_out)), dim3(256), 0, e->stream, e->part, n_chunks,
                         (int64_t)n_out, e->ovl);
      LAUNCH_CHECK();
      e->kend(K_OVERLAP, (double)(tl * tr) * (double)(2 * OV_T) * (double)e->n_sites * 16.0);
      HIPX(hipMemcpyAsync(ov.data(), e->ovl, n_out * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    }
    HIPX(hipEventRecord(e->stage_ev[4], e->stream));
    HIPX(hipStreamSynchronize(e->stream));
    e->kcollect();
    // main loop bookkeeping (coulombo.cpp:186-209)
    const bool stale = (flags & CLAT_FLAG_STALE_VALUES) != 0;
    int lastR = -1, lastRc = 0;
    double vre = 0.0, vim = 0.0;
    for (const auto& s : steps) {
      const int rc = (s.left.conj != s.right.conj) ? 1 : 0;
      if (!stale || s.right.product != lastR || rc != lastRc) {
        const double* q = ov.data() + (((size_t)slot_of[(size_t)s.left.product]) * P + s.right.product) * 4;
        const double Av = q[0], Bv = q[1], Cv = q[2], Dv = q[3];
        if (rc) { vre = Av + Bv; vim = Dv - Cv; }
        else { vre = Av - Bv; vim = Cv + Dv; }
        lastR = s.right.product;
        lastRc = rc;
      }
      out[2 * (size_t)s.id] = vre;
      out[2 * (size_t)s.id + 1] = s.left.conj ? -vim : vim;
    }
    float ms;
    HIPX(hipEventElapsedTime(&ms, e->stage_ev[1], e->stage_ev[2])); e->stage_ms[1] = ms;
    HIPX(hipEventElapsedTime(&ms, e->stage_ev[2], e->stage_ev[3])); e->stage_ms[2] = ms;
    HIPX(hipEventElapsedTime(&ms, e->stage_ev[3], e->stage_ev[4])); e->stage_ms[3] = ms;
    HIPX(hipEventElapsedTime(&ms, e->stage_ev[1], e->stage_ev[4])); e->stage_ms[4] = ms;
  });
}

int clat_compute_potential_at_atoms(clat_engine* e, int state, double* v) {
  if (!e || !v) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!e->initialized) return set_err(CLAT_ESTATE, "calculator is not initialized");
  if (state < 1 || state > e->n_states) return set_err(CLAT_EINVAL, "state index out of range");
  return guarded([&] {
    build_products(e);
    const int f = state - 1;
    scatter_product(e, f * (f + 1) / 2 + f);
    prepare(e);
    double2* tmp = (double2*)e->dalloc((size_t)std::max<int64_t>(1, e->n_atoms) * sizeof(double2));
    gather_V(e, e->d_cell, e->n_atoms, tmp);
    hipError_t ce = hipMemcpyAsync(v, tmp, (size_t)e->n_atoms * sizeof(double2), hipMemcpyDeviceToHost, e->stream);
    hipError_t se = hipStreamSynchronize(e->stream);
    void* t = tmp;
    e->dfree(t, (size_t)std::max<int64_t>(1, e->n_atoms) * sizeof(double2));
    HIPX(ce);
    HIPX(se);
    e->kcollect();
  });
}

int clat_last_stage_ms(clat_engine* e, double ms[5]) {
  if (!e || !ms) return set_err(CLAT_EINVAL, "invalid arguments");
  for (int i = 0; i < 5; ++i) ms[i] = e->stage_ms[i];
  return CLAT_OK;
}

int clat_set_kernel_timing(clat_engine* e, int enable) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  e->ktiming = enable != 0;
  return CLAT_OK;
}

int clat_kernel_kind_count(void) { return K_COUNT; }

const char* clat_kernel_kind_name(int kind) { return (kind >= 0 && kind < K_COUNT) ? kKindNames[kind] : ""; }

int clat_kernel_stats(clat_engine* e, int kind, int64_t* launches, double* total_ms, double* bytes) {
  if (!e || kind < 0 || kind >= K_COUNT) return set_err(CLAT_EINVAL, "invalid arguments");
  if (launches) *launches = e->kcount[kind];
  if (total_ms) *total_ms = e->kms[kind];
  if (bytes) *bytes = e->kbytes[kind];
  return CLAT_OK;
}

int clat_reset_kernel_stats(clat_engine* e) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  for (int k = 0; k < K_COUNT; ++k) {
    e->kcount[k] = 0;
    e->kms[k] = 0;
    e->kbytes[k] = 0;
  }
  return CLAT_OK;
}

int clat_engine_info(clat_engine* e, int dim[3], int64_t* device_bytes) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  if (dim) { dim[0] = e->Xi; dim[1] = e->Yi; dim[2] = e->Zi; }
  if (device_bytes) *device_bytes = e->device_bytes;
  return CLAT_OK;
}

}  // extern "C"
