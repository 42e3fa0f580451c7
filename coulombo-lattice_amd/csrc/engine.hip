// The C-ABI engine: device buffers, FFT plans, launch sequences of
// prepare()/calculate() and the batched integral path.  Built for gfx950.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/clat.h"
#include "host_logic.hpp"
#include "kernels.hpp"
#include "linepass.hpp"

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

// a failed call also clears the runtime's last-error state, so the failure
// is reported once and later launch checks start clean
#define HIPX(call)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      (void)hipGetLastError();                                                           \
      throw HipError{e_ == hipErrorOutOfMemory ? CLAT_ENOMEM : CLAT_EDEVICE,             \
                     std::string(#call) + ": " + hipGetErrorString(e_)};                 \
    }                                                                                    \
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

// Stockham plan (generic pass kernels and the DCT-I)
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

// two-level split L = R1 * R2 (R1 >= R2, most balanced), false if none
bool split2(int L, int& R1, int& R2) {
  int best = -1;
  for (int r2 = 1; r2 * r2 <= L; ++r2) {
    if (L % r2) continue;
    const int r1 = L / r2;
    if (reg_size_ok(r1) && reg_size_ok(r2)) best = r2;
  }
  if (best < 0) return false;
  R2 = best;
  R1 = L / best;
  return true;
}

bool col_pair_compiled(int R1, int R2);

// engine axis length for a user length: the smallest size whose two-level
// split has compiled kernels (every 2*3*5-smooth length from 32 to 400 that
// splits has them; the open-boundary convolution does not depend on the
// padding), else the smallest two-level size, else the next 2*3*5*7-smooth
// size (generic kernels)
int engine_len(int n) {
  if (n >= 32)
    for (int m = n; m <= 400; ++m) {
      int a, b;
      if (split2(m, a, b) && col_pair_compiled(a, b)) return m;
    }
  for (int m = n; m <= 400; ++m) {
    int a, b;
    if (split2(m, a, b)) return m;
  }
  return host::round2357(n);
}

// e^{-2 pi i m / den} for m in [0, n), from long double trigonometry
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
int ceil_pow2(int v) {
  int p = 1;
  while (p < v) p *= 2;
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

// launch shape of one two-level pass
struct PassCfg {
  int R1 = 0, R2 = 0, R1p = 0, C = 1, lgC = 0, NT = 64;
  size_t lds = 0;
};

int round64(int v) { return (v + 63) / 64 * 64; }

PassCfg cfg_line(int L, int nsub, size_t c2bytes, bool sites, size_t budget) {
  PassCfg p;
  split2(L, p.R1, p.R2);
  p.R1p = (p.R1 > 1 && p.R1 % 2 == 0) ? p.R1 + 1 : p.R1;
  const size_t per = (size_t)std::max(nsub * p.R2 * p.R1p, L) * c2bytes;
  int C = floor_pow2((int)std::max<size_t>(1, budget / per));
  C = std::min(C, 64);
  if (sites) C = std::min(C, std::max(1, floor_pow2(256 / p.R1)));
  p.C = C;
  p.lgC = log2i(C);
  p.NT = std::min(256, std::max(64, round64(std::max(C * p.R1, C * nsub * p.R2))));
  p.lds = (size_t)C * per;
  return p;
}

PassCfg cfg_col(int L, int nsub, size_t c2bytes, int W, size_t budget) {
  PassCfg p;
  split2(L, p.R1, p.R2);
  p.R1p = p.R1;
  const size_t per = (size_t)nsub * L * c2bytes;
  int C = floor_pow2((int)std::max<size_t>(1, budget / per));
  C = std::min(C, c2bytes == 8 ? 32 : 16);
  C = std::min(C, ceil_pow2(W));
  p.C = C;
  p.lgC = log2i(C);
  p.NT = std::min(256, std::max(64, round64(std::max(C * p.R1, C * nsub * p.R2))));
  p.lds = (size_t)C * per;
  return p;
}

}  // namespace

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
struct clat_engine {
  int X = 0, Y = 0, Z = 0;     // user grid
  int Xi = 0, Yi = 0, Zi = 0;  // engine grid
  int prec = 0, sched = 0;     // requested precision and schedule
  int rprec = 0;               // precision the passes run in
  int mix = 0;                 // mixed mode: passes computing in FP64 (0: last only .. 3: all)
  bool fast = false;           // two-level pass kernels on all three axes
  int device = 0;
  hipStream_t stream = nullptr;
  AxisPlan px{}, py{}, pz{};   // Stockham plans of the generic kernels
  AxisPlan dx{}, dy{}, dz{};   // DCT-I through 2M-point transforms
  size_t lds_budget = 65536;
  int tile_budget = 0;         // generic kernels: elements of one LDS tile
  PassCfg cxf, cyf, czc, cyi, cxi, cxf_s, cxi_s;  // two-level launch shapes (_s: atom sites)
  // element sizes (bytes per complex) of the buffers
  size_t s_in = 16, s_a = 16, s_b = 16, s_d = 16, s_v = 16;
  // device tables
  void *tw32[3] = {}, *h32[3] = {}, *tw64[3] = {}, *h64[3] = {};
  double2* dtw[3] = {};         // W_2L^m, m < 2L, per axis (FP64)
  float2* dtw32[3] = {};        // same, FP32 (column passes in complex64)
  bool colk = true;             // compile-time column kernels for the y and z passes
  bool conv_tiled = false;      // conv pass on the column kernel (tiled kernel table)
  bool btile = false;           // B in strips of C columns x Z rows (shared schedule)
  bool atile = false;           // A and D in strips of C columns x Y rows (atom path, with btile)
  int gc_c = 16, gc_ncb = 1, gc_npar = 1, gc_w = 0;  // kernel-table tiling
  double* G64 = nullptr;  // Ghat octant (Z+1)(Y+1)(X+1), FP64
  float* G32 = nullptr;   // same, FP32 (generic complex64 kernels)
  double* Gc64 = nullptr; // Ghat as [r(fy)][r(fz)][fx], fx in [0, 2X): conv pass tables
  float* Gc32 = nullptr;
  size_t gc_elems = 0;
  bool initialized = false;
  // grids
  void *F = nullptr, *V = nullptr;  // dense input / potential (grid protocol, generic path)
  void *A = nullptr, *B = nullptr, *D = nullptr;
  size_t a_bytes = 0, b_bytes = 0, d_bytes = 0, f_bytes = 0, v_bytes = 0;
  double2* Fin = nullptr;  // user-layout FP64 copy of the input (calculate())
  bool input_set = false, prepared = false;
  double* dpart = nullptr;
  size_t dpart_cap = 0;
  // atoms and states
  int64_t n_atoms = 0;
  int n_orb = 0;
  int64_t n_sites = 0;               // effective atoms (last atom of each cell)
  int32_t* d_site_atom = nullptr;    // atom of each site (sites ordered by cell)
  int32_t* d_site_cell = nullptr;    // engine-grid cell of each site
  int32_t* d_site_ptr = nullptr;     // CSR of sites over engine x-lines (Yi*Zi + 1)
  int32_t* d_atom_site = nullptr;    // site holding each atom's cell
  // occupied x-lines (atom path of the shared schedule)
  int64_t n_live = 0;
  int32_t* d_lines = nullptr;        // line id of each occupied line, ascending
  int32_t* d_live_ptr = nullptr;     // CSR of the sites over occupied lines (n_live + 1)
  int32_t* d_site_ccell = nullptr;   // occupied-line index * Xi + x of each site
  uint32_t* d_linemask = nullptr;    // y passes: word (z, n1), bit n2 = line y = n1 + R1y n2 holds a site
  size_t linemask_words = 0;
  // batched prepares of the atom path (line kernels): up to nb_max per launch
  int nb_max = 1, cur_nb = 1;
  const double2* b_in[LINE_NB] = {};
  double2* b_out[LINE_NB] = {};
  const double2* b_in2[LINE_NB] = {};  // second real density packed into the prepare (or null)
  double2* b_out2[LINE_NB] = {};
  bool pack = true;                     // pairs of real (diagonal) densities share a prepare
  // prepares per launch: measured on the first run of a shape (autotuner) or
  // the fixed rule
  int nb_rule = 1;
  int tune = 1;                         // 1: measure (the default), 0: rule only
  int tuned_nb = 0;                     // 0: not measured for the current atoms
  double tune_ms[8] = {};               // prepare-stage ms of k + 1 prepares per launch (0: not tried)
  hipEvent_t tune_ev[9] = {};
  int64_t atoms_gen = 0;                // bumped by every clat_set_atoms
  // the device stage of a run (prepares, overlaps) replayed as one HIP graph
  bool use_graph = true;
  hipGraphExec_t gexec[2] = {};         // prepares, overlaps
  hipGraphExec_t ggroup = nullptr;      // device stages of a group this engine leads
  std::vector<int64_t> ggroup_key;
  hipEvent_t join_ev = nullptr;         // stream joins of group runs
  std::vector<int64_t> gkey;            // what the graphs' kernel arguments depend on
  int64_t graph_replays = 0, graph_captures = 0;
  int64_t serial = 0;                   // unique per engine of the process (group graph keys)
  int n_cu = 256;     // compute units of the device
  // Two-coset path of the atom path (shared schedule, line kernels): when
  // every site lies on x = y = z (mod 2) of an even grid, the sites split
  // into the cosets 2m and 2m + (1,1,1) of the coarse grid with half the
  // points per axis.  The coarse engine `sub` (on this engine's stream) runs
  // the five passes on the two coset planes of each prepare, its conv
  // coupling them through the 2 x 2 kernel (k_colconv2); sites are ordered
  // coset 0, then coset 1, each over the n_u coarse cells either occupies.
  clat_engine* sub = nullptr;
  bool sub_failed = false;      // the coarse engine cannot run this grid
  bool coset = false;           // the current atoms run on the two-coset path
  int64_t n_u = 0;
  bool fine_pending = false;    // G64 holds the real-space table: its DCT-I and conv table wait for first use
  bool g64_real = false;        // with fine_pending: G64 filled (else the table is gdefer's, not yet sampled)
  // kept by the coarse engine of a two-coset path
  clat_engine* kparent = nullptr;  // kernel timing goes to this engine's events
  bool own_stream = true;
  bool pair_conv = false;          // the conv couples planes 2b and 2b + 1
  void* Q = nullptr;               // cross-coset kernel table (k_coarse_q / k_coarse_q_half)
  size_t q_cap = 0;                // its capacity in elements
  double* Gt64 = nullptr;          // octant of the cross-coset kernel: samples, then DCT-II
  double2* ph64 = nullptr;         // e^{i pi k / (2 L)}, k < 2 L, per axis (the cross-coset phases)
  // half-box form (every coarse site on mx + my + mz even): x frequencies
  // f < fmax carried in wcols columns (the strip width's multiple), kernels
  // summed over the half-box shift; otherwise wcols = fmax = 2 Xi
  bool half = false;
  int wcols = 0, fmax = 0;
  // half-length x lines (coarse engine of a half box): every site has
  // x = (y + z) mod 2, so a line is carried as its Xi / 2 sites (radix pair
  // hx of Xi / 2) and the line kernels add the x parity's phase.  The x
  // inverse (FP64 in FP32X) always takes them; the FP32X forward keeps full
  // lines, whose complex64 rounding pattern keeps the 1M-atom workload's
  // smallest integrals within 1e-5 of FP64 (9.1e-6 against 1.03e-5 on half
  // lines: the noise of FP32 arithmetic at 1e-8 of the largest integral)
  bool xhalf = false;
  int hx_r1 = 0, hx_r2 = 0;
  size_t q_elems = 0;
  // host buffers of a run, kept with their pages between runs
  std::vector<host::PlanStep> plan_steps;
  std::vector<char> plan_in_cover, plan_leads;
  double* ov_host = nullptr;            // overlap table on the host, pinned (a faster copy back)
  size_t ov_host_cap = 0;
  bool ovl_sym = true;  // symmetric MFMA overlap table (CLAT_OVL_SYM=0: every tile)
  int ovl_mode = 0;  // overlap kernel: 0 by size, 1 vector FMA, 2 FP64 MFMA, 3 vector FMA in 48 x 48 tiles,
                    // 4 the same with every site sum Kahan-compensated
  bool table_pending = false;           // stage_ms[0] of the last table build not read yet
  bool table_deferred = false;          // a set interaction whose table is not enqueued yet
  GArgs gdefer{};                       // its arguments
  size_t a_set = 0, b_set = 0, d_set = 0;  // elements of A, B, D per prepare
  bool xline = false;                // x passes on the line kernel, y passes skip empty lines
  int n_states = 0, cap_states = 0;
  double2* coeffs = nullptr;
  double2* rho = nullptr;
  size_t rho_bytes = 0;
  int rho_states = -1;
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
  uint32_t kmask = 0;  // kinds whose launches are timed (bit = kernel kind)
  std::vector<hipEvent_t> kev;
  std::vector<std::pair<int, double>> kpending;
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
  template <typename P>
  void dfree(P*& p, size_t bytes) {
    if (p) {
      (void)hipFree((void*)p);
      device_bytes -= (int64_t)bytes;
      p = nullptr;
    }
  }
  size_t N() const { return (size_t)Xi * Yi * Zi; }
  size_t NG() const { return (size_t)(Xi + 1) * (Yi + 1) * (Zi + 1); }
  size_t NU() const { return (size_t)X * Y * Z; }
  bool shared() const { return sched == CLAT_SCHED_SHARED; }

  // kernel timing: a coarse engine records into its parent's events
  clat_engine* kown() { return kparent ? kparent : this; }
  bool ktimed(int kind) const { return ((kparent ? kparent->kmask : kmask) >> kind) & 1u; }
  void kbegin(int kind) {
    if (!ktimed(kind)) return;
    clat_engine* o = kown();
    o->ensure_events(o->kused + 2);
    HIPX(hipEventRecord(o->kev[o->kused], stream));
  }
  void kend(int kind, double bytes) {
    if (!ktimed(kind)) return;
    clat_engine* o = kown();
    HIPX(hipEventRecord(o->kev[o->kused + 1], stream));
    o->kpending.push_back({kind, bytes});
    o->kused += 2;
  }
  void ensure_events(size_t n) {
    while (kev.size() < n) {
      hipEvent_t ev;
      HIPX(hipEventCreate(&ev));
      kev.push_back(ev);
    }
  }
  void kcollect() {
    if (kpending.empty()) return;
    HIPX(hipStreamSynchronize(stream));
    // (the table-build events of settle_table are complete from here on)
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

// ===========================================================================
// two-level pass launches
// ===========================================================================
template <typename TC>
const C2<TC>* tw_of(const clat_engine* e, int axis) {
  return (const C2<TC>*)(sizeof(TC) == 8 ? e->tw64[axis] : e->tw32[axis]);
}
template <typename TC>
const C2<TC>* h_of(const clat_engine* e, int axis) {
  return (const C2<TC>*)(sizeof(TC) == 8 ? e->h64[axis] : e->h32[axis]);
}
// octant table (generic kernels)
template <typename TC>
const TC* g_of(const clat_engine* e) {
  return (const TC*)(sizeof(TC) == 8 ? (const void*)e->G64 : (const void*)e->G32);
}
// conv-layout table [r(fy)][r(fz)][fx] of the two-level kernels
template <typename TC>
const TC* gc_of(const clat_engine* e) {
  return (const TC*)(sizeof(TC) == 8 ? (const void*)e->Gc64 : (const void*)e->Gc32);
}

template <typename TI, typename TC, typename TO>
constexpr int combo_of() {
  constexpr bool di = sizeof(TI) == 8, dc = sizeof(TC) == 8, dout = sizeof(TO) == 8;
  return (di && dc && dout) ? CB_DDD : (!di && !dc && !dout) ? CB_FFF : (di && !dc && !dout) ? CB_DFF
       : (!di && !dc && dout) ? CB_FFD : (di && dc && !dout) ? CB_DDF : (!di && dc && !dout) ? CB_FDF
       : CB_FDD;
}

bool col_pair_compiled(int R1, int R2) {
  switch (R1 * 64 + R2) {
#define CLAT_C2_HAVE(ra, rb) case ra * 64 + rb: return true;
    CLAT_P2_PAIRS(CLAT_C2_HAVE)
#undef CLAT_C2_HAVE
    default: return false;
  }
}

// row offsets of a column pass stay below the 2^32 - 256 bytes its buffer
// resource covers (4 KiB of margin for the column offset)
bool col_fits(int L, int nsub, long long in_is, size_t si, long long out_is, size_t so) {
  const double lim = 4294967296.0 - 4096.0, rows = (double)L * nsub;
  return rows * (double)in_is * (double)si < lim && rows * (double)out_is * (double)so < lim;
}

// Column pass on the compile-time column kernel; false when the radix pair
// has no code object or an offset would not fit 32 bits (the general kernel
// then runs it).  The conv pass follows the kernel-table layout chosen at
// engine creation.
template <typename TI, typename TC, typename TO, int MODE>
bool launch_col(clat_engine* e, const PassCfg& pc, const P2Args& a, int n_outer, int axis, const void* in,
                void* out, int kind, double bytes) {
  if (MODE == PM_CONV) {
    if (!e->conv_tiled) return false;
  } else if (!e->colk || !col_pair_compiled(pc.R1, pc.R2) ||
             !col_fits(a.L, a.nsub, a.in_is, sizeof(C2<TI>), a.out_is, sizeof(C2<TO>))) {
    return false;
  }
  const int C = col_c<TC>(pc.R1 * pc.R2);
  ColLaunch L;
  const unsigned ncb = (unsigned)((a.W + C - 1) / C);
  const unsigned nbz = (unsigned)std::max(1, a.nb);
  // conv: the batched prepares sit next to each other in blockIdx.x, so the
  // blocks that share a kernel-table tile (same fy, same columns) run at the
  // same time on one XCD (block ids ncb apart, ncb a multiple of 8) and the
  // tile comes from HBM once per launch (conv reads 1142 MB per four-prepare
  // launch against 1348); the y inverse after it also runs faster (1.86 ms
  // per step against 2.02, same box)
  const bool fold = MODE == PM_CONV && !a.outer_x && nbz > 1 && ncb % 8 == 0;
  if (fold) L.grid = dim3(ncb * nbz, (unsigned)n_outer, 1);
  else
    L.grid = a.outer_x ? dim3((unsigned)n_outer, ncb, nbz) : dim3(ncb, (unsigned)n_outer, nbz);
  L.stream = e->stream;
  L.in = in;
  L.out = out;
  ColArgs2& c = L.a;
  c.W = a.W;
  c.od = 1;
  c.in_os_hi = a.in_os;
  c.out_os_hi = a.out_os;
  c.in_os_lo = c.out_os_lo = 0;
  c.in_is_lo = (unsigned)(a.in_is * (long long)sizeof(C2<TI>));
  c.out_is_lo = (unsigned)(a.out_is * (long long)sizeof(C2<TO>));
  c.in_is_hi = c.in_is_lo * (unsigned)pc.R1;
  c.out_is_hi = c.out_is_lo * (unsigned)pc.R1;
  c.gym = a.gym; c.gya = a.gya; c.zpar = a.zpar; c.Y2 = a.Y2; c.GZ1 = a.GZ1;
  c.ncb = e->gc_ncb; c.npar = e->gc_npar; c.par = e->gc_npar == 2 ? a.gxa : 0;
  c.pair = (MODE == PM_CONV && e->gc_npar == 1) ? 1 : 0;
  c.rowmask = a.rowmask;
  c.mask_os = a.mask_os;
  c.in_bs = a.in_bs;
  c.out_bs = a.out_bs;
  c.in_cbs = a.in_cbs ? a.in_cbs : C;
  c.out_cbs = a.out_cbs ? a.out_cbs : C;
  c.outer_x = a.outer_x;
  c.nxb = fold ? (int)ncb : 0;
  c.cset = 0;
  L.T2L = sizeof(TC) == 8 ? (const void*)e->dtw[axis] : (const void*)e->dtw32[axis];
  L.G = gc_of<TC>(e);
  L.Q = nullptr;
  constexpr int combo = combo_of<TI, TC, TO>();
  e->kbegin(kind);
  hipError_t err = hipErrorInvalidValue;
  switch (pc.R1 * 64 + pc.R2) {
#define CLAT_C2_CASE(ra, rb) \
  case ra * 64 + rb: err = clat_c2_##ra##x##rb(combo, MODE, a.nsub, L); break;
    CLAT_P2_PAIRS(CLAT_C2_CASE)
#undef CLAT_C2_CASE
    default: break;
  }
  HIPX(err);
  e->kend(kind, bytes);
  return true;
}

// Coupled conv of a coarse two-coset engine: a.nb prepares a.in_bs elements
// apart, each with its coset planes e->b_set apart in B; otherwise the
// launch shape of the conv on the column kernel (fy pairs, prepares folded
// into blockIdx.x).
template <typename TB, typename TC>
void launch_conv2(clat_engine* e, const PassCfg& pc, const P2Args& a, int n_outer, double bytes) {
  if (!e->Q || (sizeof(TC) == 4) != (e->Gc32 != nullptr))
    throw host::Error(CLAT_EDEVICE, "coupled conv without its kernel tables");
  const int C = col_c<TC>(pc.R1 * pc.R2);
  ColLaunch L;
  const unsigned ncb = (unsigned)((a.W + C - 1) / C);
  const unsigned nbz = (unsigned)std::max(1, a.nb);
  const bool fold = nbz > 1 && ncb % 8 == 0;
  L.grid = fold ? dim3(ncb * nbz, (unsigned)n_outer, 1) : dim3(ncb, (unsigned)n_outer, nbz);
  L.stream = e->stream;
  L.in = e->B;
  L.out = e->B;
  ColArgs2& c = L.a;
  std::memset(&c, 0, sizeof c);
  c.W = a.W;
  c.od = 1;
  c.in_os_hi = a.in_os;
  c.out_os_hi = a.out_os;
  c.in_is_lo = (unsigned)(a.in_is * (long long)sizeof(C2<TB>));
  c.out_is_lo = (unsigned)(a.out_is * (long long)sizeof(C2<TB>));
  c.in_is_hi = c.in_is_lo * (unsigned)pc.R1;
  c.out_is_hi = c.out_is_lo * (unsigned)pc.R1;
  c.gym = 1;
  c.Y2 = a.Y2;
  c.GZ1 = a.GZ1;
  c.ncb = e->gc_ncb;
  c.npar = 1;
  c.pair = 1;
  c.in_bs = a.in_bs;
  c.out_bs = a.out_bs;
  c.in_cbs = a.in_cbs ? a.in_cbs : C;
  c.out_cbs = a.out_cbs ? a.out_cbs : C;
  c.nxb = fold ? (int)ncb : 0;
  c.cset = (long long)e->b_set;
  L.T2L = sizeof(TC) == 8 ? (const void*)e->dtw[2] : (const void*)e->dtw32[2];
  L.G = gc_of<TC>(e);
  L.Q = e->Q;
  constexpr int combo = combo_of<TB, TC, TB>();
  e->kbegin(K_ZCONV);
  hipError_t err = hipErrorInvalidValue;
  switch (pc.R1 * 64 + pc.R2) {
#define CLAT_C2_CASE(ra, rb) \
  case ra * 64 + rb: err = clat_c2_##ra##x##rb(combo, PM_CONV2, 2, L); break;
    CLAT_P2_PAIRS(CLAT_C2_CASE)
#undef CLAT_C2_CASE
    default: break;
  }
  HIPX(err);
  e->kend(K_ZCONV, bytes);
}

template <typename TI, typename TC, typename TO, int MODE, bool LINE>
void launch2(clat_engine* e, const PassCfg& pc, P2Args a, int n_outer, int axis, const void* in, void* out,
             int kind, double bytes) {
  if constexpr (!LINE) {
    if (launch_col<TI, TC, TO, MODE>(e, pc, a, n_outer, axis, in, out, kind, bytes)) return;
  }
  if (a.rowmask || a.nb > 1 || a.in_cbs || a.out_cbs || a.outer_x)
    throw host::Error(CLAT_EDEVICE, "row mask, batch or tiled plane on the general pass kernel");
  a.R1 = pc.R1;
  a.R2 = pc.R2;
  a.R1p = pc.R1p;
  a.mR1 = magic_of((uint32_t)pc.R1);
  a.mR2 = magic_of((uint32_t)pc.R2);
  a.mL = magic_of((uint32_t)a.L);
  a.C = pc.C;
  a.lgC = pc.lgC;
  P2Launch L;
  L.grid = dim3((unsigned)((a.W + pc.C - 1) / pc.C), (unsigned)n_outer);
  L.block = dim3(pc.NT);
  L.lds = pc.lds;
  L.stream = e->stream;
  L.in = in;
  L.out = out;
  L.a = a;
  L.tw = tw_of<TC>(e, axis);
  L.h = h_of<TC>(e, axis);
  L.hy = h_of<TC>(e, 1);
  L.hz = h_of<TC>(e, 2);
  L.G = gc_of<TC>(e);
  constexpr int combo = combo_of<TI, TC, TO>();
  e->kbegin(kind);
  hipError_t err;
  switch (pc.R1 * 64 + pc.R2) {
#define CLAT_P2_CASE(ra, rb) \
  case ra * 64 + rb: err = clat_p2_##ra##x##rb(combo, MODE, LINE, L); break;
    CLAT_P2_PAIRS(CLAT_P2_CASE)
#undef CLAT_P2_CASE
    default: err = p2_launch<0, 0>(combo, MODE, LINE, L); break;
  }
  HIPX(err);
  e->kend(kind, bytes);
}

// x pass of the atom path on the line kernel (occupied lines only)
// in / out: dense side of prepare 0, bs: its element stride to the next
// prepare of the batch (e->cur_nb prepares, sites in e->b_in / e->b_out)
template <typename TI, typename TC, typename TO, int MODE>
void launch_line(clat_engine* e, const void* in, void* out, long long bs, int kind, double bytes) {
  LineLaunch L;
  L.stream = e->stream;
  L.in = in;
  L.out = out;
  L.nb = (unsigned)e->cur_nb;
  LineArgs& a = L.a;
  a.n_live = (int)e->n_live;
  a.lines = e->d_lines;
  a.live_ptr = e->d_live_ptr;
  a.site_ccell = e->d_site_ccell;
  for (int b = 0; b < LINE_NB; ++b) {
    a.site_in[b] = MODE == PM_FWD ? e->b_in[b] : nullptr;
    a.site_out[b] = MODE == PM_INV ? e->b_out[b] : nullptr;
    a.site_in2[b] = MODE == PM_FWD && b < e->cur_nb ? e->b_in2[b] : nullptr;
    a.site_out2[b] = MODE == PM_INV && b < e->cur_nb ? e->b_out2[b] : nullptr;
  }
  if ((double)e->a_set * (double)std::max(e->s_a, e->s_d) >= (double)COL_RANGE - 4096.0)
    throw host::Error(CLAT_EUNSUP, "grid too large for the line kernels");
  const int W = e->wcols;
  a.Yl = e->Yi;
  a.lz = e->Yi * W;
  a.fmax = e->fmax;
  if (e->atile) {
    a.ly = e->gc_c;
    a.lgsw = log2i(e->gc_c);
    a.strip = e->Yi * e->gc_c;
  } else {
    a.ly = W;
    a.lgsw = 0;
    a.strip = 0;
  }
  a.in_bs = a.out_bs = bs;
  // half-length lines: the Xi / 2-point lines' table W_Xi^m, m < Xi, the
  // phase table W_2Xi^f, and the site cells halved (x = 2j + p -> j).  The
  // FP32X forward keeps full lines (see xhalf)
  const bool hl = e->xhalf && (MODE == PM_INV || e->rprec != CLAT_PREC_FP32X);
  L.hl = hl;
  L.ho = !hl && e->fmax == e->Xi;  // half box on full lines: the forward forms f < Xi only
  if (hl) {
    L.T2L = sizeof(TC) == 8 ? e->tw64[0] : e->tw32[0];
    L.PH = sizeof(TC) == 8 ? (const void*)e->dtw[0] : (const void*)e->dtw32[0];
  } else {
    L.T2L = sizeof(TC) == 8 ? (const void*)e->dtw[0] : (const void*)e->dtw32[0];
    L.PH = nullptr;
  }
  constexpr int combo = combo_of<TI, TC, TO>();
  const int R1 = hl ? e->hx_r1 : e->cxf_s.R1, R2 = hl ? e->hx_r2 : e->cxf_s.R2;
  int lines = 0;
  switch (R1 * 64 + R2) {
#define CLAT_L2_N(ra, rb) case ra * 64 + rb: lines = clat_l2_lines_##ra##x##rb(combo); break;
    CLAT_P2_PAIRS(CLAT_L2_N)
#undef CLAT_L2_N
    default: break;
  }
  if (lines <= 0) throw host::Error(CLAT_EDEVICE, "line kernel missing for the x radix pair");
  if (e->n_live == 0) return;
  L.blocks = (unsigned)((e->n_live + lines - 1) / lines);
  e->kbegin(kind);
  hipError_t err = hipErrorInvalidValue;
  switch (R1 * 64 + R2) {
#define CLAT_L2_CASE(ra, rb) case ra * 64 + rb: err = clat_l2_##ra##x##rb(combo, MODE, L); break;
    CLAT_P2_PAIRS(CLAT_L2_CASE)
#undef CLAT_L2_CASE
    default: break;
  }
  HIPX(err);
  e->kend(kind, bytes);
}

P2Args base_args(const clat_engine* e) {
  P2Args a;
  std::memset(&a, 0, sizeof a);
  a.X2 = 2 * e->Xi;
  a.Y2 = 2 * e->Yi;
  a.Z2 = 2 * e->Zi;
  a.GZ1 = e->Zi + 1;
  a.Yl = e->Yi;
  return a;
}

// One prepare() on the two-level kernels.
//   TIN: dense input type, TA: x-forward output, TB: y-forward / conv output,
//   TD: y-inverse output, TV: dense potential type, TC: arithmetic of every
//   pass but the last, TCX: arithmetic of the last (inverse x) pass.
// sites_in / sites_out: pair density at the sites and potential at the sites
// (atom path); when null, the dense F / V grids are used.
template <typename TIN, typename TA, typename TB, typename TD, typename TV, typename TC, typename TCX,
          typename TCZ = TC, typename TCY = TC>
void prepare_fast(clat_engine* e, const double2* sites_in, double2* sites_out) {
  const int Xi = e->Xi, Yi = e->Yi, Zi = e->Zi;
  const double N = (double)e->N();
  const double sa = sizeof(C2<TA>), sb = sizeof(C2<TB>), sd = sizeof(C2<TD>), sv = sizeof(C2<TV>),
               si = sizeof(C2<TIN>);
  const double site_bytes = (double)e->n_sites * 20.0;  // value + cell index
  const int nlines = Yi * Zi;
  const bool atoms = sites_in != nullptr;
  // atom path of the shared schedule: x passes over occupied lines only, y
  // passes never touch the rows of empty lines
  const bool lk = atoms && e->shared() && e->xline;
  const double live = lk ? (double)e->n_live / (double)nlines : 1.0;
  // batched prepares (line-kernel path only): e->cur_nb of them, sites in e->b_in / e->b_out
  const int nb = lk ? e->cur_nb : 1;
  if (nb == 1) {
    e->b_in[0] = sites_in;
    e->b_out[0] = sites_out;
  }
  auto batch = [&](P2Args& a, long long in_bs, long long out_bs) {
    a.nb = nb;
    a.in_bs = in_bs;
    a.out_bs = out_bs;
  };

  auto xfwd = [&](int nsub, int kx, int ky, int kz, void* out) {
    P2Args a = base_args(e);
    a.L = Xi;
    a.nsub = nsub;
    a.W = nlines;
    a.in_os = Xi; a.in_is = 1;
    a.out_os = (long long)nsub * Xi; a.out_is = 1;
    a.rphase = nsub == 1 ? 1 : 0;
    a.kx = kx; a.ky = ky; a.kz = kz;
    const double bytes = (atoms ? site_bytes : si * N) + sa * (nsub == 2 ? (double)e->a_set : N) * live;
    if (lk) {
      launch_line<double, TC, TA, PM_FWD>(e, nullptr, out, (long long)e->a_set, K_XFWD, bytes * nb);
    } else if (atoms) {
      a.site_ptr = e->d_site_ptr;
      a.site_cell = e->d_site_cell;
      a.site_in = sites_in;
      launch2<double, TC, TA, PM_FWD, true>(e, e->cxf_s, a, 1, 0, nullptr, out, K_XFWD, bytes);
    } else {
      launch2<TIN, TC, TA, PM_FWD, true>(e, e->cxf, a, 1, 0, e->F, out, K_XFWD, bytes);
    }
  };
  auto xinv = [&](int nsub, int kx, int ky, int kz, int acc, const void* in) {
    P2Args a = base_args(e);
    a.L = Xi;
    a.nsub = nsub;
    a.W = nlines;
    a.in_os = (long long)nsub * Xi; a.in_is = 1;
    a.out_os = Xi; a.out_is = 1;
    a.rphase = nsub == 1 ? 1 : 0;
    a.kx = kx; a.ky = ky; a.kz = kz;
    a.acc = acc;
    if (lk) {
      launch_line<TD, TCX, double, PM_INV>(e, in, nullptr, (long long)e->d_set, K_XINV,
                                           (sd * (double)e->d_set * live + site_bytes) * nb);
    } else if (atoms) {
      a.site_ptr = e->d_site_ptr;
      a.site_cell = e->d_site_cell;
      a.site_out = sites_out;
      const double bytes = sd * N * nsub + site_bytes + (acc ? 16.0 * e->n_sites : 0.0);
      launch2<TD, TCX, double, PM_INV, true>(e, e->cxi_s, a, 1, 0, in, nullptr, K_XINV, bytes);
    } else {
      const double bytes = sd * N * nsub + sv * N * (acc ? 2 : 1);
      launch2<TD, TCX, TV, PM_INV, true>(e, e->cxi, a, 1, 0, in, e->V, K_XINV, bytes);
    }
  };

  if (e->shared()) {
    const long long W = e->wcols;  // x-frequency columns carried (2 Xi, or the half box)
    // B holds the y-forward output as [fy][z][fx] (the z pass reads rows W
    // apart) or, with e->btile, as [fy][fx / C][z][fx % C]: element (fy, z)
    // of column strip j at fy Z W + j Z C + z C
    const bool bt = e->btile;
    const long long Cb = bt ? e->gc_c : 0;
    if (bt && (col_c<TC>(Yi) != Cb || col_c<TCZ>(Zi) != Cb || col_c<TCY>(Yi) != Cb))
      throw host::Error(CLAT_EDEVICE, "tiled B with column passes of different widths");
    const long long bz = bt ? Cb : W;       // B: z stride
    const long long bcs = bt ? Zi * Cb : 0; // B: column-strip stride (0: C)
    // A and D in strips ([z][fx / C][y][fx % C]) on the line-kernel path
    const bool at = bt && lk && e->atile;
    // with strips the y inverse runs z fastest, so the blocks in flight read
    // whole strips (y inverse 2.16 ms per step against 2.25 column-block
    // fastest; with A in plain rows the y forward measured the other way
    // round, 2.64 against 2.48)
    const int yo = bt ? 1 : 0;
    xfwd(2, 0, 0, 0, e->A);
    {
      P2Args a = base_args(e);
      a.L = Yi; a.nsub = 2; a.W = (int)W;
      a.in_os = (long long)Yi * W; a.in_is = W; a.out_os = bz; a.out_is = (long long)Zi * W;
      a.out_cbs = bcs;
      // with A in strips too, the y forward also walks z fastest (1.98 ms per
      // step against 2.12; the z conv after it 3.81 against 3.90)
      if (at) { a.in_is = Cb; a.in_cbs = (long long)Yi * Cb; a.outer_x = 1; }
      if (lk) { a.rowmask = e->d_linemask; a.mask_os = e->cyf.R1; batch(a, (long long)e->a_set, (long long)e->b_set); }
      launch2<TA, TC, TB, PM_FWD, false>(e, e->cyf, a, Zi, 1, e->A, e->B, K_YFWD,
                                         (sa * (double)e->a_set * live + sb * (double)e->b_set) * nb);
    }
    {
      P2Args a = base_args(e);
      a.L = Zi; a.nsub = 2; a.W = (int)W;
      a.in_os = a.out_os = (long long)Zi * W; a.in_is = a.out_is = bz;
      a.in_cbs = a.out_cbs = bcs;
      a.gxm = 1; a.gym = 1;
      if (lk) batch(a, (long long)e->b_set, (long long)e->b_set);
      // algorithmic bytes: B read and written per prepare, the kernel-table
      // octant once per launch (the batched prepares share it)
      if (e->pair_conv) {
        // two coset planes per prepare: nb / 2 prepares, 2 b_set apart; the
        // cross-coset table read once per launch besides the octant
        if (!lk || nb % 2) throw host::Error(CLAT_EDEVICE, "coupled conv outside a two-coset batch");
        a.nb = nb / 2;
        a.in_bs = a.out_bs = 2LL * (long long)e->b_set;
        launch_conv2<TB, TCZ>(e, e->czc, a, 2 * Yi,
                              sb * 2 * (double)e->b_set * nb + (double)sizeof(TCZ) * (double)e->gc_elems +
                                  (double)sizeof(C2<TCZ>) * (double)e->q_elems);
      } else {
        launch2<TB, TCZ, TB, PM_CONV, false>(e, e->czc, a, 2 * Yi, 2, e->B, e->B, K_ZCONV,
                                             sb * 8 * N * nb + (double)sizeof(TCZ) * (double)e->NG());
      }
    }
    {
      P2Args a = base_args(e);
      a.L = Yi; a.nsub = 2; a.W = (int)W;
      a.in_os = bz; a.in_is = (long long)Zi * W; a.out_os = (long long)Yi * W; a.out_is = W;
      a.in_cbs = bcs;
      if (at) { a.out_is = Cb; a.out_cbs = (long long)Yi * Cb; }
      a.outer_x = yo;
      if (lk) { a.rowmask = e->d_linemask; a.mask_os = e->cyi.R1; batch(a, (long long)e->b_set, (long long)e->d_set); }
      launch2<TB, TCY, TD, PM_INV, false>(e, e->cyi, a, Zi, 1, e->B, e->D, K_YINV,
                                          (sb * (double)e->b_set + sd * (double)e->d_set * live) * nb);
    }
    xinv(2, 0, 0, 0, 0, e->D);
  } else {
    const long long W = Xi;
    for (int round = 0; round < 8; ++round) {
      const int kx = round & 1, ky = (round >> 1) & 1, kz = (round >> 2) & 1;
      xfwd(1, kx, ky, kz, e->A);
      P2Args a = base_args(e);
      a.L = Yi; a.nsub = 1; a.W = (int)W;
      a.in_os = a.out_os = (long long)Yi * W; a.in_is = a.out_is = W;
      launch2<TA, TC, TA, PM_FWD, false>(e, e->cyf, a, Zi, 1, e->A, e->A, K_YFWD, 2 * sa * N);
      P2Args c = base_args(e);
      c.L = Zi; c.nsub = 1; c.W = (int)W;
      c.in_os = c.out_os = W; c.in_is = c.out_is = (long long)Yi * W;
      c.zpar = kz; c.gxm = 2; c.gxa = kx; c.gym = 2; c.gya = ky;
      launch2<TA, TCZ, TA, PM_CONV, false>(e, e->czc, c, Yi, 2, e->A, e->A, K_ZCONV,
                                           2 * sa * N + (double)sizeof(TCZ) * (double)e->NG());
      launch2<TA, TCY, TD, PM_INV, false>(e, e->cyi, a, Zi, 1, e->A, e->D, K_YINV, (sa + sd) * N);
      xinv(1, kx, ky, kz, round > 0 ? 1 : 0, e->D);
    }
  }
}

// ===========================================================================
// generic (Stockham, single precision type) path
// ===========================================================================
int tile_lines(const clat_engine* e, int nsub, int L, int cap = 64) {
  const int c = floor_pow2(std::max(1, e->tile_budget / (nsub * L)));
  return std::min(c, cap);
}

template <typename T, int EMAX>
void prepare_generic(clat_engine* e) {
  using CT = C2<T>;
  const int Xi = e->Xi, Yi = e->Yi, Zi = e->Zi;
  const size_t N = e->N();
  const double s = (double)sizeof(CT);
  const CT* twx = tw_of<T>(e, 0);
  const CT* twy = tw_of<T>(e, 1);
  const CT* twz = tw_of<T>(e, 2);
  const CT* hx = h_of<T>(e, 0);
  const CT* hy = h_of<T>(e, 1);
  const CT* hz = h_of<T>(e, 2);
  const T* G = g_of<T>(e);
  CT* F = (CT*)e->F;
  CT* V = (CT*)e->V;
  CT* A = (CT*)e->A;
  CT* B = (CT*)e->B;
  const int nlines = Yi * Zi;

  auto xfwd = [&](int nsub, int kx, int ky, int kz, CT* out) {
    const int C = tile_lines(e, nsub, Xi);
    const size_t lds = (size_t)nsub * C * Xi * sizeof(CT);
    dim3 grid((nlines + C - 1) / C);
    e->kbegin(K_XFWD);
    hipLaunchKernelGGL((k_xfwd<T, EMAX>), grid, dim3(NT), lds, e->stream, F, out, e->px, twx, hx, nlines, C,
                       log2i(C), nsub, kx, ky, kz, hy, hz, Yi);
    LAUNCH_CHECK();
    e->kend(K_XFWD, s * (double)N * (1 + nsub));
  };
  auto xinv = [&](int nsub, int kx, int ky, int kz, int acc, const CT* in) {
    const int C = tile_lines(e, nsub, Xi);
    const size_t lds = (size_t)nsub * C * Xi * sizeof(CT);
    dim3 grid((nlines + C - 1) / C);
    e->kbegin(K_XINV);
    hipLaunchKernelGGL((k_xinv<T, EMAX>), grid, dim3(NT), lds, e->stream, in, V, e->px, twx, hx, nlines, C,
                       log2i(C), nsub, kx, ky, kz, hy, hz, Yi, acc);
    LAUNCH_CHECK();
    e->kend(K_XINV, s * (double)N * (1 + nsub + (acc ? 1 : 0)));
  };
  auto col = [&](int mode, const AxisPlan& p, const CT* tw, const CT* h, ColArgs a, int n_outer, const CT* in,
                 CT* out) {
    const int L = p.L;
    a.C = tile_lines(e, a.nsub, L, 64);
    a.C = std::min(a.C, ceil_pow2(a.W));
    a.lgC = log2i(a.C);
    const size_t lds = (size_t)a.nsub * a.C * L * sizeof(CT);
    dim3 grid((a.W + a.C - 1) / a.C, n_outer);
    const int kind0 = mode == COL_FWD ? K_YFWD : (mode == COL_CONV ? K_ZCONV : K_YINV);
    e->kbegin(kind0);
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

  if (e->shared()) {
    const long long W = e->wcols;  // x-frequency columns carried (2 Xi, or the half box)
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

// ===========================================================================
// buffers and prepare dispatch
// ===========================================================================
void zero_grid(clat_engine* e, void* p, size_t n, size_t csz) {
  if (csz == 8)
    hipLaunchKernelGGL(k_zero<float>, dim3(grid1d(n)), dim3(256), 0, e->stream, (float2*)p, n);
  else
    hipLaunchKernelGGL(k_zero<double>, dim3(grid1d(n)), dim3(256), 0, e->stream, (double2*)p, n);
  LAUNCH_CHECK();
}

// dense F and V of the grid protocol (and of the generic path)
void ensure_dense(clat_engine* e) {
  if (!e->F) {
    e->F = e->dalloc(e->f_bytes);
    zero_grid(e, e->F, e->N(), e->s_in);
  }
  if (!e->V) e->V = e->dalloc(e->v_bytes);
}

// The mixed mode: complex64 storage; FP64 arithmetic in the last pass only
// (mix 0) or also in the y-inverse (1), the conv (2) and the forward (3) passes.
void fp32x_prepare(clat_engine* e, const double2* si, double2* so) {
  switch (e->mix) {
    case 0: prepare_fast<double, float, float, float, double, float, double, float, float>(e, si, so); break;
    case 1: prepare_fast<double, float, float, double, double, float, double, float, double>(e, si, so); break;
    case 2: prepare_fast<double, float, float, double, double, float, double, double, double>(e, si, so); break;
    default: prepare_fast<double, float, float, double, double, double, double, double, double>(e, si, so); break;
  }
}

void ensure_fine_tables(clat_engine* e);

// dense F -> dense V
bool ensure_table(clat_engine* e);

void prepare_dense(clat_engine* e) {
  ensure_table(e);
  ensure_dense(e);
  ensure_fine_tables(e);
  if (e->fast) {
    if (e->rprec == CLAT_PREC_FP64) prepare_fast<double, double, double, double, double, double, double>(e, nullptr, nullptr);
    else if (e->rprec == CLAT_PREC_FP32) prepare_fast<float, float, float, float, float, float, float>(e, nullptr, nullptr);
    else fp32x_prepare(e, nullptr, nullptr);
  } else {
    if (e->rprec == CLAT_PREC_FP32) prepare_generic<float, 32>(e);
    else prepare_generic<double, 16>(e);
  }
}

void coset_prepare(clat_engine* e, const double2* rho_sites, double2* v_sites);

// pair density at the sites -> potential at the sites (FP64)
void prepare_sites(clat_engine* e, const double2* rho_sites, double2* v_sites) {
  if (e->coset) {
    coset_prepare(e, rho_sites, v_sites);
    return;
  }
  if (e->fast) {
    if (e->rprec == CLAT_PREC_FP64) prepare_fast<double, double, double, double, double, double, double>(e, rho_sites, v_sites);
    else if (e->rprec == CLAT_PREC_FP32) prepare_fast<float, float, float, float, float, float, float>(e, rho_sites, v_sites);
    else fp32x_prepare(e, rho_sites, v_sites);
    return;
  }
  ensure_dense(e);
  zero_grid(e, e->F, e->N(), e->s_in);
  e->kbegin(K_SCATTER);
  if (e->s_in == 8)
    hipLaunchKernelGGL(k_scatter<float>, dim3(grid1d(e->n_sites)), dim3(256), 0, e->stream, (float2*)e->F,
                       e->d_site_cell, rho_sites, e->n_sites);
  else
    hipLaunchKernelGGL(k_scatter<double>, dim3(grid1d(e->n_sites)), dim3(256), 0, e->stream, (double2*)e->F,
                       e->d_site_cell, rho_sites, e->n_sites);
  LAUNCH_CHECK();
  e->kend(K_SCATTER, (double)e->n_sites * (20.0 + (double)e->s_in));
  if (e->rprec == CLAT_PREC_FP32) prepare_generic<float, 32>(e);
  else prepare_generic<double, 16>(e);
  e->kbegin(K_GATHER);
  if (e->s_v == 8)
    hipLaunchKernelGGL(k_gather<float>, dim3(grid1d(e->n_sites)), dim3(256), 0, e->stream, (const float2*)e->V,
                       e->d_site_cell, v_sites, e->n_sites);
  else
    hipLaunchKernelGGL(k_gather<double>, dim3(grid1d(e->n_sites)), dim3(256), 0, e->stream, (const double2*)e->V,
                       e->d_site_cell, v_sites, e->n_sites);
  LAUNCH_CHECK();
  e->kend(K_GATHER, (double)e->n_sites * (20.0 + (double)e->s_v));
}

// ===========================================================================
// interaction table: G -> DCT-I -> /(8 Xi Yi Zi)
// ===========================================================================
void dct_axis(clat_engine* e, int axis, double scale) {
  const int X1 = e->Xi + 1, Y1 = e->Yi + 1, Z1 = e->Zi + 1;
  const AxisPlan& p = axis == 0 ? e->dx : (axis == 1 ? e->dy : e->dz);
  const int L2 = p.L;
  const int C = std::min(64, floor_pow2(std::max(1, 4096 / L2)));
  const size_t lds = (size_t)C * L2 * sizeof(double2);
  e->kbegin(K_DCT);
  if (axis == 0) {
    const int ncols = Y1 * Z1;
    dim3 grid((ncols + 2 * C - 1) / (2 * C));
    hipLaunchKernelGGL((k_dct1<16, true>), grid, dim3(NT), lds, e->stream, e->G64, p, e->dtw[0], ncols, C,
                       log2i(C), 0LL, 1LL, (long long)X1, ncols, scale);
  } else if (axis == 1) {
    const int ncols = X1;
    dim3 grid(((ncols + 2 * C - 1) / (2 * C)) * Z1);
    hipLaunchKernelGGL((k_dct1<16, false>), grid, dim3(NT), lds, e->stream, e->G64, p, e->dtw[1], ncols, C,
                       log2i(C), (long long)X1 * Y1, (long long)X1, 1LL, ncols, scale);
  } else {
    const int ncols = X1 * Y1;
    dim3 grid((ncols + 2 * C - 1) / (2 * C));
    hipLaunchKernelGGL((k_dct1<16, false>), grid, dim3(NT), lds, e->stream, e->G64, p, e->dtw[2], ncols, C,
                       log2i(C), 0LL, (long long)X1 * Y1, 1LL, ncols, scale);
  }
  LAUNCH_CHECK();
  e->kend(K_DCT, 16.0 * (double)e->NG());
}

// DCT-I of e->G64 and DCT-II of e->Gt64 along one axis in one launch (the
// two-coset tables; same line layout as dct_axis)
void dct12_axis(clat_engine* e, int axis, double scale) {
  const int X1 = e->Xi + 1, Y1 = e->Yi + 1, Z1 = e->Zi + 1;
  const AxisPlan& p = axis == 0 ? e->dx : (axis == 1 ? e->dy : e->dz);
  const int L2 = p.L;
  const int C = std::min(64, floor_pow2(std::max(1, 4096 / L2)));
  const size_t lds = (size_t)C * L2 * sizeof(double2);
  e->kbegin(K_DCT);
  if (axis == 0) {
    const int ncols = Y1 * Z1, nblk = (ncols + 2 * C - 1) / (2 * C);
    hipLaunchKernelGGL((k_dct12<16, true>), dim3(2 * nblk), dim3(NT), lds, e->stream, e->G64, e->Gt64, p, e->dtw[0],
                       nblk, C, log2i(C), 0LL, 1LL, (long long)X1, ncols, scale);
  } else if (axis == 1) {
    const int ncols = X1, nblk = ((ncols + 2 * C - 1) / (2 * C)) * Z1;
    hipLaunchKernelGGL((k_dct12<16, false>), dim3(2 * nblk), dim3(NT), lds, e->stream, e->G64, e->Gt64, p,
                       e->dtw[1], nblk, C, log2i(C), (long long)X1 * Y1, (long long)X1, 1LL, ncols, scale);
  } else {
    const int ncols = X1 * Y1, nblk = (ncols + 2 * C - 1) / (2 * C);
    hipLaunchKernelGGL((k_dct12<16, false>), dim3(2 * nblk), dim3(NT), lds, e->stream, e->G64, e->Gt64, p,
                       e->dtw[2], nblk, C, log2i(C), 0LL, (long long)X1 * Y1, 1LL, ncols, scale);
  }
  LAUNCH_CHECK();
  e->kend(K_DCT, 32.0 * (double)e->NG());
}

void tile_conv_table(clat_engine* e);

// Kernel tables of the coarse engine of a two-coset path.  From the
// real-space fine table (G64 before its DCT): the within-coset kernel G(2d)
// through the coarse engine's DCT-I, the cross-coset kernel G(2d + (1,1,1))
// through a DCT-II, both on the coarse octant (an eighth of the fine one).
// From a fine table already transformed: the eight half-box shifts of it
// (k_coarse_g0, k_coarse_st).  Same octants either way; they are then tiled
// for the coarse engine's column layout (all 2 Xc x frequencies, or the half
// box with the kernels summed over the half-box shift).
void build_coset_tables(clat_engine* e) {
  clat_engine* s = e->sub;
  s->stream = e->stream;
  if (e->fine_pending) {
    if (e->g64_real)
      hipLaunchKernelGGL(k_coset_samples, dim3(grid1d(s->NG())), dim3(256), 0, e->stream, e->G64, e->Xi, e->Yi,
                         e->Zi, s->G64, s->Gt64);
    else
      hipLaunchKernelGGL(k_coset_gtable, dim3(grid1d(s->NG())), dim3(256), 0, e->stream, e->gdefer, e->Xi, e->Yi,
                         e->Zi, s->G64, s->Gt64);
    LAUNCH_CHECK();
    const double sc = 1.0 / (8.0 * (double)s->Xi * (double)s->Yi * (double)s->Zi);
    dct12_axis(s, 0, 1.0);
    dct12_axis(s, 1, 1.0);
    dct12_axis(s, 2, sc);
  } else {
    hipLaunchKernelGGL(k_coarse_g0, dim3(grid1d(s->NG())), dim3(256), 0, e->stream, e->G64, e->Xi, e->Yi, e->Zi,
                       s->G64);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_coarse_st, dim3(grid1d(s->NG())), dim3(256), 0, e->stream, e->G64, e->Xi, e->Yi, e->Zi,
                       s->Gt64);
    LAUNCH_CHECK();
  }
  // the conv tables of the current column layout from the two octants
  const unsigned gq = grid1d(s->q_elems), gg = grid1d(s->gc_elems);
  if (s->half) {
    if (s->Gc32) {
      hipLaunchKernelGGL(k_gconv_tiled_half<float>, dim3(gg), dim3(256), 0, e->stream, s->G64, s->Xi, s->Yi, s->Zi,
                         s->gc_c, s->gc_ncb, s->Gc32);
      hipLaunchKernelGGL(k_coarse_q_half<float>, dim3(gq), dim3(256), 0, e->stream, s->Gt64, s->ph64, s->Xi, s->Yi,
                         s->Zi, s->gc_c, s->gc_ncb, (float2*)s->Q);
    } else {
      hipLaunchKernelGGL(k_gconv_tiled_half<double>, dim3(gg), dim3(256), 0, e->stream, s->G64, s->Xi, s->Yi,
                         s->Zi, s->gc_c, s->gc_ncb, s->Gc64);
      hipLaunchKernelGGL(k_coarse_q_half<double>, dim3(gq), dim3(256), 0, e->stream, s->Gt64, s->ph64, s->Xi, s->Yi,
                         s->Zi, s->gc_c, s->gc_ncb, (double2*)s->Q);
    }
  } else {
    tile_conv_table(s);
    if (s->Gc32)
      hipLaunchKernelGGL(k_coarse_q<float>, dim3(gq), dim3(256), 0, e->stream, s->Gt64, s->ph64, s->Xi, s->Yi, s->Zi,
                         s->gc_c, s->gc_ncb, (float2*)s->Q);
    else
      hipLaunchKernelGGL(k_coarse_q<double>, dim3(gq), dim3(256), 0, e->stream, s->Gt64, s->ph64, s->Xi, s->Yi,
                         s->Zi, s->gc_c, s->gc_ncb, (double2*)s->Q);
  }
  LAUNCH_CHECK();
  s->initialized = true;
}

// the fine table's DCT-I (and FP32 copy) and conv table, once per interaction
void ensure_fine_tables(clat_engine* e) {
  if (!e->fine_pending) return;
  if (!e->g64_real) {  // the interaction's table, not built on the two-coset path
    hipLaunchKernelGGL(k_gtable, dim3(grid1d(e->NG())), dim3(256), 0, e->stream, e->G64, e->gdefer);
    LAUNCH_CHECK();
    e->g64_real = true;
  }
  dct_axis(e, 0, 1.0);
  dct_axis(e, 1, 1.0);
  dct_axis(e, 2, 1.0 / (8.0 * (double)e->Xi * (double)e->Yi * (double)e->Zi));
  if (e->G32) {
    hipLaunchKernelGGL(k_d2f, dim3(grid1d(e->NG())), dim3(256), 0, e->stream, e->G64, e->G32, e->NG());
    LAUNCH_CHECK();
  }
  tile_conv_table(e);
  e->fine_pending = false;
}

// G64 holds a new real-space table: on the two-coset path only the coarse
// tables are built (the fine ones wait for a single-grid prepare)
void finish_interaction(clat_engine* e) {
  e->fine_pending = true;
  if (e->coset) build_coset_tables(e);
  else ensure_fine_tables(e);
  e->initialized = true;
  e->prepared = false;
}

// the conv pass's copy of the kernel table (tiled per column block, or the
// row layout of the general kernel)
void tile_conv_table(clat_engine* e) {
  if (e->conv_tiled) {
    if (e->Gc64)
      hipLaunchKernelGGL(k_gconv_tiled<double>, dim3(grid1d(e->gc_elems)), dim3(256), 0, e->stream, e->G64,
                         e->Gc64, e->Xi, e->Yi + 1, e->Zi + 1, e->gc_c, e->gc_ncb, e->gc_npar, e->gc_w);
    else
      hipLaunchKernelGGL(k_gconv_tiled<float>, dim3(grid1d(e->gc_elems)), dim3(256), 0, e->stream, e->G64,
                         e->Gc32, e->Xi, e->Yi + 1, e->Zi + 1, e->gc_c, e->gc_ncb, e->gc_npar, e->gc_w);
    LAUNCH_CHECK();
  } else if (e->Gc64) {
    hipLaunchKernelGGL(k_gconv<double>, dim3(grid1d(e->gc_elems)), dim3(256), 0, e->stream, e->G64, e->Gc64,
                       e->Xi, e->Yi + 1, e->Zi + 1);
    LAUNCH_CHECK();
  } else if (e->Gc32) {
    hipLaunchKernelGGL(k_gconv<float>, dim3(grid1d(e->gc_elems)), dim3(256), 0, e->stream, e->G64, e->Gc32,
                       e->Xi, e->Yi + 1, e->Zi + 1);
    LAUNCH_CHECK();
  }
}

GArgs base_gargs(const clat_engine* e) {
  GArgs g{};
  g.X1 = e->Xi + 1; g.Y1 = e->Yi + 1; g.Z1 = e->Zi + 1;
  g.ux = e->X + 1; g.uy = e->Y + 1; g.uz = e->Z + 1;
  return g;
}

// A new interaction: its G table + DCT-I are enqueued by the first call
// that needs them (ensure_table), which in an integral run is after the
// run's pair densities, so the device has work while the host launches the
// table's kernels
void build_table(clat_engine* e, GArgs g) {
  e->gdefer = g;
  e->table_deferred = true;
  e->initialized = true;
  e->prepared = false;
}

// the deferred G table + DCT-I on the engine stream, timed into stage_ms[0];
// true when it was enqueued by this call
bool ensure_table(clat_engine* e) {
  if (!e->table_deferred) return false;
  HIPX(hipEventRecord(e->stage_ev[0], e->stream));
  // the fine real-space table is built only where a fine-grid pass needs it
  // (the two-coset path samples G(2d) and G(2d + 1) from the arguments)
  e->g64_real = false;
  finish_interaction(e);
  e->table_deferred = false;
  HIPX(hipEventRecord(e->stage_ev[5], e->stream));
  // no host wait here: the next call that waits for the stream reads the time
  e->table_pending = true;
  return true;
}

// time of the last table build, once the stream has passed it
void settle_table(clat_engine* e) {
  if (!e->table_pending) return;
  float ms = 0.f;
  HIPX(hipEventSynchronize(e->stage_ev[5]));
  HIPX(hipEventElapsedTime(&ms, e->stage_ev[0], e->stage_ev[5]));
  e->stage_ms[0] = ms;
  e->table_pending = false;
}

void ensure_dpart(clat_engine* e, size_t n) {
  if (e->dpart_cap < n) {
    e->dfree(e->dpart, e->dpart_cap * sizeof(double));
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

// ===========================================================================
// atom-site path helpers
// ===========================================================================
void build_products(clat_engine* e) {
  const int P = e->n_states * (e->n_states + 1) / 2;
  if (e->rho_states == e->n_states && e->rho) return;
  // the buffer is kept between runs: a free here would synchronise the device
  const size_t need = std::max<size_t>(16, (size_t)P * e->n_sites * sizeof(double2));
  if (e->rho_bytes < need) {
    e->dfree(e->rho, e->rho_bytes);
    e->rho = (double2*)e->dalloc(need);
    e->rho_bytes = need;
  }
  e->rho_states = e->n_states;
  if (e->n_sites == 0) return;
  // S sites per block (power of two, <= 64): all states of S + 1 sites in at
  // most 22.5 KB of LDS, so that six blocks share a CU (8 states: 0.59 ms per
  // step against 0.72 with 44 KB tiles), but at least two sites, or too few
  // threads load (32 states: 2.96 ms against 4.10 with one site), and four
  // where they fit 64 KB (32 states x 20 orbitals: 1.6 against 2.8 ms with two)
  const size_t per_site = (size_t)e->n_states * e->n_orb * sizeof(double2);
  int S = 64;
  while (S > 1 && (size_t)(S + 1) * per_site > 23040) S /= 2;
  if (S == 1 && 3 * per_site <= 45056) S = 2;
  if (S < 4 && 5 * per_site <= 65536) S = 4;
  e->kbegin(K_PRODUCTS);
  if ((size_t)(S + 1) * per_site <= 65536) {
    const unsigned nb = (unsigned)((e->n_sites + S - 1) / S);
    if (e->n_states <= 8)
      hipLaunchKernelGGL(k_products<false>, dim3(nb), dim3(NT), (size_t)(S + 1) * per_site, e->stream, e->coeffs,
                         e->n_sites, e->n_states, e->n_orb, S, e->rho);
    else
      hipLaunchKernelGGL(k_products<true>, dim3(nb), dim3(NT), (size_t)(S + 1) * per_site, e->stream, e->coeffs,
                         e->n_sites, e->n_states, e->n_orb, S, e->rho);
  } else {
    hipLaunchKernelGGL(k_products_direct, dim3(grid1d((size_t)P * e->n_sites)), dim3(256), 0, e->stream,
                       e->coeffs, e->n_sites, e->n_states, e->n_orb, e->rho);
  }
  LAUNCH_CHECK();
  e->kend(K_PRODUCTS, (double)e->n_sites * (double)per_site + (double)P * e->n_sites * 16.0);
}

// out[a] = vs[atom_site[a]]
__global__ void k_atoms_from_sites(const double2* __restrict__ vs, const int32_t* __restrict__ atom_site,
                                   double2* __restrict__ out, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    out[t] = vs[atom_site[t]];
}

// One plain prepare per launch with no packed second density: the state the
// pass launches read outside a batched run.
void reset_batch(clat_engine* e) {
  e->cur_nb = 1;
  for (int b = 0; b < LINE_NB; ++b) {
    e->b_in2[b] = nullptr;
    e->b_out2[b] = nullptr;
  }
}

struct BatchScope {
  clat_engine* e;
  explicit BatchScope(clat_engine* x) : e(x) {}
  ~BatchScope() { reset_batch(e); }
  BatchScope(const BatchScope&) = delete;
  BatchScope& operator=(const BatchScope&) = delete;
};

// The e->cur_nb prepares of a two-coset engine (sites in e->b_in / e->b_out,
// or rho_sites / v_sites for a single prepare) on its coarse engine: the
// coset-0 and coset-1 halves of each prepare's site arrays are planes 2b and
// 2b + 1 of one coarse batch.
void coset_prepare(clat_engine* e, const double2* rho_sites, double2* v_sites) {
  clat_engine* s = e->sub;
  const int nb = e->cur_nb;
  if (!s || 2 * nb > LINE_NB || 2 * nb > s->nb_max)
    throw host::Error(CLAT_EDEVICE, "two-coset batch exceeds the coarse engine's buffers");
  // the coarse passes go wherever this engine's launches go now (a group
  // capture points the member's stream at the lead's)
  s->stream = e->stream;
  if (nb == 1) {
    e->b_in[0] = rho_sites;
    e->b_out[0] = v_sites;
  }
  BatchScope guard(s);
  s->cur_nb = 2 * nb;
  const size_t nu = (size_t)e->n_u;
  for (int b = 0; b < nb; ++b)
    for (int c = 0; c < 2; ++c) {
      const int j = 2 * b + c;
      s->b_in[j] = e->b_in[b] + c * nu;
      s->b_out[j] = e->b_out[b] + c * nu;
      s->b_in2[j] = e->b_in2[b] ? e->b_in2[b] + c * nu : nullptr;
      s->b_out2[j] = e->b_out2[b] ? e->b_out2[b] + c * nu : nullptr;
    }
  prepare_sites(s, s->b_in[0], s->b_out[0]);
}

// V of state f's self density at every atom's cell (potentials.cpp:140-158):
// the self density at the sites into rho, its potential at the sites into vs,
// the per-atom values into va and from there to the host array out.  The
// three device buffers hold n_sites, n_sites and n_atoms elements; the caller
// owns them.  Synchronous.
void potential_run(clat_engine* e, int f, double2* rho, double2* vs, double2* va, double* out) {
  reset_batch(e);
  if (e->n_sites > 0) {
    hipLaunchKernelGGL(k_self_density, dim3(grid1d(e->n_sites)), dim3(256), 0, e->stream, e->coeffs, e->n_sites,
                       e->n_orb, f, rho);
    LAUNCH_CHECK();
    ensure_table(e);
    prepare_sites(e, rho, vs);
  }
  if (e->n_atoms > 0) {
    hipLaunchKernelGGL(k_atoms_from_sites, dim3(grid1d(e->n_atoms)), dim3(256), 0, e->stream, vs, e->d_atom_site,
                       va, e->n_atoms);
    LAUNCH_CHECK();
    HIPX(hipMemcpyAsync(out, va, (size_t)e->n_atoms * sizeof(double2), hipMemcpyDeviceToHost, e->stream));
  }
  HIPX(hipStreamSynchronize(e->stream));
}

}  // namespace

// ---------------------------------------------------------------------------
// scratch of the allocation-free potential entry point: one static device
// array, held by every GPU the library runs on until the process ends.  Host
// code never uses its name as an address (that is a host-side stand-in); it
// uses the per-device address that hipGetSymbolAddress returns.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_ELEMS = (size_t)1 << 22;  // double2 elements: 64 MiB
static __device__ __align__(256) double2 scratch[NOALLOC_ELEMS];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double2* g_scratch_ptr[MAX_DEVICES] = {};

// The address of `scratch` on device dev, looked up on the first call for
// that device with the device current (each GPU holds its own copy).
double2* scratch_of(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) throw host::Error(CLAT_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    HIPX(hipGetDevice(&cur));
    if (cur != dev) HIPX(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    HIPX(rc);
    if (!p) throw host::Error(CLAT_EDEVICE, "scratch symbol has no device address");
    g_scratch_ptr[dev] = (double2*)p;
  }
  return g_scratch_ptr[dev];
}

// ===========================================================================
// device stage of clat_compute_integrals: prepares of the cover, overlaps
// ===========================================================================
using Units = std::vector<std::array<int, 2>>;  // (slot, second real slot or -1) per prepare

// Overlap launch shape: 64 x 64 tiles on FP64 MFMA when both sides fill
// them; else 36 x 36 tiles on vector FMA when both sides fit one, 48 x 48
// otherwise (or Kahan-compensated 48 x 48 when asked for)
struct OvlCfg {
  bool mfma = false, kahan = false, t36 = false;
  bool sym = false;  // MFMA tiles on and above the diagonal only (every product a left slot)
  int OV_T = 48, tl = 0, tr = 0, groups = 1;
  int64_t n_chunks = 1, chunk = 1, n_slices = 1;
  size_t n_out = 0;
};

OvlCfg overlap_cfg(const clat_engine* e, int nL, int P) {
  OvlCfg c;
  c.mfma = e->ovl_mode == 2 || (e->ovl_mode == 0 && nL >= OVM_MIN && P >= OVM_MIN);
  c.kahan = !c.mfma && e->ovl_mode == 4;
  c.t36 = !c.mfma && !c.kahan && e->ovl_mode != 3 && nL <= 36 && P <= 36;
  c.OV_T = c.mfma ? OVM_T : c.t36 ? 36 : 48;
  const int ovs = c.mfma ? OVM_S : c.t36 ? 24 : OV_S;
  c.groups = c.t36 ? 3 : 1;  // site groups per block, a partial slice each
  c.tl = (nL + c.OV_T - 1) / c.OV_T;
  c.tr = (P + c.OV_T - 1) / c.OV_T;
  c.sym = c.mfma && nL == P && e->ovl_sym;
  const int tiles = c.sym ? c.tl * (c.tl + 1) / 2 : c.tl * c.tr;
  // blocks: about 2048 for the 48 x 48 and MFMA tiles; for the 36 x 36
  // tiles (246 VGPRs: two blocks per CU) one full round, two per CU
  // (0.33 ms against 0.37 at 2.7 blocks per CU and 0.45 at 8, 1M atoms)
  const int want_blocks = c.t36 ? 2 * e->n_cu : 2048;
  int64_t n_chunks = std::max<int64_t>(1, want_blocks / std::max(1, tiles));
  n_chunks = std::min<int64_t>(n_chunks, (e->n_sites + ovs - 1) / ovs);
  n_chunks = std::max<int64_t>(1, n_chunks);
  int64_t chunk = (e->n_sites + n_chunks - 1) / n_chunks;
  chunk = std::max<int64_t>(ovs, (chunk + ovs - 1) / ovs * ovs);
  c.chunk = chunk;
  c.n_chunks = std::max<int64_t>(1, (e->n_sites + chunk - 1) / chunk);
  c.n_slices = c.n_chunks * c.groups;
  c.n_out = (size_t)std::max(1, nL) * P * 4;
  return c;
}

void ensure_overlap_buffers(clat_engine* e, const OvlCfg& c) {
  if (e->part_cap < c.n_out * c.n_slices) {
    e->dfree(e->part, e->part_cap * sizeof(double));
    e->part = (double*)e->dalloc(c.n_out * c.n_slices * sizeof(double));
    e->part_cap = c.n_out * c.n_slices;
  }
  if (e->ovl_cap < c.n_out) {
    e->dfree(e->ovl, e->ovl_cap * sizeof(double));
    e->ovl = (double*)e->dalloc(c.n_out * sizeof(double));
    e->ovl_cap = c.n_out;
  }
}

// the (slot, product) overlap table into e->ovl (chunk partials, then their
// compensated sum)
void launch_overlaps(clat_engine* e, int nL, int P, const OvlCfg& c) {
  e->kbegin(K_OVERLAP);
  const dim3 og(c.tl, c.tr, (unsigned)c.n_chunks);
  if (c.mfma && c.sym)
    hipLaunchKernelGGL(k_overlap_partial_mfma, dim3((unsigned)(c.tl * (c.tl + 1) / 2), 1, (unsigned)c.n_chunks),
                       dim3(256), 0, e->stream, e->Vat, e->rho, nL, P, e->n_sites, c.chunk, e->part, c.tl);
  else if (c.mfma)
    hipLaunchKernelGGL(k_overlap_partial_mfma, og, dim3(256), 0, e->stream, e->Vat, e->rho, nL, P, e->n_sites,
                       c.chunk, e->part, 0);
  else if (c.t36)
    hipLaunchKernelGGL((k_overlap_partial<9, 4, 3, 24>), og, dim3(243), 0, e->stream, e->Vat, e->rho, nL, P,
                       e->n_sites, c.chunk, e->part);
  else if (c.kahan)
    hipLaunchKernelGGL((k_overlap_partial<16, 3, 1, OV_S, true>), og, dim3(256), 0, e->stream, e->Vat, e->rho, nL,
                       P, e->n_sites, c.chunk, e->part);
  else
    hipLaunchKernelGGL((k_overlap_partial<16, 3, 1, OV_S>), og, dim3(256), 0, e->stream, e->Vat, e->rho, nL, P,
                       e->n_sites, c.chunk, e->part);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_overlap_reduce, dim3((unsigned)((c.n_out + 63) / 64)), dim3(1024), 0, e->stream, e->part,
                     c.n_slices, (int64_t)c.n_out, e->ovl);
  LAUNCH_CHECK();
  e->kend(K_OVERLAP, (double)(c.sym ? c.tl * (c.tl + 1) / 2 : c.tl * c.tr) * (double)(2 * c.OV_T) *
                        (double)e->n_sites * 16.0);
}

// every prepare of the run, nbt per pass launch: pair density of slot
// slots[s] -> potential of slot s at the sites (Vat)
void run_prepares(clat_engine* e, const Units& units, const std::vector<int>& slots, int nbt) {
  const int nU = (int)units.size();
  for (int u = 0; u < nU;) {
    const int nb = std::min(nbt, nU - u);
    // the batch lives in engine state for the pass launches; the guard
    // returns it to one plain prepare however this scope is left
    BatchScope batch(e);
    e->cur_nb = nb;
    for (int b = 0; b < nb; ++b) {
      const int s0 = units[(size_t)(u + b)][0], s1 = units[(size_t)(u + b)][1];
      e->b_in[b] = e->rho + (size_t)slots[(size_t)s0] * e->n_sites;
      e->b_out[b] = e->Vat + (size_t)s0 * e->n_sites;
      e->b_in2[b] = s1 >= 0 ? e->rho + (size_t)slots[(size_t)s1] * e->n_sites : nullptr;
      e->b_out2[b] = s1 >= 0 ? e->Vat + (size_t)s1 * e->n_sites : nullptr;
    }
    prepare_sites(e, e->b_in[0], e->b_out[0]);
    u += nb;
  }
}

// Measured batch sizes, shared by every engine of the process: the key is
// what the pass launches depend on (grid, arithmetic, layout, occupied lines,
// buffer capacity, device size).
std::mutex g_tune_mu;
std::map<std::vector<int64_t>, int> g_tuned;

std::vector<int64_t> tune_key(const clat_engine* e) {
  return {e->Xi, e->Yi, e->Zi, e->rprec, e->mix, e->sched, e->nb_max, e->n_live, e->n_sites,
          e->atile ? 1 : 0, e->btile ? 1 : 0, e->n_cu, e->coset ? e->sub->n_live : -1,
          e->coset && e->sub->half ? 1 : 0, e->coset && e->sub->xhalf ? 1 : 0};
}

// The fixed rule's prepares per launch: the engine's own, or on the
// two-coset path as many as keep a launch's B planes (both cosets) within 3/4
// of the 256 MiB Infinity Cache, so that the conv reads what the y forward
// just wrote from it (1M-atom step, 32 MiB of B per prepare: conv 0.45 ms
// per step at six, 0.51 at seven, 0.57 at eight; step 2.84, 2.86 and 2.87
// ms on one box); where fewer than four fit, eight where sixteen coarse B
// planes take at most 1.1 GB, else four
constexpr size_t MALL_BYTES = (size_t)256 << 20;
int batch_rule(const clat_engine* e) {
  if (!e->coset || std::getenv("CLAT_BATCH")) return e->nb_rule;
  const size_t per = 2 * e->sub->b_set * e->sub->s_b;
  const int fit = (int)std::min<size_t>(8, MALL_BYTES / 4 * 3 / std::max<size_t>(1, per));
  if (fit >= 4) return std::min(e->nb_max, fit);
  return std::min(e->nb_max, 8 * per <= ((size_t)1100 << 20) ? 8 : 4);
}

// Prepares per pass launch for this run (the autotuner, the counterpart of
// the reference's FFTW_MEASURE plans, CoulombCalculator.cpp:23,31).  Batches
// of 1, 2, 4 and 8 prepares run the same kernels on the same data, so they
// give bit-identical potentials and only their time differs.  On the first
// run of a shape the whole prepare sequence is timed once per candidate, in
// two rounds (forward and reverse order, after one untimed warm-up with the
// rule's batch), and the fastest is kept for the engine and for the process.
int choose_batch(clat_engine* e, const Units& units, const std::vector<int>& slots) {
  if (!(e->xline && e->shared())) return 1;  // batching lives on the line-kernel path only
  if (!e->tune) return batch_rule(e);
  if (e->tuned_nb) return e->tuned_nb;
  const std::vector<int64_t> key = tune_key(e);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) return e->tuned_nb = it->second;
  }
  const int nU = (int)units.size();
  // 1, 2, 4, 8 and the rule's count (up to the units of the run)
  const int rule0 = batch_rule(e);
  std::vector<int> cand;
  for (int c = 1; c <= e->nb_max; c *= 2) {
    if (rule0 > c / 2 && rule0 < c && rule0 / 2 < nU) cand.push_back(rule0);
    if (c == 1 || c / 2 < nU) cand.push_back(c);
  }
  if (cand.size() == 1) return e->tuned_nb = cand[0];
  for (auto& ev : e->tune_ev)
    if (!ev) HIPX(hipEventCreate(&ev));
  const uint32_t kmask = e->kmask;
  e->kmask = 0;  // the trials stay out of the kernel statistics
  std::vector<double> best(cand.size(), 1e30);
  try {
    run_prepares(e, units, slots, batch_rule(e));  // warm-up: code objects loaded, caches filled
    for (int round = 0; round < 2; ++round) {
      const int nc = (int)cand.size();
      HIPX(hipEventRecord(e->tune_ev[0], e->stream));
      for (int i = 0; i < nc; ++i) {
        const int k = round ? nc - 1 - i : i;
        run_prepares(e, units, slots, cand[(size_t)k]);
        HIPX(hipEventRecord(e->tune_ev[i + 1], e->stream));
      }
      HIPX(hipEventSynchronize(e->tune_ev[nc]));
      for (int i = 0; i < nc; ++i) {
        const int k = round ? nc - 1 - i : i;
        float ms = 0.f;
        HIPX(hipEventElapsedTime(&ms, e->tune_ev[i], e->tune_ev[i + 1]));
        best[(size_t)k] = std::min(best[(size_t)k], (double)ms);
      }
    }
  } catch (...) {
    e->kmask = kmask;
    throw;
  }
  e->kmask = kmask;
  size_t pick = 0;
  for (double& t : e->tune_ms) t = 0.0;
  for (size_t i = 0; i < cand.size(); ++i) {
    if (best[i] < best[pick]) pick = i;
    e->tune_ms[cand[i] - 1] = best[i];
  }
  // the rule's count stays unless another is more than 2 % faster (the trials
  // of neighbouring counts often differ by less than their run-to-run noise)
  const int rule = batch_rule(e);
  for (size_t i = 0; i < cand.size(); ++i)
    if (cand[i] == rule && best[i] <= 1.02 * best[pick]) pick = i;
  e->tuned_nb = cand[pick];
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tuned[key] = e->tuned_nb;
  return e->tuned_nb;
}

// what the captured kernel arguments of a device stage depend on
std::vector<int64_t> stage_key(const clat_engine* e, const Units& units, const std::vector<int>& slots, int nbt,
                               const OvlCfg& c) {
  std::vector<int64_t> k = {nbt, (int64_t)units.size(), (int64_t)slots.size(), (int64_t)(uintptr_t)e->rho,
                            (int64_t)(uintptr_t)e->Vat, (int64_t)(uintptr_t)e->part, (int64_t)(uintptr_t)e->ovl,
                            e->n_sites, e->atoms_gen, c.mfma, c.kahan, c.t36, c.sym, c.tl, c.tr, c.n_chunks, c.chunk,
                            (int64_t)c.n_out};
  for (const auto& u : units) {
    k.push_back(u[0]);
    k.push_back(u[1]);
  }
  for (int s : slots) k.push_back(s);
  return k;
}

void drop_graphs(clat_engine* e) {
  for (auto& g : e->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  e->gkey.clear();
}

// Capture what f launches on the engine stream into an executable graph;
// nullptr when the runtime refuses (nothing is then left launched).
template <typename F>
hipGraphExec_t capture(clat_engine* e, F&& f) {
  if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  try {
    f();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(e->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    throw;
  }
  hipGraph_t g = nullptr;
  hipError_t st = hipStreamEndCapture(e->stream, &g);
  hipGraphExec_t x = nullptr;
  if (st == hipSuccess && g) st = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (st != hipSuccess || !x) {
    (void)hipGetLastError();
    if (x) (void)hipGraphExecDestroy(x);
    return nullptr;
  }
  return x;
}

// Prepares, the stage event, overlaps.  On the line-kernel path with kernel
// timing off, the prepares and the overlaps are each captured once into a
// HIP graph and replayed on later runs with the same plan and buffers (two
// launches instead of about five per batch of prepares and two for the
// overlaps); the stage event between them stays outside the graphs, whose
// own event records the runtime cannot time.  A capture the runtime refuses
// falls back to direct launches for good.
void device_stage(clat_engine* e, const Units& units, const std::vector<int>& slots, int nbt, const OvlCfg& c) {
  const int nL = (int)slots.size(), P = e->n_states * (e->n_states + 1) / 2;
  auto prepares = [&] { run_prepares(e, units, slots, nbt); };
  auto overlaps = [&] { launch_overlaps(e, nL, P, c); };
  if (!(e->use_graph && e->kmask == 0 && e->xline && e->shared())) {
    prepares();
    HIPX(hipEventRecord(e->stage_ev[3], e->stream));
    overlaps();
    return;
  }
  const std::vector<int64_t> key = stage_key(e, units, slots, nbt, c);
  if (!(e->gexec[0] && e->gexec[1] && key == e->gkey)) {
    drop_graphs(e);
    e->gexec[0] = capture(e, prepares);
    e->gexec[1] = e->gexec[0] ? capture(e, overlaps) : nullptr;
    if (!e->gexec[1]) {
      drop_graphs(e);
      e->use_graph = false;
      prepares();
      HIPX(hipEventRecord(e->stage_ev[3], e->stream));
      overlaps();
      return;
    }
    e->gkey = key;
    ++e->graph_captures;
  } else {
    ++e->graph_replays;
  }
  HIPX(hipGraphLaunch(e->gexec[0], e->stream));
  HIPX(hipEventRecord(e->stage_ev[3], e->stream));
  HIPX(hipGraphLaunch(e->gexec[1], e->stream));
}

// ===========================================================================
// one integral run, in phases (a group of engines shares the middle one)
// ===========================================================================
struct RunPlan {
  int P = 0, nL = 0, nbt = 1;
  bool stale = false, device = false, table = false;
  std::vector<int> slot_of, slots;
  Units units;
  OvlCfg oc;
};

// Pair densities (queued first: they run while the host plans), the cover,
// the left-product slots, the prepare units, the buffers and the batch.
void run_begin(clat_engine* e, const int32_t* quads, int64_t n, int flags, RunPlan& rp) {
  HIPX(hipEventRecord(e->stage_ev[1], e->stream));
  e->rho_states = -1;  // the pair densities are part of every run's compute stage
  build_products(e);
  HIPX(hipEventRecord(e->stage_ev[2], e->stream));
  rp.table = ensure_table(e);  // a new interaction's table, queued behind the densities
  // the stale-values bookkeeping walks the steps in execution order; without
  // it every value depends on its own step only and id order serves
  rp.stale = (flags & CLAT_FLAG_STALE_VALUES) != 0;
  // the planner's cover fixes which products lead steps and so the prepare
  // order; the steps themselves are built while the prepares run
  host::plan_cover(quads, n, e->n_states, e->plan_in_cover, e->plan_leads);
  rp.P = e->n_states * (e->n_states + 1) / 2;
  // left products in execution order (ascending product index) -> slots
  rp.slot_of.assign((size_t)rp.P, -1);
  for (int p = 0; p < rp.P; ++p)
    if (e->plan_leads[(size_t)p]) {
      rp.slot_of[(size_t)p] = (int)rp.slots.size();
      rp.slots.push_back(p);
    }
  rp.nL = (int)rp.slots.size();
  const size_t vat_need = (size_t)std::max(1, rp.nL) * (size_t)std::max<int64_t>(1, e->n_sites);
  if (e->vat_cap < vat_need) {
    e->dfree(e->Vat, e->vat_cap * sizeof(double2));
    e->Vat = (double2*)e->dalloc(vat_need * sizeof(double2));
    e->vat_cap = vat_need;
  }
  // prepare units: on the line-kernel path two real (diagonal) densities
  // share one complex prepare, their potentials being its real and
  // imaginary parts (the kernel is real)
  {
    std::vector<int> reals;
    const bool pk = e->pack && e->xline && e->shared();
    for (int sl = 0; sl < rp.nL; ++sl) {
      int L = 0;
      while ((L + 1) * (L + 2) / 2 <= rp.slots[(size_t)sl]) ++L;
      const bool diag = rp.slots[(size_t)sl] == L * (L + 1) / 2 + L;
      if (pk && diag) reals.push_back(sl);
      else rp.units.push_back({sl, -1});
    }
    for (size_t i = 0; i < reals.size(); i += 2)
      rp.units.push_back({reals[i], i + 1 < reals.size() ? reals[i + 1] : -1});
  }
  // overlap launch shape; its buffers are allocated here, before any
  // capture of the device stage
  rp.oc = overlap_cfg(e, rp.nL, rp.P);
  ensure_overlap_buffers(e, rp.oc);
  if (e->ov_host_cap < rp.oc.n_out) {  // kept between runs, like the plan buffer
    if (e->ov_host) (void)hipHostFree(e->ov_host);
    e->ov_host = nullptr;
    e->ov_host_cap = 0;
    HIPX(hipHostMalloc((void**)&e->ov_host, rp.oc.n_out * sizeof(double), hipHostMallocDefault));
    e->ov_host_cap = rp.oc.n_out;
  }
  rp.device = rp.nL > 0 && e->n_sites > 0;
  // with no sites every overlap is zero (the buffer holds the last run's)
  if (!rp.device) std::fill(e->ov_host, e->ov_host + rp.oc.n_out, 0.0);
  // prepares per launch: measured on the first run of this shape (the
  // choices give bit-identical potentials)
  else rp.nbt = choose_batch(e, rp.units, rp.slots);
}

// The steps (on the host while the device stage runs), the overlap table to
// the host, the main loop's bookkeeping (coulombo.cpp:186-209), stage times.
void run_finish(clat_engine* e, const int32_t* quads, int64_t n, const RunPlan& rp, double* out) {
  HIPX(hipEventRecord(e->stage_ev[4], e->stream));
  // the copy of the overlap table below waits for the queued prepares and
  // overlaps
  host::make_plan_into(quads, n, e->n_states, rp.stale, e->plan_steps, &e->plan_in_cover);
  const std::vector<host::PlanStep>& steps = e->plan_steps;
  const double* ov = e->ov_host;
  if (rp.device)
    HIPX(hipMemcpyAsync(e->ov_host, e->ovl, rp.oc.n_out * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  HIPX(hipStreamSynchronize(e->stream));
  e->kcollect();
  const int P = rp.P;
  auto value_of = [&](const host::PlanStep& s, int rc, double& vre, double& vim) {
    const int sl = rp.slot_of[(size_t)s.left.product], pr = s.right.product;
    double Av, Bv, Cv, Dv;
    if (rp.oc.sym && sl / OVM_T > pr / OVM_T) {
      // below the diagonal: the mirror entry (slot pr, product sl), its
      // mixed sums exchanged (every product is a slot: slot = product)
      const double* q = ov + ((size_t)pr * P + sl) * 4;
      Av = q[0]; Bv = q[1]; Cv = q[3]; Dv = q[2];
    } else {
      const double* q = ov + ((size_t)sl * P + pr) * 4;
      Av = q[0]; Bv = q[1]; Cv = q[2]; Dv = q[3];
    }
    if (rc) { vre = Av + Bv; vim = Dv - Cv; }
    else { vre = Av - Bv; vim = Cv + Dv; }
  };
  if (!rp.stale) {
    // every value from its own step: host threads over id ranges
    host::parallel_ranges(steps.size(), [&](size_t k0, size_t k1, size_t) {
      for (size_t k = k0; k < k1; ++k) {
        const auto& s = steps[k];
        double vre, vim;
        value_of(s, (s.left.conj != s.right.conj) ? 1 : 0, vre, vim);
        out[2 * (size_t)s.id] = vre;
        out[2 * (size_t)s.id + 1] = s.left.conj ? -vim : vim;
      }
    });
  } else {
    int lastR = -1, lastRc = 0;
    double vre = 0.0, vim = 0.0;
    for (const auto& s : steps) {
      const int rc = (s.left.conj != s.right.conj) ? 1 : 0;
      if (s.right.product != lastR || rc != lastRc) {
        value_of(s, rc, vre, vim);
        lastR = s.right.product;
        lastRc = rc;
      }
      out[2 * (size_t)s.id] = vre;
      out[2 * (size_t)s.id + 1] = s.left.conj ? -vim : vim;
    }
  }
  float ms;
  HIPX(hipEventElapsedTime(&ms, e->stage_ev[1], e->stage_ev[2])); e->stage_ms[1] = ms;
  HIPX(hipEventElapsedTime(&ms, e->stage_ev[rp.table ? 5 : 2], e->stage_ev[3])); e->stage_ms[2] = ms;
  HIPX(hipEventElapsedTime(&ms, e->stage_ev[3], e->stage_ev[4])); e->stage_ms[3] = ms;
  HIPX(hipEventElapsedTime(&ms, e->stage_ev[1], e->stage_ev[4])); e->stage_ms[4] = ms;
}

// The device stages of several engines (systems of any grid sizes) as one
// HIP graph on the first engine's stream: captured once, then launched once
// per group run while every engine's plan and buffers stay the same.  The
// first engine's stream waits for every engine's pair densities first, and
// every engine's stream waits for the graph before its overlap table is
// copied.  With kernel timing on, or when the runtime refuses the capture,
// each engine runs its own device stage on its own stream instead.  A group
// run reports its whole device stage as every member's prepare stage.
void group_stage(clat_engine* const* es, int count, std::vector<RunPlan>& rps) {
  clat_engine* lead = es[0];
  bool graph = lead->use_graph;
  for (int i = 0; i < count; ++i) {
    const clat_engine* e = es[i];
    graph = graph && e->use_graph && e->kmask == 0 && (!rps[(size_t)i].device || (e->xline && e->shared()));
  }
  auto each_direct = [&] {
    for (int i = 0; i < count; ++i) {
      clat_engine* e = es[i];
      const RunPlan& rp = rps[(size_t)i];
      if (rp.device) device_stage(e, rp.units, rp.slots, rp.nbt, rp.oc);
      else HIPX(hipEventRecord(e->stage_ev[3], e->stream));
    }
  };
  if (!graph) return each_direct();
  for (int i = 0; i < count; ++i)
    if (!es[i]->join_ev) HIPX(hipEventCreateWithFlags(&es[i]->join_ev, hipEventDisableTiming));
  // the graph's key: every member and what its captured arguments depend on
  std::vector<int64_t> key;
  for (int i = 0; i < count; ++i) {
    const RunPlan& rp = rps[(size_t)i];
    key.push_back(es[i]->serial);  // not the address: a later engine may reuse it
    key.push_back(rp.device ? 1 : 0);
    if (rp.device) {
      const std::vector<int64_t> k = stage_key(es[i], rp.units, rp.slots, rp.nbt, rp.oc);
      key.push_back((int64_t)k.size());
      key.insert(key.end(), k.begin(), k.end());
    }
  }
  hipStream_t gs = lead->stream;
  for (int i = 1; i < count; ++i) {
    HIPX(hipEventRecord(es[i]->join_ev, es[i]->stream));
    HIPX(hipStreamWaitEvent(gs, es[i]->join_ev, 0));
  }
  if (!(lead->ggroup && key == lead->ggroup_key)) {
    if (lead->ggroup) {
      (void)hipGraphExecDestroy(lead->ggroup);
      lead->ggroup = nullptr;
      lead->ggroup_key.clear();
    }
    auto all = [&] {
      for (int i = 0; i < count; ++i) {
        clat_engine* e = es[i];
        const RunPlan& rp = rps[(size_t)i];
        if (!rp.device) continue;
        // this member's launches go to the lead stream for the capture
        struct StreamSwap {
          clat_engine* e;
          hipStream_t own;
          ~StreamSwap() { e->stream = own; }
        } swap{e, e->stream};
        e->stream = gs;
        run_prepares(e, rp.units, rp.slots, rp.nbt);
        launch_overlaps(e, rp.nL, rp.P, rp.oc);
      }
    };
    lead->ggroup = capture(lead, all);
    if (!lead->ggroup) {
      lead->use_graph = false;
      return each_direct();
    }
    lead->ggroup_key = key;
    ++lead->graph_captures;
  } else {
    ++lead->graph_replays;
  }
  HIPX(hipGraphLaunch(lead->ggroup, gs));
  HIPX(hipEventRecord(lead->join_ev, gs));
  for (int i = 0; i < count; ++i) {
    if (i > 0) HIPX(hipStreamWaitEvent(es[i]->stream, lead->join_ev, 0));
    HIPX(hipEventRecord(es[i]->stage_ev[3], es[i]->stream));
  }
}

// The coarse engine of the two-coset path (created on first use): half the
// engine points per axis, the same precision and schedule, this engine's
// stream and kernel timing, buffers for two coset planes per prepare.  False
// when the coarse grid cannot run every pass on the compile-time kernels,
// or the coupled conv's two LDS tiles would exceed 64 KiB.
bool ensure_sub(clat_engine* e) {
  if (e->sub) return true;
  if (e->sub_failed) return false;
  clat_engine* s = nullptr;
  if (clat_create(e->Xi / 2, e->Yi / 2, e->Zi / 2, e->prec, e->sched, &s) != CLAT_OK) {
    e->sub_failed = true;
    return false;
  }
  const size_t csz = s->Gc32 ? sizeof(float2) : sizeof(double2);
  const size_t conv2_lds = (2 * (size_t)s->Zi * s->gc_c + 2 * (size_t)s->Zi) * csz;
  const bool ok = 2 * s->Xi == e->Xi && 2 * s->Yi == e->Yi && 2 * s->Zi == e->Zi && s->xline && s->conv_tiled &&
                  s->btile && s->atile && s->rprec == e->rprec && s->mix == e->mix && conv2_lds <= 65536 &&
                  2 * e->nb_max <= LINE_NB;
  if (!ok) {
    clat_destroy(s);
    e->sub_failed = true;
    return false;
  }
  try {
    HIPX(hipStreamSynchronize(s->stream));
    hipStream_t own = s->stream;
    s->stream = e->stream;
    s->own_stream = false;
    HIPX(hipStreamDestroy(own));
    s->kparent = e;
    s->pair_conv = true;
    s->tune = 0;
    // plane buffers for every prepare count the parent may batch
    if (s->D != s->A) s->dfree(s->D, s->d_bytes);
    s->D = nullptr;
    s->dfree(s->A, s->a_bytes);
    s->dfree(s->B, s->b_bytes);
    s->nb_max = s->nb_rule = 2 * e->nb_max;
    s->a_bytes = s->nb_max * s->a_set * s->s_a;
    s->b_bytes = s->nb_max * s->b_set * s->s_b;
    s->A = s->dalloc(s->a_bytes);
    s->B = s->dalloc(s->b_bytes);
    if (s->s_d == s->s_a) {
      s->D = s->A;
    } else {
      s->d_bytes = s->nb_max * s->d_set * s->s_d;
      s->D = s->dalloc(s->d_bytes);
    }
    // cross-coset table: a tile per fy in [0, 2 Yc) and column block, rows fz
    // in [0, 2 Zc), over the 2 Xc columns (the half box carries fewer)
    s->q_cap = (size_t)2 * s->Yi * 2 * s->Xi * (2 * s->Zi);
    s->q_elems = (size_t)2 * s->Yi * s->gc_ncb * (2 * s->Zi) * s->gc_c;
    s->Q = s->dalloc(s->q_cap * csz);
    s->Gt64 = (double*)s->dalloc(s->NG() * sizeof(double));
    {
      const int dims[3] = {s->Xi, s->Yi, s->Zi};
      std::vector<double> ph;
      for (int d = 0; d < 3; ++d)
        for (int k = 0; k < 2 * dims[d]; ++k) {
          const double a = M_PI * (double)k / (2.0 * dims[d]);
          ph.push_back(std::cos(a));
          ph.push_back(std::sin(a));
        }
      s->ph64 = (double2*)s->dalloc(ph.size() * sizeof(double));
      HIPX(hipMemcpy(s->ph64, ph.data(), ph.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  } catch (...) {
    clat_destroy(s);
    e->sub_failed = true;
    throw;
  }
  e->sub = s;
  return true;
}

// The coarse engine's column layout for the current atoms: the half box
// (x frequencies below Xc only, in whole strips) or all 2 Xc of them.  The
// buffers and tables were sized for the larger of the two.
void set_half(clat_engine* s, bool half) {
  const int C = s->gc_c;
  s->half = half;
  s->wcols = half ? (s->Xi + C - 1) / C * C : 2 * s->Xi;
  s->fmax = half ? s->Xi : 2 * s->Xi;
  s->gc_w = s->wcols;
  s->gc_ncb = s->wcols / C;
  s->gc_elems = (size_t)(s->Yi + 1) * s->gc_ncb * (s->Zi + 1) * C;
  s->q_elems = (size_t)2 * s->Yi * s->gc_ncb * (2 * s->Zi) * C;
  s->a_set = s->d_set = (size_t)s->Zi * s->wcols * s->Yi;
  s->b_set = (size_t)2 * s->Yi * s->wcols * s->Zi;
  int r1 = 0, r2 = 0;
  s->xhalf = half && s->Xi % 2 == 0 && split2(s->Xi / 2, r1, r2) && col_pair_compiled(r1, r2);
  if (const char* g = std::getenv("CLAT_HALFLINES")) s->xhalf = s->xhalf && std::atoi(g) != 0;
  s->hx_r1 = s->xhalf ? r1 : 0;
  s->hx_r2 = s->xhalf ? r2 : 0;
}

// Sites of the two-coset path: coset 0 (all coordinates even), then coset 1,
// each over the ascending coarse cells (of the coarse grid Xi/2 x Yi/2 x
// Zi/2) that either coset occupies; a coarse cell occupied in one coset only
// is a site without atom in the other (zero coefficients, so zero density;
// its potential is never read).  False when a site lies off x = y = z (mod 2).
struct CosetOrder {
  std::vector<int32_t> ucell;              // coarse cell of each of the n_u positions
  std::vector<int32_t> site_atom, site_cell;  // 2 n_u sites: atom (or -1), engine cell
  std::vector<int32_t> new_of;             // cell-ordered site -> two-coset site
};

bool coset_order(const clat_engine* e, const std::vector<int32_t>& site_atom, const std::vector<int32_t>& site_cell,
                 CosetOrder& co) {
  const int Xi = e->Xi, Yi = e->Yi, Xc = e->Xi / 2, Yc = e->Yi / 2, Zc = e->Zi / 2;
  const int64_t nc = (int64_t)Xc * Yc * Zc;
  std::vector<int32_t> slot((size_t)(2 * nc), -1);  // (coset, coarse cell) -> cell-ordered site
  for (size_t k = 0; k < site_cell.size(); ++k) {
    const int64_t cell = site_cell[k], line = cell / Xi;
    const int ix = (int)(cell - line * Xi), iy = (int)(line % Yi), iz = (int)(line / Yi);
    const int p = ix & 1;
    if ((iy & 1) != p || (iz & 1) != p) return false;
    slot[(size_t)(p * nc + ((int64_t)(iz >> 1) * Yc + (iy >> 1)) * Xc + (ix >> 1))] = (int32_t)k;
  }
  for (int64_t u = 0; u < nc; ++u)
    if (slot[(size_t)u] >= 0 || slot[(size_t)(nc + u)] >= 0) co.ucell.push_back((int32_t)u);
  const size_t nu = co.ucell.size();
  co.site_atom.assign(2 * nu, -1);
  co.site_cell.assign(2 * nu, 0);
  co.new_of.assign(site_cell.size(), -1);
  for (int p = 0; p < 2; ++p)
    for (size_t r = 0; r < nu; ++r) {
      const int64_t u = co.ucell[r];
      const int mx = (int)(u % Xc), my = (int)((u / Xc) % Yc), mz = (int)(u / ((int64_t)Xc * Yc));
      const size_t j = (size_t)p * nu + r;
      const int32_t k = slot[(size_t)(p * nc + u)];
      co.site_atom[j] = k >= 0 ? site_atom[(size_t)k] : -1;
      co.site_cell[j] = (int32_t)(((int64_t)(2 * mz + p) * Yi + (2 * my + p)) * Xi + (2 * mx + p));
      if (k >= 0) co.new_of[(size_t)k] = (int32_t)j;
    }
  return true;
}

void free_atoms(clat_engine* e) {
  e->dfree(e->d_site_atom, (size_t)e->n_sites * 4);
  e->dfree(e->d_site_cell, (size_t)e->n_sites * 4);
  e->dfree(e->d_site_ptr, ((size_t)e->Yi * e->Zi + 1) * 4);
  e->dfree(e->d_atom_site, (size_t)e->n_atoms * 4);
  e->dfree(e->d_lines, std::max<size_t>(1, (size_t)e->n_live) * 4);
  e->dfree(e->d_live_ptr, ((size_t)e->n_live + 1) * 4);
  e->dfree(e->d_site_ccell, std::max<size_t>(1, (size_t)e->n_sites) * 4);
  e->dfree(e->d_linemask, e->linemask_words * 4);
  e->linemask_words = 0;
  e->n_live = 0;
  e->dfree(e->coeffs, std::max<size_t>(1, (size_t)e->cap_states * e->n_sites * e->n_orb * 16));
  e->dfree(e->rho, e->rho_bytes);
  e->rho_bytes = 0;
  e->rho_states = -1;
  e->n_states = e->cap_states = 0;
  e->n_sites = 0;
  e->n_atoms = 0;
  e->coset = false;
  e->n_u = 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* clat_last_error(void) { return g_err.c_str(); }

const char* clat_version(void) { return "clat 1.1 (gfx950)"; }

int clat_round_up(int n, int c, const int* primes, int n_primes, int* out) {
  return guarded([&] {
    if (!primes || n_primes <= 0 || !out) throw host::Error(CLAT_EINVAL, "invalid arguments");
    *out = host::round_up(n, c, std::vector<int>(primes, primes + n_primes));
  });
}

int clat_engine_length(int n, int* out) {
  return guarded([&] {
    if (n <= 0 || !out) throw host::Error(CLAT_EINVAL, "invalid arguments");
    *out = engine_len(n);
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
  static std::atomic<int64_t> next_serial{1};
  e->serial = next_serial++;
  int rc = guarded([&] {
    if (X <= 0 || Y <= 0 || Z <= 0) throw host::Error(CLAT_EINVAL, "dimensions must be positive");
    if ((double)X * Y * Z >= 2147483647.0) throw host::Error(CLAT_EINVAL, "dimensions are too large");
    if (precision != CLAT_PREC_FP64 && precision != CLAT_PREC_FP32 && precision != CLAT_PREC_FP32X)
      throw host::Error(CLAT_EINVAL, "unknown precision");
    if (schedule != CLAT_SCHED_SHARED && schedule != CLAT_SCHED_ROUNDS)
      throw host::Error(CLAT_EINVAL, "unknown schedule");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
      throw host::Error(CLAT_EDEVICE, "no HIP device available");
    HIPX(hipGetDevice(&e->device));
    HIPX(hipDeviceGetAttribute(&e->n_cu, hipDeviceAttributeMultiprocessorCount, e->device));
    e->n_cu = std::max(1, e->n_cu);
    e->X = X; e->Y = Y; e->Z = Z;
    e->Xi = engine_len(X); e->Yi = engine_len(Y); e->Zi = engine_len(Z);
    if ((double)e->Xi * e->Yi * e->Zi * 4.0 >= 4294967295.0)
      throw host::Error(CLAT_EUNSUP, "grid too large for one engine");
    e->prec = precision;
    e->sched = schedule;
    int r1, r2;
    e->fast = split2(e->Xi, r1, r2) && split2(e->Yi, r1, r2) && split2(e->Zi, r1, r2) &&
              e->Xi <= 400 && e->Yi <= 400 && e->Zi <= 400;
    if (const char* g = std::getenv("CLAT_GENERIC"))
      if (std::atoi(g)) e->fast = false;
    if (const char* g = std::getenv("CLAT_COLK")) e->colk = std::atoi(g) != 0;
    // the mixed mode needs the two-level kernels; otherwise it runs in FP64
    e->rprec = (precision == CLAT_PREC_FP32X && !e->fast) ? CLAT_PREC_FP64 : precision;
    if (const char* lb = std::getenv("CLAT_LDS_BYTES")) {
      const long v = std::atol(lb);
      if (v >= 4096 && v <= 163840) e->lds_budget = (size_t)v;
    }
    const int nsub = schedule == CLAT_SCHED_SHARED ? 2 : 1;
    const size_t dbl = 16, flt = 8;
    switch (e->rprec) {
      case CLAT_PREC_FP64: e->s_in = e->s_a = e->s_b = e->s_d = e->s_v = dbl; break;
      case CLAT_PREC_FP32: e->s_in = e->s_a = e->s_b = e->s_d = e->s_v = flt; break;
      default: e->s_in = dbl; e->s_a = flt; e->s_b = flt; e->s_d = flt; e->s_v = dbl; break;
    }
    if (const char* m = std::getenv("CLAT_MIX")) e->mix = std::max(0, std::min(3, std::atoi(m)));
    if (e->rprec == CLAT_PREC_FP32X && e->mix >= 1) e->s_d = dbl;
    const size_t tc = (e->rprec == CLAT_PREC_FP64 || (e->rprec == CLAT_PREC_FP32X && e->mix >= 3)) ? dbl : flt;
    const size_t tcz = (e->rprec == CLAT_PREC_FP64 || (e->rprec == CLAT_PREC_FP32X && e->mix >= 2)) ? dbl : flt;
    const size_t tcy = (e->rprec == CLAT_PREC_FP64 || (e->rprec == CLAT_PREC_FP32X && e->mix >= 1)) ? dbl : flt;
    const size_t tcx = e->rprec == CLAT_PREC_FP32 ? flt : dbl;   // arithmetic of the inverse x pass
    const int dims[3] = {e->Xi, e->Yi, e->Zi};
    AxisPlan* lp[3] = {&e->px, &e->py, &e->pz};
    AxisPlan* dp[3] = {&e->dx, &e->dy, &e->dz};
    e->tile_budget = e->rprec == CLAT_PREC_FP32 ? 32 * NT : 16 * NT;
    for (int d = 0; d < 3; ++d) {
      if (!make_plan(dims[d], *lp[d]) || !make_plan(2 * dims[d], *dp[d]))
        throw host::Error(CLAT_EUNSUP, "axis length has no radix plan");
      if (!e->fast && nsub * dims[d] > e->tile_budget)
        throw host::Error(CLAT_EUNSUP, "axis length exceeds the single-tile FFT limit");
      if (2 * dims[d] > 16 * NT) throw host::Error(CLAT_EUNSUP, "axis length exceeds the DCT-I tile limit");
    }
    if (e->fast) {
      const int W = nsub * e->Xi;
      e->cxf = cfg_line(e->Xi, nsub, tc, false, e->lds_budget);
      e->cxf_s = cfg_line(e->Xi, nsub, tc, true, e->lds_budget);
      e->cxi = cfg_line(e->Xi, nsub, tcx, false, e->lds_budget);
      e->cxi_s = cfg_line(e->Xi, nsub, tcx, true, e->lds_budget);
      e->cyf = cfg_col(e->Yi, nsub, tc, W, e->lds_budget);
      // the complex64 conv pass runs best with half-size tiles (more blocks in flight)
      size_t zb = tcz == flt ? e->lds_budget / 2 : e->lds_budget;
      if (const char* zl = std::getenv("CLAT_LDS_CONV")) {
        const long v = std::atol(zl);
        if (v >= 4096 && v <= 65536) zb = (size_t)v;
      }
      e->czc = cfg_col(e->Zi, nsub, tcz, W, zb);
      e->cyi = cfg_col(e->Yi, nsub, tcy, W, e->lds_budget);
      const PassCfg* all[] = {&e->cxf, &e->cxf_s, &e->cxi, &e->cxi_s, &e->cyf, &e->czc, &e->cyi};
      for (const PassCfg* p : all)
        if (p->lds > 65536) throw host::Error(CLAT_EUNSUP, "pass tile exceeds the LDS budget");
      if (e->cxf_s.C * e->cxf_s.R1 > e->cxf_s.NT || e->cxi_s.C * e->cxi_s.R1 > e->cxi_s.NT)
        throw host::Error(CLAT_EUNSUP, "site pass shape");
    }
    if (e->fast && e->shared() && e->colk) {
      const long long W = 2LL * e->Xi;
      e->xline = col_pair_compiled(e->cxf_s.R1, e->cxf_s.R2) && col_pair_compiled(e->cyf.R1, e->cyf.R2) &&
                 e->cyf.R1 == e->cyi.R1 && e->cyf.R2 == e->cyi.R2 &&
                 col_fits(e->Yi, 2, W, e->s_a, (long long)e->Zi * W, e->s_b) &&
                 col_fits(e->Yi, 2, (long long)e->Zi * W, e->s_b, W, e->s_d);
    }
    if (const char* g = std::getenv("CLAT_XLINE")) e->xline = e->xline && std::atoi(g) != 0;
    HIPX(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    for (int i = 0; i < 6; ++i) HIPX(hipEventCreate(&e->stage_ev[i]));
    for (int d = 0; d < 3; ++d) {
      const int L = dims[d];
      std::vector<double> tw = unit_roots(L, L), hf = unit_roots(L, 2 * L), dt = unit_roots(2 * L, 2 * L);
      std::vector<float> a(tw.begin(), tw.end()), b(hf.begin(), hf.end());
      e->tw32[d] = e->dalloc(a.size() * sizeof(float));
      e->h32[d] = e->dalloc(b.size() * sizeof(float));
      e->tw64[d] = e->dalloc(tw.size() * sizeof(double));
      e->h64[d] = e->dalloc(hf.size() * sizeof(double));
      e->dtw[d] = (double2*)e->dalloc(dt.size() * sizeof(double));
      std::vector<float> dt32(dt.begin(), dt.end());
      e->dtw32[d] = (float2*)e->dalloc(dt32.size() * sizeof(float));
      HIPX(hipMemcpy(e->dtw32[d], dt32.data(), dt32.size() * sizeof(float), hipMemcpyHostToDevice));
      HIPX(hipMemcpy(e->tw32[d], a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice));
      HIPX(hipMemcpy(e->h32[d], b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
      HIPX(hipMemcpy(e->tw64[d], tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
      HIPX(hipMemcpy(e->h64[d], hf.data(), hf.size() * sizeof(double), hipMemcpyHostToDevice));
      HIPX(hipMemcpy(e->dtw[d], dt.data(), dt.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    e->G64 = (double*)e->dalloc(e->NG() * sizeof(double));
    if (e->rprec == CLAT_PREC_FP32) e->G32 = (float*)e->dalloc(e->NG() * sizeof(float));
    if (e->fast) {
      // conv pass on the column kernel when its radix pair is compiled and its rows fit 32-bit offsets
      const long long zrow = e->shared() ? 2LL * e->Xi : (long long)e->Xi * e->Yi;
      const size_t zb = e->shared() ? e->s_b : e->s_a;
      e->conv_tiled = e->colk && col_pair_compiled(e->czc.R1, e->czc.R2) &&
                      col_fits(e->Zi, nsub, zrow, zb, zrow, zb);
      if (e->conv_tiled) {
        e->gc_c = col_c_bytes(tcz, e->Zi);
        e->gc_npar = e->shared() ? 1 : 2;
        e->gc_w = e->shared() ? 2 * e->Xi : e->Xi;
        e->gc_ncb = (e->gc_w + e->gc_c - 1) / e->gc_c;
        e->gc_elems = (size_t)(e->Yi + 1) * e->gc_npar * e->gc_ncb * (e->Zi + 1) * e->gc_c;
      } else {
        e->gc_elems = (size_t)2 * e->Xi * (e->Yi + 1) * (e->Zi + 1);
      }
      if (tcz == flt) e->Gc32 = (float*)e->dalloc(e->gc_elems * sizeof(float));
      else e->Gc64 = (double*)e->dalloc(e->gc_elems * sizeof(double));
      // B as strips of C columns, each Z rows deep ([fy][strip][z][C]): the conv
      // pass then streams one contiguous 16 KB tile per block instead of
      // 128-byte row pieces 2X apart (a copy with the two access maps: 5.96 vs
      // 4.98 TB/s); the y passes keep their access pattern.  Needs all three
      // passes on column kernels with one column count C.
      const long long W = 2LL * e->Xi;
      e->btile = e->shared() && e->conv_tiled && tc == tcz && tcy == tcz && W % e->gc_c == 0 &&
                 col_c_bytes(tc, e->Yi) == e->gc_c &&
                 col_pair_compiled(e->cyf.R1, e->cyf.R2) && col_pair_compiled(e->cyi.R1, e->cyi.R2) &&
                 col_fits(e->Yi, 2, W, e->s_a, (long long)e->Zi * W, e->s_b) &&
                 col_fits(e->Yi, 2, (long long)e->Zi * W, e->s_b, W, e->s_d);
      if (const char* g = std::getenv("CLAT_BTILE")) e->btile = e->btile && std::atoi(g) != 0;
      // A and D likewise ([z][fx / C][y][fx % C]) where the x passes run on the
      // line kernels: the y forward reads and the y inverse writes contiguous
      // strips, the line kernels write and read 128-byte pieces of them (a
      // copy with the y-forward access map: 4.66 against 3.57 TB/s; with the
      // x-forward write map: 6.36 against 4.56)
      e->atile = e->btile && e->xline && col_fits(e->Yi, 2, e->gc_c, e->s_a, (long long)e->Zi * W, e->s_b) &&
                 col_fits(e->Yi, 2, (long long)e->Zi * W, e->s_b, e->gc_c, e->s_d);
      if (const char* g = std::getenv("CLAT_ATILE")) e->atile = e->atile && std::atoi(g) != 0;
    }
    e->wcols = e->fmax = 2 * e->Xi;
    const size_t N = e->N();
    e->f_bytes = N * e->s_in;
    e->v_bytes = N * e->s_v;
    e->a_set = (e->shared() ? 2 : 1) * N;
    e->b_set = e->shared() ? 4 * N : 0;
    e->d_set = e->a_set;
    // the atom path batches up to four prepares per launch (all five passes on
    // the compile-time kernels) while the buffer sets stay under 16 GB, and
    // eight when eight B planes take at most 1.1 GB (128^3: 57 against 60 ms
    // per batched-form step; 256x256x128 with eight: 10.4 against 10.2 ms)
    const size_t set_bytes = e->a_set * e->s_a + e->b_set * e->s_b + (e->s_d == e->s_a ? 0 : e->d_set * e->s_d);
    // buffers for up to eight prepares per launch while the sets stay under
    // 16 GB; which count runs is measured on the first run of the engine's
    // shape (1, 2, 4 or 8 give bit-identical potentials).  Unmeasured, the
    // rule is four, or eight when eight B planes take at most 1.1 GB (128^3:
    // 57 against 60 ms per batched-form step; 256x256x128: 10.2 against 10.4)
    e->nb_max = 1;
    if (e->xline && e->conv_tiled) {
      while (e->nb_max < 8 && (size_t)(2 * e->nb_max) * set_bytes <= ((size_t)16 << 30)) e->nb_max *= 2;
      e->nb_rule = std::min(e->nb_max, 8 * e->b_set * e->s_b <= (size_t)1100 << 20 ? 8 : 4);
    }
    static_assert(LINE_NB >= 8, "the pointer arrays of a launch hold eight prepares");
    if (const char* g = std::getenv("CLAT_BATCH")) {
      e->nb_max = e->nb_rule = std::max(1, std::min(e->nb_max, std::atoi(g)));
      e->tune = 0;
    }
    if (const char* g = std::getenv("CLAT_TUNE")) e->tune = e->tune && std::atoi(g) != 0;
    if (const char* g = std::getenv("CLAT_GRAPH")) e->use_graph = std::atoi(g) != 0;
    if (const char* g = std::getenv("CLAT_OVL_SYM")) e->ovl_sym = std::atoi(g) != 0;
    if (const char* g = std::getenv("CLAT_PACK")) e->pack = std::atoi(g) != 0;
    if (const char* g = std::getenv("CLAT_OVL"))
      e->ovl_mode = std::strcmp(g, "valu") == 0 ? 1 : std::strcmp(g, "mfma") == 0 ? 2
                    : std::strcmp(g, "valu48") == 0 ? 3 : std::strcmp(g, "kahan") == 0 ? 4 : 0;
    e->a_bytes = e->nb_max * e->a_set * e->s_a;
    e->A = e->dalloc(e->a_bytes);
    if (e->shared()) {
      e->b_bytes = e->nb_max * e->b_set * e->s_b;
      e->B = e->dalloc(e->b_bytes);
    }
    if (e->s_d == e->s_a) {
      e->D = e->A;  // the y-inverse output reuses the x-forward buffer
    } else {
      e->d_bytes = e->nb_max * e->d_set * e->s_d;
      e->D = e->dalloc(e->d_bytes);
    }
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
  if (e->sub) {
    e->sub->stream = e->stream;  // (a group capture may have left it on another engine's stream)
    (void)clat_destroy(e->sub);
    e->sub = nullptr;
  }
  void* ptrs[] = {e->tw32[0], e->tw32[1], e->tw32[2], e->h32[0], e->h32[1], e->h32[2], e->tw64[0], e->tw64[1],
                  e->tw64[2], e->h64[0], e->h64[1], e->h64[2], e->dtw[0], e->dtw[1], e->dtw[2], e->dtw32[0], e->dtw32[1], e->dtw32[2], e->G64,
                  e->G32, e->Gc64, e->Gc32, e->F, e->V, e->A, e->B, (e->D != e->A) ? e->D : nullptr, e->Fin, e->dpart,
                  e->d_site_atom, e->d_site_cell, e->d_site_ptr, e->d_atom_site, e->d_lines, e->d_live_ptr,
                  e->d_site_ccell, e->d_linemask, e->coeffs, e->rho, e->Vat,
                  e->part, e->ovl, e->Q, e->Gt64, e->ph64};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto ev : e->stage_ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto ev : e->kev) (void)hipEventDestroy(ev);
  for (auto ev : e->tune_ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto g : e->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  if (e->ggroup) (void)hipGraphExecDestroy(e->ggroup);
  if (e->join_ev) (void)hipEventDestroy(e->join_ev);
  if (e->ov_host) (void)hipHostFree(e->ov_host);
  if (e->stream && e->own_stream) (void)hipStreamDestroy(e->stream);
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
    e->table_deferred = false;  // this table replaces a set interaction not built yet
    const size_t nu = (size_t)(e->X + 1) * (e->Y + 1) * (e->Z + 1);
    double* tmp = (double*)e->dalloc(nu * sizeof(double));
    hipError_t ce = hipMemcpyAsync(tmp, G, nu * sizeof(double), hipMemcpyHostToDevice, e->stream);
    if (ce == hipSuccess) {
      GArgs g = base_gargs(e);
      hipLaunchKernelGGL(k_gpad, dim3(grid1d(e->NG())), dim3(256), 0, e->stream, tmp, e->G64, g);
      ce = hipGetLastError();
    }
    hipError_t se = hipStreamSynchronize(e->stream);
    e->dfree(tmp, nu * sizeof(double));
    HIPX(ce);
    HIPX(se);
    e->g64_real = true;
    finish_interaction(e);
    HIPX(hipStreamSynchronize(e->stream));
    e->kcollect();
  });
}

int clat_set_input(clat_engine* e, const double* F) {
  if (!e || !F) return set_err(CLAT_EINVAL, "invalid arguments");
  return guarded([&] {
    ensure_dense(e);
    if (!e->Fin) e->Fin = (double2*)e->dalloc(e->NU() * sizeof(double2));
    HIPX(hipMemcpyAsync(e->Fin, F, e->NU() * sizeof(double2), hipMemcpyHostToDevice, e->stream));
    zero_grid(e, e->F, e->N(), e->s_in);
    if (e->s_in == 8)
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
    prepare_dense(e);
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
    if (e->s_v == 8)
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
    if (e->s_v == 8)
      hipLaunchKernelGGL(k_grid_out<float>, dim3(grid1d(e->NU())), dim3(256), 0, e->stream, (const float2*)e->V,
                         tmp, e->X, e->Y, e->Z, e->Xi, e->Yi);
    else
      hipLaunchKernelGGL(k_grid_out<double>, dim3(grid1d(e->NU())), dim3(256), 0, e->stream,
                         (const double2*)e->V, tmp, e->X, e->Y, e->Z, e->Xi, e->Yi);
    hipError_t le = hipGetLastError();
    hipError_t ce = hipMemcpyAsync(V, tmp, e->NU() * sizeof(double2), hipMemcpyDeviceToHost, e->stream);
    hipError_t se = hipStreamSynchronize(e->stream);
    e->dfree(tmp, e->NU() * sizeof(double2));
    HIPX(le);
    HIPX(ce);
    HIPX(se);
  });
}

int clat_set_atoms(clat_engine* e, const int32_t* cell, int64_t n_atoms, int n_orbitals) {
  if (!e || (!cell && n_atoms > 0) || n_atoms < 0) return set_err(CLAT_EINVAL, "invalid arguments");
  if (n_orbitals <= 0) return set_err(CLAT_EINVAL, "invalid value for orbitals");
  if (n_atoms >= 2147483647) return set_err(CLAT_EUNSUP, "too many atoms");
  return guarded([&] {
    const int64_t NU = (int64_t)e->NU();
    // effective sites: the last atom of every occupied cell, ordered by cell
    std::vector<int32_t> last((size_t)NU, -1);
    for (int64_t a = 0; a < n_atoms; ++a) {
      if (cell[a] < 0 || cell[a] >= NU) throw host::Error(CLAT_EINVAL, "atom cell index outside the grid");
      last[(size_t)cell[a]] = (int32_t)a;
    }
    const int64_t nlines = (int64_t)e->Yi * e->Zi;
    std::vector<int32_t> site_atom, site_cell, site_ptr((size_t)nlines + 1, 0), atom_site((size_t)n_atoms);
    for (int64_t c = 0; c < NU; ++c) {
      if (last[(size_t)c] < 0) continue;
      const int ix = (int)(c % e->X), iy = (int)((c / e->X) % e->Y), iz = (int)(c / ((int64_t)e->X * e->Y));
      const int64_t line = (int64_t)iz * e->Yi + iy;
      site_ptr[(size_t)line + 1] += 1;
      site_atom.push_back(last[(size_t)c]);
      site_cell.push_back((int32_t)(line * e->Xi + ix));
      last[(size_t)c] = (int32_t)(site_atom.size() - 1);  // cell -> site from here on
    }
    for (int64_t l = 0; l < nlines; ++l) site_ptr[(size_t)l + 1] += site_ptr[(size_t)l];
    for (int64_t a = 0; a < n_atoms; ++a) atom_site[(size_t)a] = last[(size_t)cell[a]];
    // two-coset path: every site on x = y = z (mod 2) of an even grid, on the
    // atom path of the shared schedule with every pass on compile-time kernels
    bool cos = !e->pair_conv && e->shared() && e->xline && e->conv_tiled && e->btile && e->atile &&
               e->Xi % 2 == 0 && e->Yi % 2 == 0 && e->Zi % 2 == 0 && !site_atom.empty();
    if (const char* g = std::getenv("CLAT_COSETS")) cos = cos && std::atoi(g) != 0;
    CosetOrder co;
    if (cos) cos = coset_order(e, site_atom, site_cell, co);
    if (cos) cos = ensure_sub(e);
    // half box: every coarse position on mx + my + mz even (both cosets on
    // the coarse grid's fcc sublattice, as zincblende's are)
    bool half = cos;
    if (const char* g = std::getenv("CLAT_HALFBOX")) half = half && std::atoi(g) != 0;
    if (half) {
      const int64_t Xc = e->Xi / 2, Yc = e->Yi / 2;
      for (int32_t u : co.ucell)
        if (((u % Xc) + ((u / Xc) % Yc) + (u / (Xc * Yc))) & 1) {
          half = false;
          break;
        }
    }
    // occupied x-lines: compact list, CSR over it, compact cells, bit mask
    std::vector<int32_t> lines, live_ptr(1, 0), ccell(site_cell.size());
    // y-pass mask: one word per (z, n1) of the y radix split, bit n2 for y = n1 + R1 n2
    const int yr1 = e->cyf.R1, yr2 = e->cyf.R2;
    const bool ymask = e->xline && yr2 <= 32;
    std::vector<uint32_t> mask(ymask ? (size_t)e->Zi * yr1 : 1, 0u);
    // (on half-length lines the kernels take ccell >> 1 = rank Xi / 2 + j)
    for (int64_t l = 0; l < nlines; ++l) {
      const int32_t b = site_ptr[(size_t)l], n = site_ptr[(size_t)l + 1] - b;
      if (n == 0) continue;
      const int64_t rank = (int64_t)lines.size();
      const int par = (int)((l % e->Yi + l / e->Yi) & 1);
      for (int32_t k = b; k < b + n; ++k) {
        const int x = (int)(site_cell[(size_t)k] - l * e->Xi);
        if (e->xhalf && (x & 1) != par) throw host::Error(CLAT_EINVAL, "half-length lines: a site off x = y + z (mod 2)");
        ccell[(size_t)k] = (int32_t)(rank * e->Xi + x);
      }
      lines.push_back((int32_t)l);
      live_ptr.push_back(b + n);
      if (ymask) {
        const int64_t z = l / e->Yi, y = l % e->Yi;
        mask[(size_t)(z * yr1 + y % yr1)] |= 1u << (y / yr1);
      }
    }
    free_atoms(e);  // also drops the loaded states (stored in the old site order)
    e->atoms_gen += 1;  // a new geometry: new measured batch, new graph
    e->tuned_nb = 0;
    e->n_atoms = n_atoms;
    e->n_orb = n_orbitals;
    if (cos) {
      // the coarse engine takes the n_u positions as its atoms (one per
      // coarse cell, so its sites are the positions in order); this engine
      // keeps the two-coset site order and no fine line structures
      set_half(e->sub, half);
      const int rc = clat_set_atoms(e->sub, co.ucell.data(), (int64_t)co.ucell.size(), n_orbitals);
      if (rc != CLAT_OK) throw host::Error(rc, std::string("coarse engine: ") + clat_last_error());
      e->coset = true;
      e->n_u = (int64_t)co.ucell.size();
      for (int64_t a = 0; a < n_atoms; ++a) atom_site[(size_t)a] = co.new_of[(size_t)atom_site[(size_t)a]];
      site_atom.swap(co.site_atom);
      site_cell.swap(co.site_cell);
      std::fill(site_ptr.begin(), site_ptr.end(), 0);
      lines.clear();
      live_ptr.assign(1, 0);
      ccell.assign(site_atom.size(), 0);
      std::fill(mask.begin(), mask.end(), 0u);
      if (e->initialized && !e->table_deferred) build_coset_tables(e);
    } else if (e->initialized && !e->table_deferred) {
      ensure_fine_tables(e);
    }
    e->n_sites = (int64_t)site_atom.size();
    e->d_site_atom = (int32_t*)e->dalloc((size_t)e->n_sites * 4);
    e->d_site_cell = (int32_t*)e->dalloc((size_t)e->n_sites * 4);
    e->d_site_ptr = (int32_t*)e->dalloc(((size_t)nlines + 1) * 4);
    e->d_atom_site = (int32_t*)e->dalloc((size_t)n_atoms * 4);
    HIPX(hipMemcpy(e->d_site_atom, site_atom.data(), (size_t)e->n_sites * 4, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(e->d_site_cell, site_cell.data(), (size_t)e->n_sites * 4, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(e->d_site_ptr, site_ptr.data(), ((size_t)nlines + 1) * 4, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(e->d_atom_site, atom_site.data(), (size_t)n_atoms * 4, hipMemcpyHostToDevice));
    e->n_live = (int64_t)lines.size();
    e->d_lines = (int32_t*)e->dalloc(std::max<size_t>(1, lines.size()) * 4);
    e->d_live_ptr = (int32_t*)e->dalloc(live_ptr.size() * 4);
    e->d_site_ccell = (int32_t*)e->dalloc(std::max<size_t>(1, ccell.size()) * 4);
    e->d_linemask = (uint32_t*)e->dalloc(mask.size() * 4);
    e->linemask_words = mask.size();
    if (!lines.empty()) HIPX(hipMemcpy(e->d_lines, lines.data(), lines.size() * 4, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(e->d_live_ptr, live_ptr.data(), live_ptr.size() * 4, hipMemcpyHostToDevice));
    if (!ccell.empty()) HIPX(hipMemcpy(e->d_site_ccell, ccell.data(), ccell.size() * 4, hipMemcpyHostToDevice));
    HIPX(hipMemcpy(e->d_linemask, mask.data(), mask.size() * 4, hipMemcpyHostToDevice));
  });
}

int clat_add_state(clat_engine* e, const double* coeffs) {
  if (!e || !coeffs) return set_err(CLAT_EINVAL, "invalid arguments");
  if (e->n_orb <= 0) return set_err(CLAT_ESTATE, "atoms have not been set");
  return guarded([&] {
    // stored in site order (the effective atom of every occupied cell)
    const size_t per = (size_t)e->n_sites * e->n_orb * sizeof(double2);
    const size_t in_bytes = (size_t)e->n_atoms * e->n_orb * sizeof(double2);
    if (e->n_states == e->cap_states) {
      const int ncap = std::max(4, 2 * e->cap_states);
      double2* nc = (double2*)e->dalloc(std::max<size_t>(1, per * ncap));
      if (e->n_states) HIPX(hipMemcpy(nc, e->coeffs, per * e->n_states, hipMemcpyDeviceToDevice));
      e->dfree(e->coeffs, std::max<size_t>(1, per * e->cap_states));
      e->coeffs = nc;
      e->cap_states = ncap;
    }
    double2* stage = (double2*)e->dalloc(std::max<size_t>(1, in_bytes));
    hipError_t ce = hipMemcpyAsync(stage, coeffs, in_bytes, hipMemcpyHostToDevice, e->stream);
    if (ce == hipSuccess && e->n_sites > 0)
      hipLaunchKernelGGL(k_coeffs_to_sites, dim3(grid1d((size_t)e->n_sites * e->n_orb)), dim3(256), 0, e->stream,
                         stage, e->d_site_atom, e->n_sites, e->n_orb, e->coeffs + (size_t)e->n_states * e->n_sites * e->n_orb);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(e->stream);
    e->dfree(stage, std::max<size_t>(1, in_bytes));
    HIPX(ce);
    HIPX(le);
    HIPX(se);
    e->n_states += 1;
    e->rho_states = -1;
  });
}

int clat_clear_states(clat_engine* e) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  e->n_states = 0;
  e->rho_states = -1;
  return CLAT_OK;
}

int clat_compute_integrals(clat_engine* e, const int32_t* quads, int64_t n, int flags, double* out) {
  if (!e || (n > 0 && (!quads || !out)) || n < 0) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!e->initialized) return set_err(CLAT_ESTATE, "calculator is not initialized");
  if (e->n_states <= 0) return set_err(CLAT_ESTATE, "no states loaded");
  return guarded([&] {
    RunPlan rp;
    run_begin(e, quads, n, flags, rp);
    if (rp.device) device_stage(e, rp.units, rp.slots, rp.nbt, rp.oc);
    else HIPX(hipEventRecord(e->stage_ev[3], e->stream));
    run_finish(e, quads, n, rp, out);
  });
}

int clat_compute_integrals_group(clat_engine* const* engines, int count, const int32_t* const* quads,
                                 const int64_t* n, int flags, double* const* out) {
  if (!engines || count <= 0 || !quads || !n || !out) return set_err(CLAT_EINVAL, "invalid arguments");
  for (int i = 0; i < count; ++i) {
    clat_engine* e = engines[i];
    if (!e || (n[i] > 0 && (!quads[i] || !out[i])) || n[i] < 0) return set_err(CLAT_EINVAL, "invalid arguments");
    for (int j = 0; j < i; ++j)
      if (engines[j] == e) return set_err(CLAT_EINVAL, "an engine appears twice in the group");
    if (!e->initialized) return set_err(CLAT_ESTATE, "calculator is not initialized");
    if (e->n_states <= 0) return set_err(CLAT_ESTATE, "no states loaded");
    if (e->device != engines[0]->device) return set_err(CLAT_EINVAL, "the engines of a group share one device");
  }
  return guarded([&] {
    std::vector<RunPlan> rps((size_t)count);
    for (int i = 0; i < count; ++i) run_begin(engines[i], quads[i], n[i], flags, rps[(size_t)i]);
    group_stage(engines, count, rps);
    for (int i = 0; i < count; ++i) run_finish(engines[i], quads[i], n[i], rps[(size_t)i], out[i]);
  });
}

int clat_compute_potential_at_atoms(clat_engine* e, int state, double* v) {
  if (!e || !v) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!e->initialized) return set_err(CLAT_ESTATE, "calculator is not initialized");
  if (state < 1 || state > e->n_states) return set_err(CLAT_EINVAL, "state index out of range");
  return guarded([&] {
    // self density, V at the sites and V at the atoms: 2 n_sites + n_atoms
    // elements, freed at return
    const size_t ns = (size_t)e->n_sites, na = (size_t)e->n_atoms;
    const size_t bytes = std::max<size_t>(1, 2 * ns + na) * sizeof(double2);
    double2* buf = (double2*)e->dalloc(bytes);
    try {
      potential_run(e, state - 1, buf, buf + ns, buf + 2 * ns, v);
    } catch (...) {
      (void)hipStreamSynchronize(e->stream);
      e->dfree(buf, bytes);
      throw;
    }
    e->dfree(buf, bytes);
    e->kcollect();
  });
}

int clat_compute_potential_at_atoms_noalloc(clat_engine* e, int state, double* v) {
  if (!e || !v) return set_err(CLAT_EINVAL, "invalid arguments");
  if (!e->initialized) return set_err(CLAT_ESTATE, "calculator is not initialized");
  if (state < 1 || state > e->n_states) return set_err(CLAT_EINVAL, "state index out of range");
  if (!e->fast)
    return set_err(CLAT_EUNSUP, "the allocation-free potential needs the two-level pass kernels; this grid "
                                "runs on the general kernels, whose dense grids are allocated on first use");
  const double need = 2.0 * (double)e->n_sites + (double)e->n_atoms;
  if (need > (double)NOALLOC_ELEMS)
    return set_err(CLAT_EUNSUP, "the allocation-free potential needs 2 x %lld sites + %lld atoms = %.0f "
                                "scratch elements; its maximum is %zu (64 MiB)",
                   (long long)e->n_sites, (long long)e->n_atoms, need, NOALLOC_ELEMS);
  return guarded([&] {
    double2* s = scratch_of(e->device);
    const size_t ns = (size_t)e->n_sites;
    // no kernel timing here: its events and records would allocate
    const uint32_t kmask = e->kmask;
    e->kmask = 0;
    try {
      potential_run(e, state - 1, s, s + ns, s + 2 * ns, v);
    } catch (...) {
      e->kmask = kmask;
      throw;
    }
    e->kmask = kmask;
  });
}

int clat_last_stage_ms(clat_engine* e, double ms[5]) {
  if (!e || !ms) return set_err(CLAT_EINVAL, "invalid arguments");
  return guarded([&] {
    settle_table(e);
    for (int i = 0; i < 5; ++i) ms[i] = e->stage_ms[i];
  });
}

int clat_set_kernel_timing(clat_engine* e, int enable) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  e->kmask = enable ? ~0u : 0u;
  return CLAT_OK;
}

int clat_set_kernel_timing_kinds(clat_engine* e, uint32_t kinds) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  e->kmask = kinds;
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
  if (device_bytes) *device_bytes = e->device_bytes + (e->sub ? e->sub->device_bytes : 0);
  return CLAT_OK;
}

int clat_engine_kernels(clat_engine* e, uint32_t* flags) {
  if (!e || !flags) return set_err(CLAT_EINVAL, "invalid arguments");
  uint32_t f = 0;
  if (e->fast) f |= CLAT_KERNELS_TWO_LEVEL;
  if (e->fast && e->colk && col_pair_compiled(e->cyf.R1, e->cyf.R2)) f |= CLAT_KERNELS_COLUMN;
  if (e->xline) f |= CLAT_KERNELS_LINE;
  if (e->conv_tiled) f |= CLAT_KERNELS_CONV_TILED;
  if (e->btile) f |= CLAT_KERNELS_B_STRIPS;
  if (e->atile) f |= CLAT_KERNELS_AD_STRIPS;
  if (e->coset) f |= CLAT_KERNELS_COSETS;
  if (e->coset && e->sub->half) f |= CLAT_KERNELS_HALF_BOX;
  if (e->coset && e->sub->xhalf) f |= CLAT_KERNELS_HALF_LINES;
  *flags = f;
  return CLAT_OK;
}

int clat_set_tuning(clat_engine* e, int measure) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  e->tune = measure ? 1 : 0;
  return CLAT_OK;
}

int clat_tuning_info(clat_engine* e, int* batch, double trial_ms[8], int64_t* graph_replays,
                     int64_t* graph_captures) {
  if (!e) return set_err(CLAT_EINVAL, "invalid arguments");
  if (batch) *batch = e->tune ? e->tuned_nb : batch_rule(e);
  if (trial_ms)
    for (int i = 0; i < 8; ++i) trial_ms[i] = e->tune_ms[i];
  if (graph_replays) *graph_replays = e->graph_replays;
  if (graph_captures) *graph_captures = e->graph_captures;
  return CLAT_OK;
}

}  // extern "C"
