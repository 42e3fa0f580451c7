/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.  FP64 CPU restatement of the
 * reference's FFT Coulomb-integral path, used by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg as the checker.  Never linked into the
 * product library.
 *
 * Functions and the reference code they follow (paths relative to the
 * reference repository root):
 *   orc_round_up            src/Round.hpp:20-27,38-49
 *   orc_lattice_grid        src/Broadcaster.cpp:8-43,53-56,136-143
 *   orc_lattice_cells       src/Broadcaster.cpp:76-88 (one rank)
 *   orc_gtable              src/Interaction.hpp:32-51, src/Interaction.cpp:8-39
 *   orc_dct1_3d             FFTW REDFT00 on every axis as planned in
 *                           src/CoulombCalculator.cpp:33-39, executed at :61;
 *                           restated from its definition as a direct O(n^2) sum
 *   orc_prepare             src/CoulombCalculator.cpp:105-163 (8 rounds), with
 *                           Gfreq = DCT/(8XYZ) (:102) and the y<->z transposed
 *                           lookup (:130-144) written in untransposed indices
 *   orc_calculate           src/CoulombCalculator.cpp:165-172
 *   orc_products            src/Product.cpp:16-30, src/FunctionCollection.cpp:91-105
 *   orc_direct_potential    direct O(N^2) lattice sum V(n) = sum_m G(|n-m|) F(m):
 *                           the quantity the 8-round method computes exactly
 *   orc_fft                 FFTW_FORWARD / FFTW_BACKWARD conventions (unnormalised)
 *
 * Complex arrays are interleaved (re, im) doubles; grids are x-fastest.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define E2_4PE0 14.39963737103201 /* src/base.hpp:28 */

typedef struct { double re, im; } cplx;

/* ------------------------------------------------------------------ Round */
static int64_t round_rec(int64_t n, int64_t c, const int* p, int np) {
  if (np == 1) {
    while (c < n) c *= p[0];
    return c;
  }
  int64_t best = round_rec(n, c, p + 1, np - 1);
  while (c < n) {
    c *= p[0];
    int64_t b = round_rec(n, c, p + 1, np - 1);
    if (b < best) best = b;
  }
  return best;
}

int orc_round_up(int n, int c, const int* primes, int np) { return (int)round_rec(n, c, primes, np); }

/* ---------------------------------------------------------------- lattice */
static double fgcd(double a, double b) {
  if (a < b) { double t = a; a = b; b = t; }
  for (;;) {
    double r = fmod(a, b);
    if (r < 1.0e-10) return b;
    if (r < 1.0e-2) return 0.0;
    a = b;
    b = r;
  }
}

static int cmp_d(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

/* returns 0 on success, 1 "could not find common grid step" */
int orc_lattice_grid(const double* xyz, int64_t n, int raw[3], int pad[3], double origin[3], double step[3]) {
  static const int p235[3] = {2, 3, 5};
  double* v = (double*)malloc(sizeof(double) * (size_t)n);
  for (int d = 0; d < 3; ++d) {
    for (int64_t a = 0; a < n; ++a) v[a] = xyz[3 * a + d];
    qsort(v, (size_t)n, sizeof(double), cmp_d);
    double s = NAN;
    for (int64_t a = 1; a < n; ++a) {
      double nx = v[a] - v[a - 1];
      if (nx > 1.0e-10) {
        s = isnan(s) ? nx : fgcd(s, nx);
        if (s == 0) { free(v); return 1; }
      }
    }
    origin[d] = v[0];
    if (isnan(s)) { raw[d] = 1; step[d] = 1.0; }
    else { raw[d] = (int)(lrint((v[n - 1] - v[0]) / s) + 1); step[d] = s; }
    pad[d] = orc_round_up(raw[d], 1, p235, 3);
  }
  free(v);
  return 0;
}

void orc_lattice_cells(const double* xyz, int64_t n, const int dim[3], const double origin[3],
                       const double step[3], int32_t* cell) {
  for (int64_t a = 0; a < n; ++a) {
    long ix = lrint((xyz[3 * a] - origin[0]) / step[0]);
    long iy = lrint((xyz[3 * a + 1] - origin[1]) / step[1]);
    long iz = lrint((xyz[3 * a + 2] - origin[2]) / step[2]);
    cell[a] = (int32_t)((iz * dim[1] + iy) * dim[0] + ix);
  }
}

/* --------------------------------------------------------- interaction G */
/* model 0: simple; model 1: Thomas-Fermi-Resta.  Table (X+1)(Y+1)(Z+1). */
static double tf_rTF(double eps, double a, double* qTF) {
  const double tol = 1.0e-12;
  double q = 2.0 / sqrt(M_PI) * cbrt(96.0 * M_PI * M_PI) / a;
  double x = sqrt(6.0 * (eps - 1.0)), dx;
  do {
    double F = sinh(x) - eps * x;
    double dF = cosh(x) - eps;
    dx = (fabs(F) < tol) ? 0 : F / dF;
    x -= dx;
  } while (dx / q > tol);
  *qTF = q;
  return x / q;
}

void orc_gtable(int X1, int Y1, int Z1, const double step[3], double onsite, double dielectric, int model,
                double lattice_constant, double* G) {
  double qTF = 0.0, rTF = 0.0;
  if (model == 1) rTF = tf_rTF(dielectric, lattice_constant, &qTF);
  for (int iz = 0; iz < Z1; ++iz) {
    double z = iz * step[2], z2 = z * z;
    for (int iy = 0; iy < Y1; ++iy) {
      double y = iy * step[1], y2z2 = z2 + y * y;
      for (int ix = 0; ix < X1; ++ix) {
        double x = ix * step[0], r = sqrt(y2z2 + x * x);
        double eps = dielectric;
        if (model == 1 && r < rTF) eps = dielectric * (qTF * rTF / (sinh(qTF * (rTF - r)) + qTF * r));
        G[((size_t)iz * Y1 + iy) * X1 + ix] = E2_4PE0 / (eps * r);
      }
    }
  }
  G[0] = onsite;
}

/* ------------------------------------------------------------------ DCT-I */
/* REDFT00 of length M+1: Y_k = x_0 + (-1)^k x_M + 2 sum_{n=1}^{M-1} x_n cos(pi n k / M) */
static void dct1_line(const double* x, double* y, int n1, int stride, const double* cosT) {
  const int M = n1 - 1;
  for (int k = 0; k < n1; ++k) {
    if (M == 0) { y[0] = x[0]; break; }
    double s = x[0] + ((k & 1) ? -x[(size_t)M * stride] : x[(size_t)M * stride]);
    for (int n = 1; n < M; ++n) s += 2.0 * x[(size_t)n * stride] * cosT[((size_t)n * k) % (2 * M)];
    y[k] = s;
  }
}

void orc_dct1_3d(double* G, int X1, int Y1, int Z1) {
  int dims[3] = {X1, Y1, Z1};
  size_t strides[3] = {1, (size_t)X1, (size_t)X1 * Y1};
  for (int ax = 0; ax < 3; ++ax) {
    const int n1 = dims[ax];
    const int M = n1 - 1;
    double* cosT = (double*)malloc(sizeof(double) * (size_t)(2 * M + 1));
    for (int j = 0; j < 2 * M; ++j) cosT[j] = (double)cosl(3.141592653589793238462643383279502884L * j / M);
    const size_t total = (size_t)X1 * Y1 * Z1;
    const size_t nlines = total / n1;
#pragma omp parallel for schedule(static)
    for (size_t l = 0; l < nlines; ++l) {
      /* base index of line l along axis ax */
      size_t base;
      if (ax == 0) base = l * X1;
      else if (ax == 1) base = (l / X1) * (size_t)X1 * Y1 + (l % X1);
      else base = l;
      double* tmp = (double*)malloc(sizeof(double) * (size_t)(2 * n1));
      for (int i = 0; i < n1; ++i) tmp[i] = G[base + (size_t)i * strides[ax]];
      dct1_line(tmp, tmp + n1, n1, 1, cosT);
      for (int i = 0; i < n1; ++i) G[base + (size_t)i * strides[ax]] = tmp[n1 + i];
      free(tmp);
    }
    free(cosT);
  }
}

/* ---------------------------------------------------------------- FFT */
/* Mixed-radix recursive Cooley-Tukey, any size (generic prime butterflies).
 * y[k] = sum_j x[j*s] W^{jk}, W = e^{sign 2 pi i / n}.  tab[j] = e^{sign 2 pi i j / N}. */
static int smallest_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

static void fft_rec(const cplx* x, int s, cplx* y, int n, int N, const cplx* tab, cplx* tmp) {
  if (n == 1) { y[0] = x[0]; return; }
  int p = smallest_factor(n), m = n / p;
  for (int r = 0; r < p; ++r) fft_rec(x + (size_t)r * s, s * p, y + (size_t)r * m, m, N, tab, tmp);
  const int step = N / n;
  for (int k1 = 0; k1 < m; ++k1) {
    for (int r = 0; r < p; ++r) tmp[r] = y[(size_t)r * m + k1];
    for (int q = 0; q < p; ++q) {
      const int k = k1 + q * m;
      double re = 0, im = 0;
      for (int r = 0; r < p; ++r) {
        const cplx w = tab[(size_t)(((int64_t)r * k) % n) * step];
        re += tmp[r].re * w.re - tmp[r].im * w.im;
        im += tmp[r].re * w.im + tmp[r].im * w.re;
      }
      y[k] = (cplx){re, im};
    }
  }
}

static cplx* make_tab(int N, int sign) {
  cplx* t = (cplx*)malloc(sizeof(cplx) * (size_t)N);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int j = 0; j < N; ++j) {
    long double a = sign * 2.0L * pi * j / N;
    t[j].re = (double)cosl(a);
    t[j].im = (double)sinl(a);
  }
  return t;
}

/* 1-D transform of a strided line in place. sign -1 forward, +1 backward. */
void orc_fft(double* data, int n, int sign) {
  cplx* tab = make_tab(n, sign);
  cplx* in = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)(n + 1));
  memcpy(in, data, sizeof(cplx) * (size_t)n);
  fft_rec(in, 1, (cplx*)data, n, n, tab, tmp);
  free(tab); free(in); free(tmp);
}

/* 3-D transform of an X*Y*Z x-fastest grid, in place (FFTW dims (z,y,x)). */
static void fft3(cplx* a, int X, int Y, int Z, int sign) {
  const int dims[3] = {X, Y, Z};
  for (int ax = 0; ax < 3; ++ax) {
    const int n = dims[ax];
    if (n == 1) continue;
    cplx* tab = make_tab(n, sign);
    const size_t stride = ax == 0 ? 1 : (ax == 1 ? (size_t)X : (size_t)X * Y);
    const size_t nlines = (size_t)X * Y * Z / n;
#pragma omp parallel
    {
      cplx* in = (cplx*)malloc(sizeof(cplx) * (size_t)n);
      cplx* out = (cplx*)malloc(sizeof(cplx) * (size_t)n);
      cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)(n + 1));
#pragma omp for schedule(static)
      for (size_t l = 0; l < nlines; ++l) {
        size_t base;
        if (ax == 0) base = l * X;
        else if (ax == 1) base = (l / X) * (size_t)X * Y + (l % X);
        else base = l;
        for (int i = 0; i < n; ++i) in[i] = a[base + i * stride];
        fft_rec(in, 1, out, n, n, tab, tmp);
        for (int i = 0; i < n; ++i) a[base + i * stride] = out[i];
      }
      free(in); free(out); free(tmp);
    }
    free(tab);
  }
}

/* -------------------------------------------------------- 8-round prepare */
/* Ghat: DCT-I'd table of (X+1)(Y+1)(Z+1), already divided by 8XYZ.
 * F, V: X*Y*Z complex.  work: X*Y*Z complex scratch. */
void orc_prepare(const double* Fd, int X, int Y, int Z, const double* Ghat, double* Vd) {
  const cplx* F = (const cplx*)Fd;
  cplx* V = (cplx*)Vd;
  const size_t N = (size_t)X * Y * Z;
  cplx* T = (cplx*)malloc(sizeof(cplx) * N);
  cplx *px = make_tab(2 * X, +1), *py = make_tab(2 * Y, +1), *pz = make_tab(2 * Z, +1); /* e^{i pi n/N} */
  const int X1 = X + 1, Y1 = Y + 1;
  memset(V, 0, sizeof(cplx) * N);
  for (int round = 0; round < 8; ++round) {
    const int kx = round & 1, ky = (round >> 1) & 1, kz = (round >> 2) & 1;
    memcpy(T, F, sizeof(cplx) * N);
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Z; ++iz)
      for (int iy = 0; iy < Y; ++iy)
        for (int ix = 0; ix < X; ++ix) {
          cplx ph = {1.0, 0.0};
          if (kx) ph = px[ix];
          if (ky) ph = (cplx){ph.re * py[iy].re - ph.im * py[iy].im, ph.re * py[iy].im + ph.im * py[iy].re};
          if (kz) ph = (cplx){ph.re * pz[iz].re - ph.im * pz[iz].im, ph.re * pz[iz].im + ph.im * pz[iz].re};
          cplx* t = &T[((size_t)iz * Y + iy) * X + ix];
          const cplx v = *t; /* multiply by conj(ph) */
          *t = (cplx){v.re * ph.re + v.im * ph.im, v.im * ph.re - v.re * ph.im};
        }
    fft3(T, X, Y, Z, -1);
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Z; ++iz) {
      int fz = 2 * iz + kz; if (fz > 2 * Z - fz) fz = 2 * Z - fz;
      for (int iy = 0; iy < Y; ++iy) {
        int fy = 2 * iy + ky; if (fy > 2 * Y - fy) fy = 2 * Y - fy;
        for (int ix = 0; ix < X; ++ix) {
          int fx = 2 * ix + kx; if (fx > 2 * X - fx) fx = 2 * X - fx;
          const double g = Ghat[((size_t)fz * Y1 + fy) * X1 + fx];
          cplx* t = &T[((size_t)iz * Y + iy) * X + ix];
          t->re *= g;
          t->im *= g;
        }
      }
    }
    fft3(T, X, Y, Z, +1);
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < Z; ++iz)
      for (int iy = 0; iy < Y; ++iy)
        for (int ix = 0; ix < X; ++ix) {
          cplx ph = {1.0, 0.0};
          if (kx) ph = px[ix];
          if (ky) ph = (cplx){ph.re * py[iy].re - ph.im * py[iy].im, ph.re * py[iy].im + ph.im * py[iy].re};
          if (kz) ph = (cplx){ph.re * pz[iz].re - ph.im * pz[iz].im, ph.re * pz[iz].im + ph.im * pz[iz].re};
          const size_t i = ((size_t)iz * Y + iy) * X + ix;
          const cplx v = T[i];
          V[i].re += v.re * ph.re - v.im * ph.im;
          V[i].im += v.re * ph.im + v.im * ph.re;
        }
  }
  free(T); free(px); free(py); free(pz);
}

/* Neumaier-compensated bilinear sum  sum_n V(n) F(n) */
void orc_calculate(const double* Vd, const double* Fd, int64_t n, double out[2]) {
  const cplx* V = (const cplx*)Vd;
  const cplx* F = (const cplx*)Fd;
  double s[2] = {0, 0}, c[2] = {0, 0};
  for (int64_t i = 0; i < n; ++i) {
    double x[2] = {V[i].re * F[i].re - V[i].im * F[i].im, V[i].re * F[i].im + V[i].im * F[i].re};
    for (int k = 0; k < 2; ++k) {
      double t = s[k] + x[k];
      c[k] += (fabs(s[k]) >= fabs(x[k])) ? (s[k] - t) + x[k] : (x[k] - t) + s[k];
      s[k] = t;
    }
  }
  out[0] = s[0] + c[0];
  out[1] = s[1] + c[1];
}

/* rho[p][a] for p = fL(fL+1)/2 + fR, fR <= fL:  sum_o conj(c_fL[a][o]) c_fR[a][o].
 * coeffs: n_states x n_atoms x n_orb complex. */
void orc_products(const double* cd, int n_states, int64_t n_atoms, int n_orb, double* rd) {
  const cplx* c = (const cplx*)cd;
  cplx* rho = (cplx*)rd;
  int p = 0;
  for (int fL = 0; fL < n_states; ++fL)
    for (int fR = 0; fR <= fL; ++fR, ++p) {
#pragma omp parallel for schedule(static)
      for (int64_t a = 0; a < n_atoms; ++a) {
        const cplx* l = c + ((size_t)fL * n_atoms + a) * n_orb;
        const cplx* r = c + ((size_t)fR * n_atoms + a) * n_orb;
        double re = 0, im = 0;
        for (int o = 0; o < n_orb; ++o) {
          re += l[o].re * r[o].re + l[o].im * r[o].im;
          im += l[o].re * r[o].im - l[o].im * r[o].re;
        }
        rho[(size_t)p * n_atoms + a] = (cplx){re, im};
      }
    }
}

/* Direct lattice sum on the grid: V(n) = sum_m G(|n-m|) F(m) for the cells n
 * listed in `at` (n_at of them), over the nonzero cells of F (the sources).
 * G is the first-octant table (X+1)(Y+1)(Z+1) before any transform. */
void orc_direct_potential(const double* Fd, int X, int Y, int Z, const double* G, const int32_t* at,
                          int64_t n_at, double* outd) {
  const cplx* F = (const cplx*)Fd;
  cplx* out = (cplx*)outd;
  const size_t N = (size_t)X * Y * Z;
  int64_t ns = 0;
  for (size_t i = 0; i < N; ++i)
    if (F[i].re != 0.0 || F[i].im != 0.0) ++ns;
  int32_t* src = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ns ? ns : 1));
  ns = 0;
  for (size_t i = 0; i < N; ++i)
    if (F[i].re != 0.0 || F[i].im != 0.0) src[ns++] = (int32_t)i;
  const int X1 = X + 1, Y1 = Y + 1;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < n_at; ++t) {
    const int c = at[t];
    const int ix = c % X, iy = (c / X) % Y, iz = c / (X * Y);
    double sr = 0, si = 0, cr = 0, ci = 0;
    for (int64_t k = 0; k < ns; ++k) {
      const int m = src[k];
      const int jx = m % X, jy = (m / X) % Y, jz = m / (X * Y);
      const double g = G[((size_t)abs(iz - jz) * Y1 + abs(iy - jy)) * X1 + abs(ix - jx)];
      const double xr = g * F[m].re, xi = g * F[m].im;
      double u = sr + xr;
      cr += (fabs(sr) >= fabs(xr)) ? (sr - u) + xr : (xr - u) + sr;
      sr = u;
      u = si + xi;
      ci += (fabs(si) >= fabs(xi)) ? (si - u) + xi : (xi - u) + si;
      si = u;
    }
    out[t] = (cplx){sr + cr, si + ci};
  }
  free(src);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
