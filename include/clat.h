/*
 * clat.h -- C ABI of the MI355X FFT Coulomb-integral engine.
 *
 * Every entry point returns an int status (CLAT_OK == 0) and never throws.
 * On failure clat_last_error() returns a thread-local message.  Host arrays
 * passed in are borrowed for the duration of the call only; the engine owns
 * all device memory.  One host thread per engine handle.  All calls are
 * ordered on the engine's own HIP stream and synchronous at return.
 *
 * Complex numbers are passed as interleaved (re, im) doubles.  Grids are
 * x-fastest: cell (ix, iy, iz) lives at (iz*Y + iy)*X + ix, the layout of the
 * reference's Domain views (src/Domain.hpp:32, run-tests.cpp:170-175).
 *
 * Reference interfaces each entry point replaces are cited as
 * file:line relative to the reference repository root.
 */
#ifndef CLAT_H
#define CLAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CLAT_OK        0
#define CLAT_EINVAL    1  /* bad argument (reference: std::runtime_error / logic_error) */
#define CLAT_ENOMEM    2  /* device or host allocation failed */
#define CLAT_EDEVICE   3  /* HIP runtime error or no usable GPU */
#define CLAT_ESTATE    4  /* call out of protocol order (e.g. prepare before initialize) */
#define CLAT_EUNSUP    5  /* size outside what the kernels support */

/* ---- engine options ---------------------------------------------------- */
/* Arithmetic of the FFT passes.  The interaction table, its DCT-I and every
 * overlap sum are FP64 in both modes. */
#define CLAT_PREC_FP64 0  /* complex128 passes, phases and kernel multiply:
                             every integral within 1e-10 of itself against
                             the FP64 CPU restatement of the reference      */
#define CLAT_PREC_FP32 1  /* complex64 passes, phases and kernel multiply:
                             every integral within 1e-5 of max(|I|, 1e-3
                             max|I|) against FP64                           */
#define CLAT_PREC_FP32X 2 /* complex64 storage and arithmetic in every pass
                             but the last; the last (inverse x) pass computes
                             in FP64 and writes the potential in FP64.  On the
                             1M-atom, 8-state workload every integral is
                             within 1e-5 of itself against FP64 (5.6e-6); on
                             the batched and six-box workloads every integral
                             of at least 1e-6 of the run's largest is, and
                             the rest are within 1e-11 of the largest.  In
                             general (small grids included): within 1e-5 of
                             itself or 3e-10 of the run's largest, whichever
                             is larger -- rounding a pair density to
                             complex64 at all moves integrals that cancel to
                             ~1e-8 of the largest by more than 1e-5 of
                             themselves                                     */

/* Schedule of the eight half-shift rounds (CoulombCalculator.cpp:105-163). */
#define CLAT_SCHED_SHARED 0  /* rounds share their x and y passes: the eight
                                parity classes are produced by pruned 2X/2Y/2Z
                                transforms of the unpadded grid              */
#define CLAT_SCHED_ROUNDS 1  /* eight independent rounds, each a full forward
                                and inverse 3-D FFT, as the reference loops   */

/* flags of clat_compute_integrals */
#define CLAT_FLAG_STALE_VALUES 1  /* reproduce coulombo.cpp:194-205 exactly:
                                     reuse the previous overlap when the right
                                     product is unchanged across a new left */

typedef struct clat_engine clat_engine;

/* Thread-local text of the last failure ("" if none). */
const char* clat_last_error(void);

/* Library version string, e.g. "clat 1.0 (gfx950)". */
const char* clat_version(void);

/* ---- host-side lattice and planning helpers (no GPU needed) ------------ */

/* Round::up<p1,p2,...>(n, c): smallest integer >= n of the form
 * c * p1^a * p2^b * ...   Replaces src/Round.hpp:20-27,38-49. */
int clat_round_up(int n, int c, const int* primes, int n_primes, int* out);

/* The axis length an engine computes on for a user length n: n itself when
 * its two-level split has compile-time kernels, else the smallest such
 * length above n (n >= 32; exact: the open-boundary convolution does not
 * depend on the padding), else the smallest two-level length, else the next
 * 2*3*5*7-smooth length.  No counterpart in the reference. */
int clat_engine_length(int n, int* out);

/* Lattice detection of src/Broadcaster.cpp:24-43,53-56 (computeGrid x3 +
 * round4FFT).  xyz is atom-major (x,y,z per atom, Angstrom).  Outputs raw and
 * 2*3*5-padded sizes, origin and step per axis.  Fails with CLAT_EINVAL and
 * "could not find common grid step" as Broadcaster.cpp:38. */
int clat_lattice_grid(const double* xyz, int64_t n_atoms, int dim_raw[3],
                      int dim_pad[3], double origin[3], double step[3]);

/* Atom -> linear cell index ((iz*Y + iy)*X + ix) with lrint rounding,
 * Broadcaster.cpp:76-88 at one rank.  dim is the padded dimension. */
int clat_lattice_cells(const double* xyz, int64_t n_atoms, const int dim[3],
                       const double origin[3], const double step[3],
                       int32_t* cell);

/* Enumerate, in the order of coulombo.cpp:135-148, every (i1,i2,i3,i4) in
 * [1,n_states]^4 matched by the comma-separated pattern list (src/Pattern.hpp).
 * Writes up to cap quads (4 ints each, 1-based) and the total count.
 * quads may be NULL to count only. */
int clat_pattern_enumerate(const char* spec, int n_states, int n_holes,
                           int32_t* quads, int64_t cap, int64_t* count);

/* Greedy vertex cover of src/Graph.hpp:49-91.  edges: n_edges (u,v) pairs.
 * cover receives the sorted cover (capacity n_vertices). */
int clat_vertex_cover(int n_vertices, const int32_t* edges, int64_t n_edges,
                      int32_t* cover, int* n_cover);

/* MasterPlanner (src/Planner.hpp:65-145): for each quad, left = product(i1,i4),
 * right = product(i2,i3); cover swap; sort.  steps receives n rows of
 * {id, left, left_conj, right, right_conj} in execution order. */
int clat_plan(const int32_t* quads, int64_t n, int n_states, int32_t* steps);

/* ---- engine: the CoulombCalculator protocol --------------------------- */

/* CoulombCalculator(const Dimension&, bool) -- CoulombCalculator.cpp:11-55.
 * X, Y, Z: grid dimension.  precision: CLAT_PREC_*, schedule: CLAT_SCHED_*.
 * Sizes with a prime factor above 7 are carried on the next 2*3*5*7-smooth
 * grid; results are unchanged because the open-boundary convolution is exact
 * for any padding. */
int clat_create(int X, int Y, int Z, int precision, int schedule,
                clat_engine** out);

/* Releases every device buffer of the engine. */
int clat_destroy(clat_engine* e);

/* InteractionSimple + InteractionBase::map + CoulombCalculator::initialize
 * (Interaction.hpp:32-51, Interaction.cpp:8-14, CoulombCalculator.cpp:57-103):
 * builds G(ix,iy,iz) = 14.39963737103201/(dielectric*r) on the
 * (X+1)(Y+1)(Z+1) first octant with G(0,0,0) = onsite, then its 3-D DCT-I
 * divided by 8*X*Y*Z -- all in FP64 on the GPU.  The table is built by the
 * first call that needs it (clat_prepare, clat_compute_integrals, after that
 * call's pair densities, clat_compute_potential_at_atoms*), so a failure of
 * its kernels is reported there; the same holds for the Thomas-Fermi form. */
int clat_set_interaction_simple(clat_engine* e, const double step[3],
                                double onsite, double dielectric);

/* InteractionThomasFermi (Interaction.cpp:18-39), dielectric > 1,
 * lattice_constant > 0 (coulombo.cpp:166-174). */
int clat_set_interaction_thomas_fermi(clat_engine* e, const double step[3],
                                      double onsite, double dielectric,
                                      double lattice_constant);

/* Arbitrary Interaction::map result (Interaction.hpp:14): G on the
 * (X+1)(Y+1)(Z+1) first octant, x fastest, then CoulombCalculator::initialize. */
int clat_set_interaction_table(clat_engine* e, const double* G);

/* Write the calculator's `input` domain (CoulombCalculator.hpp:48): F is
 * X*Y*Z complex numbers. */
int clat_set_input(clat_engine* e, const double* F);

/* CoulombCalculator::prepare (CoulombCalculator.cpp:105-163): V = G * input. */
int clat_prepare(clat_engine* e);

/* CoulombCalculator::calculate (CoulombCalculator.cpp:165-172):
 * sum_n V(n) * input(n), bilinear, FP64 compensated sum. */
int clat_calculate(clat_engine* e, double result[2]);

/* CoulombCalculator::potential (CoulombCalculator.hpp:90-93): copies V
 * (X*Y*Z complex) to the host. */
int clat_get_potential(clat_engine* e, double* V);

/* ---- engine: atom-site path (FunctionCollection + Product + main loop) - */

/* Atom -> cell map of the lattice (Broadcaster.cpp:76-88) and the orbital
 * count (coulombo.cpp:59-65).  Where several atoms share a cell, the last
 * one wins, as Product::map's assignment does (Product.cpp:22,27).  States
 * loaded before the call are dropped (they belong to the old atom list). */
int clat_set_atoms(clat_engine* e, const int32_t* cell, int64_t n_atoms,
                   int n_orbitals);

/* FunctionCollection::appendFile (FunctionCollection.cpp:56-88): one state,
 * n_atoms*n_orbitals complex coefficients, atom-major / orbital-minor.
 * States are numbered 1.. in call order; holes must come first. */
int clat_add_state(clat_engine* e, const double* coeffs);

/* Drop every state added so far. */
int clat_clear_states(clat_engine* e);

/* The coulombo main loop (coulombo.cpp:150-209) for the given quads
 * (n rows of 1-based i1,i2,i3,i4): plan, pair densities, prepares of the
 * cover, overlaps, conjugation bookkeeping.  out: n complex values in quad
 * order.  flags: CLAT_FLAG_*. */
int clat_compute_integrals(clat_engine* e, const int32_t* quads, int64_t n,
                           int flags, double* out);

/* clat_compute_integrals for several systems at once (one engine each, any
 * grid sizes, all on one device): quads[i], n[i] and out[i] belong to
 * engines[i], flags to all.  No counterpart in the reference, which runs one
 * system per process.  Each engine builds its pair densities and plan as in
 * clat_compute_integrals; the device stages of all of them (prepares and
 * overlaps) are then captured into one HIP graph on engines[0]'s stream and
 * launched once per call, replayed while every member's plan and buffers stay
 * the same (the graph belongs to engines[0]; a member that leads no group
 * keeps its own graphs).  With kernel timing on in any member, or when the
 * runtime refuses the capture, each engine launches its own device stage.
 * Values are bit-identical to per-engine clat_compute_integrals calls.
 * clat_last_stage_ms of each member reports the whole group device stage as
 * [2] and 0 as [3].  An engine may appear once per group; the engines must
 * not be in use by other threads during the call. */
int clat_compute_integrals_group(clat_engine* const* engines, int count, const int32_t* const* quads,
                                 const int64_t* n, int flags, double* const* out);

/* potentials.cpp:140-158: V of the self product of `state` (1-based) at every
 * atom's cell (FunctionCollection::extractAtomCellValues,
 * FunctionCollection.cpp:122-131).  v: n_atoms complex values. */
int clat_compute_potential_at_atoms(clat_engine* e, int state, double* v);

/* The same potential without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The self density, the
 * potential at the sites and the per-atom values live in one static device
 * array of 4,194,304 complex elements (64 MiB) that every GPU running the
 * library holds until the process ends: 2 * (occupied cells) + n_atoms must
 * not exceed that, else CLAT_EUNSUP with the sizes in the message.  Needs an
 * engine on the two-level pass kernels (every axis of at most 400 points
 * splitting into two factors of at most 20; else CLAT_EUNSUP).  Kernel timing
 * is not recorded during the call.  Calls on one GPU share that array: calls
 * that overlap in time on one GPU (engines driven from several threads) race
 * and must be serialised by the caller.  Same values as
 * clat_compute_potential_at_atoms (same kernels). */
int clat_compute_potential_at_atoms_noalloc(clat_engine* e, int state, double* v);

/* ---- measurement ------------------------------------------------------- */

/* Stage times in milliseconds from HIP events on the engine stream:
 * [0] the last interaction build (G table + DCT-I + scaling), and for the
 * last clat_compute_integrals call [1] pair densities, [2] all prepares,
 * [3] overlaps, [4] the whole call (with [0] when that call built the
 * table). */
int clat_last_stage_ms(clat_engine* e, double ms[5]);

/* When enabled, every FFT pass launch is bracketed by HIP events on the
 * engine stream and timed per kernel kind. */
int clat_set_kernel_timing(clat_engine* e, int enable);

/* The same for a subset of kernel kinds: bit k of `kinds` times kind k (0
 * times nothing).  A measurement hook with no counterpart in the reference:
 * the events cost a few microseconds per launch, so a run that reports one
 * kernel's duration times only that kernel. */
int clat_set_kernel_timing_kinds(clat_engine* e, uint32_t kinds);

/* Per-kind totals since the last reset: launches, summed milliseconds and the
 * algorithmic bytes those launches must move (one read of each input element
 * and one write of each output element).  kind < clat_kernel_kind_count(). */
int clat_kernel_kind_count(void);
const char* clat_kernel_kind_name(int kind);
int clat_kernel_stats(clat_engine* e, int kind, int64_t* launches,
                      double* total_ms, double* bytes);
int clat_reset_kernel_stats(clat_engine* e);

/* Grid size the engine computes on (after 2*3*5*7 padding) and the device
 * bytes it currently holds. */
int clat_engine_info(clat_engine* e, int dim[3], int64_t* device_bytes);

/* Which kernels the engine's passes run on (bits of *flags): the two-level
 * register FFT kernels (else the LDS Stockham kernels), compile-time column
 * kernels for the y passes, the x line kernels of the atom path (occupied
 * lines only), the tiled conv kernel table, the column-strip layouts of the
 * intermediate planes, and the two-coset path of the atom path (set by
 * clat_set_atoms when every site lies on x = y = z (mod 2) of an even grid,
 * as the zincblende and diamond lattices on their a/4 grid do: the
 * convolution then runs on the two cosets of the coarse grid with half the
 * points per axis, coupled by a 2 x 2 kernel; CLAT_COSETS=0 in the
 * environment keeps the single-grid path), and its half-box form (every
 * coarse position on mx + my + mz even, as zincblende's are: only the x
 * frequencies below half the coarse box are carried, against kernels summed
 * over the half-box shift; CLAT_HALFBOX=0 keeps all of them), and in the
 * half box the x passes on half-length lines (a coarse line's sites all have
 * x = y + z (mod 2): the line is carried as its Xc / 2 sites and the x
 * parity enters as a phase; CLAT_HALFLINES=0 keeps full lines).  An
 * introspection hook with no counterpart in the reference: every combination
 * gives the same results. */
#define CLAT_KERNELS_TWO_LEVEL 1u
#define CLAT_KERNELS_COLUMN 2u
#define CLAT_KERNELS_LINE 4u
#define CLAT_KERNELS_CONV_TILED 8u
#define CLAT_KERNELS_B_STRIPS 16u
#define CLAT_KERNELS_AD_STRIPS 32u
#define CLAT_KERNELS_COSETS 64u
#define CLAT_KERNELS_HALF_BOX 128u
#define CLAT_KERNELS_HALF_LINES 256u
int clat_engine_kernels(clat_engine* e, uint32_t* flags);

/* ---- planning ---------------------------------------------------------- */

/* The measureFFT choice of CoulombCalculator(const Dimension&, bool
 * measureFFT = true) (CoulombCalculator.hpp:57, CoulombCalculator.cpp:23,31),
 * for this engine's own plans.  measure = 1 (the default, as in the
 * reference): the first clat_compute_integrals of an atom list times the
 * whole prepare sequence with 1, 2, 4 and 8 prepares per pass launch and the
 * fixed rule's count, and keeps the fastest (the rule's count where it is
 * within 2 % of the fastest), for the engine and for every later engine of
 * the process with the same grid, arithmetic and occupied lines.  The counts
 * give bit-identical results; the first run takes a few prepare sequences
 * longer.  measure = 0: the fixed rule (four, or eight on grids whose eight
 * intermediate planes take at most 1.1 GB; on the two-coset path as many as
 * keep a launch's intermediate planes within 192 MiB, three quarters of the
 * device's 256 MiB Infinity Cache, where four or more fit).  CLAT_TUNE=0 in the
 * environment at clat_create turns measuring off; CLAT_BATCH=n fixes the
 * count. */
int clat_set_tuning(clat_engine* e, int measure);

/* The prepares per launch in use (0: not measured yet), the measured
 * prepare-sequence milliseconds of k prepares per launch in trial_ms[k - 1],
 * k = 1..8 (0: not tried), and how
 * often the device stage of clat_compute_integrals was captured into and
 * replayed from HIP graphs (runs with an unchanged plan and buffers replay
 * two graphs, prepares and overlaps, instead of launching every pass; runs
 * with kernel timing on, and every run after CLAT_GRAPH=0 at clat_create,
 * launch directly).  No counterpart in the reference. */
int clat_tuning_info(clat_engine* e, int* batch, double trial_ms[8], int64_t* graph_replays,
                     int64_t* graph_captures);

#ifdef __cplusplus
}
#endif

#endif /* CLAT_H */
