"""GPU parity of the HIP engine against the FP64 CPU oracle and the
reference's own known-answer values.  Every call goes through the C ABI."""
import os

import numpy as np
import pytest

import clat
import oracle as O
import zincblende as zb
from parity import TOL, check, rel_err

pytestmark = pytest.mark.gpu

COMBOS = [(clat.PREC_FP64, clat.SCHED_SHARED), (clat.PREC_FP64, clat.SCHED_ROUNDS),
          (clat.PREC_FP32, clat.SCHED_SHARED), (clat.PREC_FP32, clat.SCHED_ROUNDS),
          (clat.PREC_FP32X, clat.SCHED_SHARED), (clat.PREC_FP32X, clat.SCHED_ROUNDS)]
IDS = ["fp64-shared", "fp64-rounds", "fp32-shared", "fp32-rounds", "fp32x-shared", "fp32x-rounds"]
# per-integral bars (tests/parity.py): FP64 1e-10 and FP32X 1e-5 relative to
# each integral, no floor; pure FP32 1e-5 relative to max(|I|, 1e-3 max|I|)


def _mock(G):
    # CoulombCalculatorTest_Interaction (run-tests.cpp:13-21): G = exp(-ix-iy-iz)
    Z1, Y1, X1 = G.shape
    iz, iy, ix = np.meshgrid(np.arange(Z1), np.arange(Y1), np.arange(X1), indexing="ij")
    G[...] = np.exp(-ix - iy - iz)


@pytest.mark.parametrize("prec,sched", COMBOS, ids=IDS)
def test_two_point_charges(prec, sched):
    """run-tests.cpp:28-48: 5x5x10 grid, delta at (2,2,2) -> 1.0; at (2,2,7) -> e^-5."""
    print("case two_point_charges", prec, sched, flush=True)
    cc = clat.CoulombCalculator((5, 5, 10), prec, sched)
    cc.initialize(clat.InteractionTable(_mock))
    tol = 1e-12 if prec == clat.PREC_FP64 else 1e-6
    cc.input[...] = 0
    cc.input[2, 2, 2] = 1.0
    cc.prepare()
    r = cc.calculate()
    assert abs(r.real - 1.0) < tol and abs(r.imag) < tol
    cc.input[...] = 0
    cc.input[7, 2, 2] = 1.0
    r = cc.calculate()
    assert abs(r.real - np.exp(-5)) < tol and abs(r.imag) < tol


@pytest.mark.parametrize("prec,sched", COMBOS, ids=IDS)
@pytest.mark.parametrize("dims", [(6, 5, 7), (16, 8, 4), (9, 10, 12), (11, 13, 5), (1, 4, 3), (32, 20, 24),
                                  (401, 3, 2)])
def test_prepare_matches_oracle(prec, sched, dims):
    """Random complex input on odd, mixed-radix, non-smooth (11, 13) and
    degenerate grids: potential and calculate() vs the 8-round oracle."""
    X, Y, Z = dims
    print("case prepare", dims, prec, sched, flush=True)
    rng = np.random.default_rng(X * 10000 + Y * 100 + Z)
    F = rng.standard_normal((Z, Y, X)) + 1j * rng.standard_normal((Z, Y, X))
    step = (1.1, 0.9, 1.3)
    Gt = O.gtable(X, Y, Z, step, 0.37, 3.0)
    Vo = O.prepare(F, O.ghat(Gt))
    cc = clat.CoulombCalculator(dims, prec, sched)
    cc.initialize(clat.InteractionSimple(step, 0.37, 3.0))
    cc.input[...] = F
    cc.prepare()
    V = cc.potential()
    err = np.max(np.abs(V - Vo)) / np.max(np.abs(Vo))
    assert err < TOL[prec] * 0.1, err
    r = cc.calculate()
    ro = O.calculate(Vo, F)
    assert abs(r - ro) / abs(ro) < TOL[prec], (r, ro)


@pytest.mark.parametrize("prec,sched", COMBOS, ids=IDS)
@pytest.mark.parametrize("dims", [(3, 32, 64), (2, 45, 96), (3, 120, 160), (2, 192, 256), (4, 256, 128),
                                  (2, 400, 8), (17, 64, 45), (8, 45, 96), (24, 120, 160), (16, 32, 400)])
def test_column_kernels_match_oracle(prec, sched, dims):
    """y and z lengths on every compile-time radix pair of the column passes
    (32, 45, 64, 96, 120, 128, 160, 192, 256, 400), with fewer columns than
    one block holds (X = 2, 3), a ragged last block (X = 17), and 2X a
    multiple of the column count (X = 8, 16, 24: the y-forward output is then
    stored in column strips, which the z pass reads as contiguous tiles)."""
    X, Y, Z = dims
    print("case column", dims, prec, sched, flush=True)
    rng = np.random.default_rng(X * 1000003 + Y * 1009 + Z)
    F = rng.standard_normal((Z, Y, X)) + 1j * rng.standard_normal((Z, Y, X))
    step = (0.8, 1.2, 1.0)
    Vo = O.prepare(F, O.ghat(O.gtable(X, Y, Z, step, 0.5, 2.0)))
    cc = clat.CoulombCalculator(dims, prec, sched)
    cc.initialize(clat.InteractionSimple(step, 0.5, 2.0))
    cc.input[...] = F
    cc.prepare()
    err = np.max(np.abs(cc.potential() - Vo)) / np.max(np.abs(Vo))
    assert err < TOL[prec] * 0.1, err


# every 2*3*5-smooth length from 36 to 360 whose two-level split was added to
# the compiled radix pairs beyond the ten above, twice per grid (y and z)
ADDED_YZ = [(2, 36, 40), (3, 48, 50), (2, 54, 60), (3, 72, 75), (2, 80, 81), (3, 90, 100), (2, 108, 135),
            (3, 144, 150), (2, 162, 180), (2, 200, 216), (2, 225, 240), (2, 270, 288), (2, 300, 320),
            (2, 324, 360)]
ADDED_X = [36, 40, 48, 50, 54, 60, 72, 75, 80, 81, 90, 100, 108, 135, 144, 150, 162, 180, 200, 216, 225, 240,
           270, 288, 300, 320, 324, 360]
SHARED = [(clat.PREC_FP64, clat.SCHED_SHARED), (clat.PREC_FP32X, clat.SCHED_SHARED),
          (clat.PREC_FP32, clat.SCHED_SHARED)]


@pytest.mark.parametrize("prec,sched", SHARED, ids=["fp64", "fp32x", "fp32"])
@pytest.mark.parametrize("dims", ADDED_YZ)
def test_added_column_pairs_match_oracle(prec, sched, dims):
    """y and z lengths on the radix pairs of the other 2*3*5-smooth lengths
    (6x6 ... 20x18: 36 to 360), against the 8-round oracle."""
    X, Y, Z = dims
    print("case added column", dims, prec, flush=True)
    rng = np.random.default_rng(X * 1000003 + Y * 1009 + Z)
    F = rng.standard_normal((Z, Y, X)) + 1j * rng.standard_normal((Z, Y, X))
    step = (0.8, 1.2, 1.0)
    Vo = O.prepare(F, O.ghat(O.gtable(X, Y, Z, step, 0.5, 2.0)))
    cc = clat.CoulombCalculator(dims, prec, sched)
    # no padding (the pairs are compiled) and the compile-time column kernels
    assert {"two_level", "column", "conv_tiled"} <= cc.kernels(), cc.kernels()
    cc.initialize(clat.InteractionSimple(step, 0.5, 2.0))
    cc.input[...] = F
    cc.prepare()
    err = np.max(np.abs(cc.potential() - Vo)) / np.max(np.abs(Vo))
    assert err < TOL[prec] * 0.1, err


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32X], ids=["fp64", "fp32x"])
@pytest.mark.parametrize("X", ADDED_X)
def test_added_line_pairs_match_oracle(prec, X):
    """Atom path (x line kernels over occupied lines, y passes on the
    36-point pair skipping empty rows) on x lengths of the added radix
    pairs."""
    Y, Z = 36, 5
    print("case added lines", X, prec, flush=True)
    rng = np.random.default_rng(X * 7919 + 17)
    lines = rng.choice(Y * Z, size=(Y * Z * 3) // 5, replace=False)
    cells = []
    for l in lines:
        xs = rng.choice(X, size=int(rng.integers(1, max(2, X // 3))), replace=False)
        cells += [int(l) * X + int(x) for x in xs]
    cells = np.array(cells, dtype=np.int32)
    psi = rng.standard_normal((len(cells), 2)) + 1j * rng.standard_normal((len(cells), 2))
    step = (1.0, 1.1, 0.9)
    eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
    assert eng.info()[0] == (X, Y, Z) and "line" in eng.kernels(), (eng.info(), eng.kernels())
    eng.initialize(clat.InteractionSimple(step, 0.3, 2.5))
    eng.set_atoms(cells, 2)
    eng.add_state(psi)
    v = eng.potential_at_atoms(1)
    eng.close()
    F = np.zeros(X * Y * Z, complex)
    F[cells] = np.sum(np.abs(psi) ** 2, axis=1)
    Vo = O.prepare(F.reshape(Z, Y, X), O.ghat(O.gtable(X, Y, Z, step, 0.3, 2.5))).reshape(-1)[cells]
    assert np.max(np.abs(v - Vo)) / np.max(np.abs(Vo)) < TOL[prec]


def _rel(a, b):
    return float(np.max(rel_err(a, b)))


def _example():
    d = os.path.join(os.path.dirname(__file__), "golden", "example")
    xyz = np.loadtxt(os.path.join(d, "X.3d"), usecols=(0, 1, 2))
    c = np.loadtxt(os.path.join(d, "e1.dat"), skiprows=1, usecols=(0, 1))
    return xyz, (c[:, 0] + 1j * c[:, 1]).reshape(len(xyz), 10)


@pytest.mark.parametrize("prec,sched", COMBOS, ids=IDS)
def test_example_1111(prec, sched):
    """example/run.sh: 8610-atom Si:P, 10 orbitals, --dielectric=11.4 ->
    ' 1  1  1  1    0.090340071977' (first 29 characters of eeee.txt)."""
    print("case example", prec, sched, flush=True)
    xyz, c = _example()
    raw, pad, origin, step = clat.lattice_grid(xyz)
    assert list(pad) == [45, 45, 45]
    cells = clat.lattice_cells(xyz, pad, origin, step)
    eng = clat.Engine(pad, prec, sched)
    eng.initialize(clat.InteractionSimple(step, 0.0, 11.4))
    eng.set_atoms(cells, 10)
    eng.add_state(c)
    v = eng.compute_integrals([(1, 1, 1, 1)])[0]
    line = "%2d %2d %2d %2d   %17.14f %17.14f\n" % (1, 1, 1, 1, v.real, v.imag)
    direct = 0.09034007197752  # FP64 direct lattice sum, pinned by the CPU tests
    assert abs(v.real - direct) / direct < TOL[prec]
    if prec == clat.PREC_FP64:
        assert line[:29] == " 1  1  1  1    0.090340071977"


@pytest.mark.parametrize("prec,sched", COMBOS, ids=IDS)
def test_zincblende_integrals(prec, sched):
    """2x2x2-cell zincblende (64 atoms, 8^3 grid), 2 holes + 2 electrons,
    20 orbitals, every integral (pattern ****) vs the oracle main loop."""
    print("case zincblende", prec, sched, flush=True)
    w = zb.workload(2, 2, 2, 2, 2, config=0, sigma_frac=0.2)
    X, Y, Z = w["dims"]
    quads = clat.pattern_enumerate("****", 4, 2)
    eng = clat.Engine((X, Y, Z), prec, sched)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    eng.set_atoms(w["cells"], 20)
    for s in w["states"]:
        eng.add_state(s)
    got = eng.compute_integrals(quads)
    rho = O.products(np.stack(w["states"]))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want = O.compute_integrals([tuple(q) for q in quads], 4, rho, w["cells"], (X, Y, Z), Gh)
    check(got, want, prec, quads)


@pytest.mark.parametrize("prec,sched", [COMBOS[0], COMBOS[1], COMBOS[4]], ids=[IDS[0], IDS[1], IDS[4]])
def test_stale_values(prec, sched):
    """CLAT_FLAG_STALE_VALUES (coulombo.cpp:194-205): the value is kept across
    a change of left product when the right product and the conj flag repeat.
    This plan has one such step: (1,4,3,4) follows (3,3,2,4) with right
    product 8 and conj flag 1 under a new left product, so it takes that
    value (conjugated, its left product being conjugate)."""
    print("case stale", prec, sched, flush=True)
    w = zb.workload(2, 2, 2, 2, 2, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    quads = [(2, 2, 1, 2), (3, 2, 4, 3), (1, 4, 4, 4), (1, 4, 3, 4), (3, 3, 2, 4), (4, 3, 2, 3),
             (3, 1, 1, 2), (3, 1, 2, 1)]
    eng = clat.Engine((X, Y, Z), prec, sched)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    eng.set_atoms(w["cells"], 4)
    for s in w["states"]:
        eng.add_state(s)
    got_stale = eng.compute_integrals(quads, stale_values=True)
    got = eng.compute_integrals(quads)
    rho = O.products(np.stack(w["states"]))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want_stale = O.compute_integrals(quads, 4, rho, w["cells"], (X, Y, Z), Gh, stale=True)
    want = O.compute_integrals(quads, 4, rho, w["cells"], (X, Y, Z), Gh)
    changed = np.abs(want_stale - want) > 1e-6 * np.max(np.abs(want))
    assert list(np.nonzero(changed)[0]) == [3]
    assert want_stale[3] == np.conj(want_stale[4])
    check(got_stale, want_stale, prec, quads)
    check(got, want, prec, quads)
    assert got_stale[3] == np.conj(got_stale[4])


@pytest.mark.parametrize("prec,sched", [COMBOS[0], COMBOS[1], COMBOS[4]], ids=[IDS[0], IDS[1], IDS[4]])
def test_atoms_sharing_cells(prec, sched):
    """Product.cpp:16-30 assigns each atom's density to its cell, so when
    atoms share a cell the last one in file order is the one that counts.
    Twelve extra atoms, interleaved at random positions, land in cells that
    are already occupied; integrals and per-atom potentials vs the oracle
    built from the last atom of every cell."""
    print("case shared cells", prec, sched, flush=True)
    w = zb.workload(2, 2, 2, 2, 2, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    rng = np.random.default_rng(7)
    n0 = len(w["cells"])
    extra = rng.choice(n0, 12, replace=False)
    # atom order: the originals with the extras inserted at random places
    order = list(range(n0))
    for k, src in enumerate(extra):
        order.insert(int(rng.integers(0, len(order) + 1)), n0 + k)
    cells_ext = np.concatenate([w["cells"], w["cells"][extra]])
    states_ext = [np.concatenate([s, rng.standard_normal((12, 4)) + 1j * rng.standard_normal((12, 4))])
                  for s in w["states"]]
    cells = cells_ext[order].astype(np.int32)
    states = [s[order] for s in states_ext]
    last = {}
    for a, c in enumerate(cells):
        last[int(c)] = a
    eff = np.array(sorted(last.values()))
    assert len(eff) == n0 and len(cells) == n0 + 12
    quads = clat.pattern_enumerate("****", 4, 2)
    eng = clat.Engine((X, Y, Z), prec, sched)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    eng.set_atoms(cells, 4)
    for s in states:
        eng.add_state(s)
    got = eng.compute_integrals(quads)
    st = np.stack(states)
    rho = O.products(st[:, eff, :])
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want = O.compute_integrals([tuple(q) for q in quads], 4, rho, cells[eff], (X, Y, Z), Gh)
    check(got, want, prec, quads)
    # the same run with every atom counted differs: the collisions matter here
    naive = O.compute_integrals([tuple(q) for q in quads], 4, O.products(st), cells, (X, Y, Z), Gh)
    assert _rel(naive, want) > 1e-3
    # potential of state 1's density at every atom, overwritten ones included
    v = eng.potential_at_atoms(1)
    F = np.zeros(X * Y * Z, complex)
    F[cells[eff]] = np.sum(np.abs(st[0, eff, :]) ** 2, axis=1)
    Vo = O.prepare(F.reshape(Z, Y, X), Gh).reshape(-1)[cells]
    assert np.max(np.abs(v - Vo)) / np.max(np.abs(Vo)) < TOL[prec]


def test_atoms_set_again():
    """A second clat_set_atoms drops the loaded states; with no atoms every
    integral is zero, also on an engine whose previous run had values."""
    print("case atoms set again", flush=True)
    w = zb.workload(2, 2, 2, 1, 1, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    eng = clat.Engine((X, Y, Z), clat.PREC_FP32X, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    eng.set_atoms(w["cells"], 4)
    for s in w["states"]:
        eng.add_state(s)
    quads = clat.pattern_enumerate("****", 2, 1)
    first = eng.compute_integrals(quads)
    assert np.max(np.abs(first)) > 0
    eng.set_atoms(np.zeros(0, np.int32), 4)
    with pytest.raises(clat.ClatError):
        eng.compute_integrals(quads)  # no states loaded
    for _ in range(2):
        eng.add_state(np.zeros((0, 4), complex))
    assert np.all(eng.compute_integrals(quads) == 0)
    # and back to the first atom list: the first values again
    eng.set_atoms(w["cells"], 4)
    for s in w["states"]:
        eng.add_state(s)
    assert np.array_equal(eng.compute_integrals(quads), first)


def test_empty_and_invalid_integral_lists():
    """No integrals (an empty list, or a pattern that selects none: the
    reference then writes no files) gives an empty result; a state index past
    the loaded states is an error, and the engine stays usable after it."""
    print("case empty lists", flush=True)
    w = zb.workload(2, 2, 2, 1, 1, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    eng = clat.Engine((X, Y, Z), clat.PREC_FP32X, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    eng.set_atoms(w["cells"], 4)
    for s in w["states"]:
        eng.add_state(s)
    assert eng.compute_integrals(np.zeros((0, 4), np.int32)).shape == (0,)
    none = clat.pattern_enumerate("9999", 2, 1)
    assert len(none) == 0 and eng.compute_integrals(none).shape == (0,)
    with pytest.raises(clat.ClatError):
        eng.compute_integrals([(1, 2, 3, 1)])
    got = eng.compute_integrals([(1, 2, 2, 1), (2, 1, 1, 2)])
    rho = O.products(np.stack(w["states"]))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want = O.compute_integrals([(1, 2, 2, 1), (2, 1, 1, 2)], 2, rho, w["cells"], (X, Y, Z), Gh)
    check(got, want, clat.PREC_FP32X)


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32, clat.PREC_FP32X], ids=["fp64", "fp32", "fp32x"])
def test_potential_at_atoms(prec):
    """potentials.cpp:140-158: V of each self product at every atom's cell."""
    print("case potentials", prec, flush=True)
    w = zb.workload(2, 2, 3, 1, 1, config=0, sigma_frac=0.2, n_orb=6)
    X, Y, Z = w["dims"]
    eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionSimple(w["step"], 0.5, 4.0))
    eng.set_atoms(w["cells"], 6)
    for s in w["states"]:
        eng.add_state(s)
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.5, 4.0))
    for k, s in enumerate(w["states"]):
        v = eng.potential_at_atoms(k + 1)
        F = np.zeros(X * Y * Z, complex)
        F[w["cells"]] = np.sum(np.abs(s) ** 2, axis=1)
        Vo = O.prepare(F.reshape(Z, Y, X), Gh).reshape(-1)[w["cells"]]
        assert np.max(np.abs(v - Vo)) / np.max(np.abs(Vo)) < TOL[prec]


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32, clat.PREC_FP32X], ids=["fp64", "fp32", "fp32x"])
@pytest.mark.parametrize("dims", [(32, 32, 5), (45, 36, 3), (64, 32, 7), (96, 36, 2), (120, 32, 3),
                                  (128, 36, 4), (160, 32, 3), (192, 36, 2), (256, 32, 3), (400, 32, 2)])
def test_site_lines_match_oracle(prec, dims):
    """Atom path on x lengths of the first ten compile-time radix pairs (y on
    a compiled pair too, so the x passes run on the line kernels): random
    sites on a random subset of x-lines (empty lines skipped by the x passes
    and never touched by the y passes), several atoms per cell (the last one
    holds the cell), an occupied-line count that leaves a ragged last block."""
    X, Y, Z = dims
    print("case site_lines", dims, prec, flush=True)
    rng = np.random.default_rng(X * 7919 + Y * 31 + Z)
    lines = rng.choice(Y * Z, size=max(1, (Y * Z * 3) // 5), replace=False)
    cells = []
    for l in lines:
        xs = rng.choice(X, size=int(rng.integers(1, max(2, X // 3))), replace=False)
        cells += [int(l) * X + int(x) for x in xs]
    cells += cells[: len(cells) // 7]  # repeated cells: the later atom holds the cell
    cells = np.array(cells, dtype=np.int32)
    n_orb = 2
    psi = rng.standard_normal((len(cells), n_orb)) + 1j * rng.standard_normal((len(cells), n_orb))
    step = (1.0, 1.1, 0.9)
    eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
    assert "line" in eng.kernels(), eng.kernels()
    eng.initialize(clat.InteractionSimple(step, 0.3, 2.5))
    eng.set_atoms(cells, n_orb)
    eng.add_state(psi)
    v = eng.potential_at_atoms(1)
    F = np.zeros(X * Y * Z, complex)
    F[cells] = np.sum(np.abs(psi) ** 2, axis=1)
    Vo = O.prepare(F.reshape(Z, Y, X), O.ghat(O.gtable(X, Y, Z, step, 0.3, 2.5))).reshape(-1)[cells]
    assert np.max(np.abs(v - Vo)) / np.max(np.abs(Vo)) < TOL[prec]


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32, clat.PREC_FP32X], ids=["fp64", "fp32", "fp32x"])
def test_batched_prepares(prec, monkeypatch):
    """Several left products per launch (up to eight): 5 states on a 32^3 grid
    give a cover of 15 left products, not a multiple of the batch; the values
    match the oracle main loop and are bit-identical to one prepare per launch."""
    print("case batched", prec, flush=True)
    w = zb.workload(8, 8, 8, 3, 2, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    quads = clat.pattern_enumerate("****", 5, 3)

    def run():
        eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
        eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
        eng.set_atoms(w["cells"], 4)
        for st in w["states"]:
            eng.add_state(st)
        return eng.compute_integrals(quads)

    got = run()
    monkeypatch.setenv("CLAT_BATCH", "1")
    one = run()
    assert np.array_equal(got, one)
    # the x-forward, y-forward and y-inverse outputs in column strips (the
    # default on this grid) or in plain rows: same arithmetic, other addresses
    monkeypatch.setenv("CLAT_ATILE", "0")
    rows_ad = run()
    assert np.array_equal(got, rows_ad)
    monkeypatch.setenv("CLAT_BTILE", "0")
    rows = run()
    assert np.array_equal(got, rows)
    monkeypatch.delenv("CLAT_BTILE")
    monkeypatch.delenv("CLAT_ATILE")
    # the five diagonal densities go two to a prepare (real and imaginary
    # part); one prepare each gives the same values to rounding
    monkeypatch.setenv("CLAT_PACK", "0")
    unpacked = run()
    assert _rel(got, unpacked) < (1e-12 if prec == clat.PREC_FP64 else 2e-6), _rel(got, unpacked)
    rho = O.products(np.stack(w["states"]))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want = O.compute_integrals([tuple(q) for q in quads], 5, rho, w["cells"], (X, Y, Z), Gh)
    check(got, want, prec, quads)
    # overlap sums on FP64 MFMA tiles (forced here; 15 slots x 15 products fill
    # one 64 x 64 tile partially): the same sums in another order
    monkeypatch.delenv("CLAT_PACK")
    monkeypatch.setenv("CLAT_OVL", "mfma")
    on_mfma = run()
    assert _rel(on_mfma, got) < 1e-12, _rel(on_mfma, got)
    check(on_mfma, want, prec, quads)
    # the default here is the 36 x 36 vector tile (three site groups per
    # block); the 48 x 48 tile sums the same terms in another order
    monkeypatch.setenv("CLAT_OVL", "valu48")
    on48 = run()
    assert _rel(on48, got) < 1e-12, _rel(on48, got)
    # the same tiles with every site sum Kahan-compensated
    monkeypatch.setenv("CLAT_OVL", "kahan")
    onk = run()
    assert _rel(onk, got) < 1e-12, _rel(onk, got)
    check(onk, want, prec, quads)


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32X], ids=["fp64", "fp32x"])
def test_overlap_tiles_by_size(prec, monkeypatch):
    """14 states: 105 left slots and 105 products, where the overlap sums run
    on FP64 MFMA tiles (two 64 x 64 tiles per side, the second partial).
    Every product being a slot, the table is symmetric (sum V_a rho_b =
    sum rho_a V_b for the real, even kernel) and only the three tiles on and
    above the diagonal are formed, the fourth read from its mirror.  All four
    tiles (CLAT_OVL_SYM=0) and the vector-FMA kernel (three 48 x 48 tiles per
    side) sum the same products: the same integrals to FP64 rounding.  The
    mirrored entries sum the other product of each pair: in FP64 the same to
    rounding; in FP32X another result of the same complex64 accuracy, no
    less accurate against FP64 than the full table, in relative and in
    absolute terms (on this 16^3 grid both reach 1.5e-5 relative on
    integrals near 1e-5 of the largest).  The kernel that ran is told by the
    overlap's staged bytes, tiles x 2 x edge x sites x 16."""
    print("case overlap_tiles", prec, flush=True)
    w = zb.workload(4, 4, 4, 7, 7, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    quads = clat.pattern_enumerate("****", 14, 7)

    def run(p):
        eng = clat.Engine((X, Y, Z), p, clat.SCHED_SHARED)
        eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
        eng.set_atoms(w["cells"], 4)
        for st in w["states"]:
            eng.add_state(st)
        eng.set_kernel_timing(True)
        vals = eng.compute_integrals(quads)
        return vals, eng.kernel_stats()["overlap"][2]

    got, staged = run(prec)
    monkeypatch.setenv("CLAT_OVL_SYM", "0")
    full, staged_full = run(prec)
    monkeypatch.delenv("CLAT_OVL_SYM")
    monkeypatch.setenv("CLAT_OVL", "valu")
    ref, staged_ref = run(prec)
    monkeypatch.delenv("CLAT_OVL")
    assert staged > 0 and staged * 4 == staged_full * 3, (staged, staged_full)
    assert staged_full * (9 * 96) == staged_ref * (4 * 128), (staged_full, staged_ref)
    assert _rel(full, ref) < 1e-12, _rel(full, ref)
    if prec == clat.PREC_FP64:
        assert _rel(got, ref) < 1e-12, _rel(got, ref)
    else:
        ref64, _ = run(clat.PREC_FP64)
        e_sym, e_full = _rel(got, ref64), _rel(full, ref64)
        a_sym, a_full = np.max(np.abs(got - ref64)), np.max(np.abs(full - ref64))
        print("fp32x vs fp64: symmetric %.3e (abs %.2e), all tiles %.3e (abs %.2e)" % (e_sym, a_sym, e_full, a_full))
        assert e_sym <= 1.25 * e_full and a_sym <= 1.25 * a_full, (e_sym, e_full, a_sym, a_full)


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32, clat.PREC_FP32X], ids=["fp64", "fp32", "fp32x"])
def test_thomas_fermi(prec):
    """InteractionThomasFermi (Interaction.cpp:18-39) table vs the oracle's."""
    print("case thomas_fermi", prec, flush=True)
    w = zb.workload(2, 2, 2, 1, 1, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    quads = clat.pattern_enumerate("****", 2, 1)
    eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionThomasFermi(w["step"], 0.0, 10.89, 5.6533))
    eng.set_atoms(w["cells"], 4)
    for s in w["states"]:
        eng.add_state(s)
    got = eng.compute_integrals(quads)
    rho = O.products(np.stack(w["states"]))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89, tf_lattice=5.6533))
    want = O.compute_integrals([tuple(q) for q in quads], 2, rho, w["cells"], (X, Y, Z), Gh)
    check(got, want, prec, quads)


@pytest.mark.parametrize("prec,sched", COMBOS, ids=IDS)
def test_golden_direct_sums(prec, sched, golden_dir):
    """Engine vs the committed direct O(N^2) lattice-sum fixtures (no FFT in
    the checker), including onsite != 0 and the 4096-atom 32^3 case."""
    import json
    with open(os.path.join(golden_dir, "zincblende_direct.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases:
        print("case golden", case["name"], prec, sched, flush=True)
        w = zb.workload(*case["cells"], case["holes"], case["electrons"], config=case["config"],
                        sigma_frac=case["sigma_frac"], n_orb=case["orbitals"])
        X, Y, Z = w["dims"]
        eng = clat.Engine((X, Y, Z), prec, sched)
        eng.initialize(clat.InteractionSimple(w["step"], case["onsite"], case["dielectric"]))
        eng.set_atoms(w["cells"], case["orbitals"])
        for s in w["states"]:
            eng.add_state(s)
        got = eng.compute_integrals(case["quads"])
        want = np.array([complex(a, b) for a, b in case["values"]])
        check(got, want, prec, case["quads"])
        eng.close()
