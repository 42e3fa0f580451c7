"""The batched and multi-system workloads at their full sizes, checked through
properties that hold at any size, plus the oracle where a grid is small
enough for it:

* batched form: zincblende 32^3 cells = 262,144 atoms on a 128^3 grid,
  32 states (16 holes, 16 electrons), 20 orbitals, the full ijkl set
  (1,048,576 integrals, 528 pair densities, all 528 prepared);
* the 1M-atom workload: 64x64x32 cells on a 256x256x128 grid, 8 states,
  eeee,hhhh,ehhe,eheh -- all 1024 integrals against the oracle;
* multi-system: boxes of 16^3 ... 100^3 cells (grids 64^3 ... 400^3, the
  last 8M atoms), 8 states (4 holes, 4 electrons) and 20 orbitals each,
  eeee,hhhh,ehhe,eheh -- every integral against the oracle up to 192^3.

In the batched form eight integrals are also checked against the oracle's
eight-round prepare on the full 128^3 grid.  Every comparison is per
integral and relative to that integral (tests/parity.py), with no floor;
on the batched and multi-system workloads the FP32X path is held to 1e-5 of
each integral down to 1e-6 of the largest, and below that to the bound that
rounding the pair densities to complex64 sets (1e-11 of the largest), with
every integral past 1e-5 printed by name.
Properties: the kernel G is real and even, so
  I(i1,i2,i3,i4) = I(i2,i1,i4,i3)          (both densities swapped)
  I(i1,i2,i3,i4) = conj(I(i4,i3,i2,i1))    (Hermitian)
  I(i,j,j,i) real and >= 0                  (a Coulomb self-energy)
and the complex64 path agrees with the FP64 path."""
import numpy as np
import pytest

import clat
import oracle as O
import zincblende as zb
from parity import check, check_fp32x_model, rel_err, worst

pytestmark = pytest.mark.gpu


def _engine(w, prec, n_orb):
    X, Y, Z = w["dims"]
    eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    eng.set_atoms(w["cells"], n_orb)
    for s in w["states"]:
        eng.add_state(s)
    return eng


def _index(quads):
    return {tuple(int(v) for v in q): k for k, q in enumerate(quads)}


def _check_symmetries(quads, vals, tol):
    idx = _index(quads)
    scale = np.max(np.abs(vals))
    worst = 0.0
    for k, (a, b, c, d) in enumerate(quads):
        for partner, conj in (((b, a, d, c), False), ((d, c, b, a), True)):
            j = idx.get(partner)
            if j is None:
                continue
            other = np.conj(vals[j]) if conj else vals[j]
            worst = max(worst, abs(vals[k] - other) / scale)
        if a == d and b == c:
            assert abs(vals[k].imag) <= tol * scale and vals[k].real > -tol * scale
    assert worst <= tol, worst


def _report(label, got, want, quads):
    """Print the per-integral error distribution (by decade of |I|/max|I|)."""
    e = rel_err(got, want)
    mag = np.abs(want) / np.max(np.abs(want))
    print("%s: %d integrals, max rel err %.3e, median %.3e" % (label, len(e), np.max(e), np.median(e)))
    for lo in range(0, -12, -2):
        sel = (mag <= 10.0 ** lo) & (mag > 10.0 ** (lo - 2))
        if sel.any():
            print("  |I|/max in (1e%d, 1e%d]: %d integrals, max rel err %.3e" % (lo - 2, lo, sel.sum(), np.max(e[sel])))
    print(worst(got, want, quads, 3), flush=True)


def _oracle_rho(states, pick):
    """The oracle's pair densities that the picked quads use, keyed by the
    product index p = L(L-1)/2 + R - 1 (Planner.hpp:134-144)."""
    rho = {}
    for (a, b, c, d) in pick:
        for (i, j) in ((a, d), (b, c)):
            hi, lo = max(i, j), min(i, j)
            pair = np.stack([states[hi - 1]]) if hi == lo else np.stack([states[lo - 1], states[hi - 1]])
            rho[hi * (hi - 1) // 2 + lo - 1] = O.products(pair)[0 if hi == lo else 1]
    return rho


def test_batched_form_full_size():
    print("case config batched 128^3, 32 states, ijkl", flush=True)
    w = zb.workload(32, 32, 32, 16, 16, config=3, sigma_frac=0.15)
    quads = clat.pattern_enumerate("ijkl", 32, 16)
    assert len(quads) == 32 ** 4
    vals = {}
    for prec in (clat.PREC_FP32X, clat.PREC_FP64):
        eng = _engine(w, prec, 20)
        vals[prec] = eng.compute_integrals(quads)
        eng.close()
    # the symmetry partners go through different products and conj paths
    _check_symmetries(quads, vals[clat.PREC_FP64], 1e-12)
    # every one of the 1,048,576 FP32X integrals against FP64: within 1e-5 of
    # itself down to 1e-6 of the largest; below, the complex64 density
    # rounding bound, each integral past 1e-5 listed by name
    _report("batched FP32X vs FP64", vals[clat.PREC_FP32X], vals[clat.PREC_FP64], quads)
    check_fp32x_model(vals[clat.PREC_FP32X], vals[clat.PREC_FP64], quads, "batched")
    # eight integrals against the oracle's eight-round prepare on the same
    # 128^3 grid, given only the pair densities they use
    X, Y, Z = w["dims"]
    pick = [(1, 1, 1, 1), (1, 2, 3, 4), (20, 3, 9, 31), (32, 32, 32, 32), (17, 5, 5, 17), (16, 17, 18, 19),
            (2, 31, 7, 30), (25, 25, 1, 1)]
    rho = _oracle_rho(w["states"], pick)
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want = O.compute_integrals(pick, 32, rho, w["cells"], (X, Y, Z), Gh)
    at = _index(quads)
    check(np.array([vals[clat.PREC_FP64][at[q]] for q in pick]), want, clat.PREC_FP64, pick)
    check(np.array([vals[clat.PREC_FP32X][at[q]] for q in pick]), want, clat.PREC_FP32X, pick)


def test_one_million_atoms_full_size():
    """The 1M-atom workload as the bench runs it: 64x64x32 cells = 1,048,576
    atoms on a 256x256x128 grid, 8 states (4 holes, 4 electrons), 20 orbitals,
    eeee,hhhh,ehhe,eheh (1024 integrals).  FP64 symmetries, and every one of
    the 1024 integrals of both paths against the oracle's eight-round
    prepare on the full grid (FP64 within 1e-10, FP32X within 1e-5, each
    relative to the integral itself)."""
    print("case config 1M atoms 256x256x128, 8 states", flush=True)
    w = zb.workload(64, 64, 32, 4, 4, config=1, sigma_frac=0.15)
    X, Y, Z = w["dims"]
    assert (X, Y, Z) == (256, 256, 128) and len(w["cells"]) == 1048576
    quads = clat.pattern_enumerate("eeee,hhhh,ehhe,eheh", 8, 4)
    assert len(quads) == 1024
    vals = {}
    for prec in (clat.PREC_FP32X, clat.PREC_FP64):
        eng = _engine(w, prec, 20)
        # the path the bench measures: two cosets, half box, half-length x lines
        assert {"cosets", "half_box", "half_lines"} <= eng.kernels()
        vals[prec] = eng.compute_integrals(quads)
        eng.close()
    _check_symmetries(quads, vals[clat.PREC_FP64], 1e-12)
    _report("1M atoms FP32X vs FP64", vals[clat.PREC_FP32X], vals[clat.PREC_FP64], quads)
    check(vals[clat.PREC_FP32X], vals[clat.PREC_FP64], clat.PREC_FP32X, quads)
    print("oracle: 36 prepares on 256x256x128", flush=True)
    rho = O.products(np.stack(w["states"]))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want = O.compute_integrals([tuple(q) for q in quads], 8, rho, w["cells"], (X, Y, Z), Gh)
    for prec, name in ((clat.PREC_FP64, "FP64"), (clat.PREC_FP32X, "FP32X")):
        _report("1M atoms %s vs oracle" % name, vals[prec], want, quads)
        check(vals[prec], want, prec, quads)


@pytest.mark.parametrize("cells", [(16, 16, 16), (24, 24, 24), (30, 30, 30), (40, 40, 40), (48, 48, 48),
                                   (100, 100, 100)])
def test_multi_system_sizes(cells):
    """The multi-system boxes at their workload: 4 holes + 4 electrons, 20
    orbitals, eeee,hhhh,ehhe,eheh (1024 integrals per box).  FP64 symmetries,
    every FP32X integral against FP64 under the complex64 model (1e-5 of
    itself down to 1e-6 of the largest, 1e-11 of the largest below, those
    past 1e-5 listed), and up to the 192^3 grid every integral of both paths
    against the oracle (the 400^3 grid, 8M atoms, would take the oracle
    minutes per prepare)."""
    print("case config multi-system", cells, flush=True)
    n_orb = 20
    w = zb.workload(*cells, 4, 4, config=4, sigma_frac=0.15, n_orb=n_orb)
    X, Y, Z = w["dims"]
    assert (X, Y, Z) == tuple(4 * c for c in cells)
    quads = clat.pattern_enumerate("eeee,hhhh,ehhe,eheh", 8, 4)
    vals = {}
    for prec in (clat.PREC_FP32X, clat.PREC_FP64):
        eng = _engine(w, prec, n_orb)
        vals[prec] = eng.compute_integrals(quads)
        eng.close()
    _check_symmetries(quads, vals[clat.PREC_FP64], 1e-12)
    _report("%d^3 FP32X vs FP64" % X, vals[clat.PREC_FP32X], vals[clat.PREC_FP64], quads)
    check_fp32x_model(vals[clat.PREC_FP32X], vals[clat.PREC_FP64], quads, "%d^3" % X)
    if X * Y * Z <= 192 ** 3:
        print("oracle: 36 prepares on %dx%dx%d" % (X, Y, Z), flush=True)
        rho = O.products(np.stack(w["states"]))
        Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
        want = O.compute_integrals([tuple(q) for q in quads], 8, rho, w["cells"], (X, Y, Z), Gh)
        check(vals[clat.PREC_FP64], want, clat.PREC_FP64, quads)
        check_fp32x_model(vals[clat.PREC_FP32X], want, quads, "%d^3 vs oracle" % X)


def test_multi_system_concurrent():
    """Engines of different grid sizes run from separate host threads (one
    stream each) give exactly the values they give one after another."""
    print("case config multi-system concurrent 16^3 + 24^3 + 30^3", flush=True)
    jobs = []
    for n in (16, 24, 30):
        w = zb.workload(n, n, n, 2, 2, config=4, sigma_frac=0.15, n_orb=4)
        eng = _engine(w, clat.PREC_FP32X, 4)
        jobs.append((eng, clat.pattern_enumerate("****", 4, 2)))
    seq = [e.compute_integrals(q) for e, q in jobs]
    for _ in range(3):
        con = clat.run_concurrent([(lambda e=e, q=q: e.compute_integrals(q)) for e, q in jobs])
        for a, b in zip(seq, con):
            assert np.array_equal(a, b)
    for e, _ in jobs:
        e.close()
