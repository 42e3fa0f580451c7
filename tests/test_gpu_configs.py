"""The batched and multi-system workloads at their full sizes, checked through
properties that hold at any size, plus the oracle where a grid is small
enough for it:

* batched form: zincblende 32^3 cells = 262,144 atoms on a 128^3 grid,
  32 states (16 holes, 16 electrons), 20 orbitals, the full ijkl set
  (1,048,576 integrals, 528 pair densities, all 528 prepared);
* multi-system: boxes of 16^3 ... 100^3 cells (grids 64^3 ... 400^3, the
  last 8M atoms), 2 states each.

In the batched form three integrals are also checked against the oracle's
eight-round prepare on the full 128^3 grid.
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


def _rel_err(got, want):
    scale = np.max(np.abs(want))
    return np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-3 * scale))


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
    sub = np.arange(0, len(quads), 97)
    _check_symmetries(quads, vals[clat.PREC_FP64], 1e-12)
    assert _rel_err(vals[clat.PREC_FP32X], vals[clat.PREC_FP64]) < 1e-5
    assert np.all(np.isfinite(vals[clat.PREC_FP32X][sub]))
    # three integrals against the oracle's eight-round prepare on the same
    # 128^3 grid, given only the pair densities they use (p = L(L-1)/2 + R - 1)
    X, Y, Z = w["dims"]
    st = w["states"]
    pick = [(1, 1, 1, 1), (1, 2, 3, 4), (20, 3, 9, 31)]
    rho = {}
    for (a, b, c, d) in pick:
        for (i, j) in ((a, d), (b, c)):
            hi, lo = max(i, j), min(i, j)
            pair = np.stack([st[hi - 1]]) if hi == lo else np.stack([st[lo - 1], st[hi - 1]])
            rho[hi * (hi - 1) // 2 + lo - 1] = O.products(pair)[0 if hi == lo else 1]
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want = O.compute_integrals(pick, 32, rho, w["cells"], (X, Y, Z), Gh)
    at = _index(quads)
    for prec, tol in ((clat.PREC_FP64, 1e-10), (clat.PREC_FP32X, 1e-5)):
        got = np.array([vals[prec][at[q]] for q in pick])
        assert _rel_err(got, want) < tol, (prec, got, want)


def test_one_million_atoms_full_size():
    """The 1M-atom workload as the bench runs it: 64x64x32 cells = 1,048,576
    atoms on a 256x256x128 grid, 8 states (4 holes, 4 electrons), 20 orbitals,
    eeee,hhhh,ehhe,eheh (1024 integrals).  FP64 symmetries, FP32X within 1e-5
    of FP64, and one integral of each kind against the oracle's eight-round
    prepare on the full grid."""
    print("case config 1M atoms 256x256x128, 8 states", flush=True)
    w = zb.workload(64, 64, 32, 4, 4, config=1, sigma_frac=0.15)
    X, Y, Z = w["dims"]
    assert (X, Y, Z) == (256, 256, 128) and len(w["cells"]) == 1048576
    quads = clat.pattern_enumerate("eeee,hhhh,ehhe,eheh", 8, 4)
    assert len(quads) == 1024
    vals = {}
    for prec in (clat.PREC_FP32X, clat.PREC_FP64):
        eng = _engine(w, prec, 20)
        vals[prec] = eng.compute_integrals(quads)
        eng.close()
    _check_symmetries(quads, vals[clat.PREC_FP64], 1e-12)
    assert _rel_err(vals[clat.PREC_FP32X], vals[clat.PREC_FP64]) < 1e-5
    st = w["states"]
    pick = [(5, 5, 5, 5), (1, 2, 3, 4), (6, 2, 2, 7), (6, 2, 7, 3)]
    at = _index(quads)
    assert all(q in at for q in pick)
    rho = {}
    for (a, b, c, d) in pick:
        for (i, j) in ((a, d), (b, c)):
            hi, lo = max(i, j), min(i, j)
            pair = np.stack([st[hi - 1]]) if hi == lo else np.stack([st[lo - 1], st[hi - 1]])
            rho[hi * (hi - 1) // 2 + lo - 1] = O.products(pair)[0 if hi == lo else 1]
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    want = O.compute_integrals(pick, 8, rho, w["cells"], (X, Y, Z), Gh)
    for prec, tol in ((clat.PREC_FP64, 1e-10), (clat.PREC_FP32X, 1e-5)):
        got = np.array([vals[prec][at[q]] for q in pick])
        assert _rel_err(got, want) < tol, (prec, got, want)


@pytest.mark.parametrize("cells", [(16, 16, 16), (24, 24, 24), (30, 30, 30), (40, 40, 40), (48, 48, 48),
                                   (100, 100, 100)])
def test_multi_system_sizes(cells):
    print("case config multi-system", cells, flush=True)
    n_orb = 4
    w = zb.workload(*cells, 1, 1, config=4, sigma_frac=0.15, n_orb=n_orb)
    X, Y, Z = w["dims"]
    assert (X, Y, Z) == tuple(4 * c for c in cells)
    quads = clat.pattern_enumerate("****", 2, 1)
    vals = {}
    for prec in (clat.PREC_FP32X, clat.PREC_FP64):
        eng = _engine(w, prec, n_orb)
        vals[prec] = eng.compute_integrals(quads)
        eng.close()
    _check_symmetries(quads, vals[clat.PREC_FP64], 1e-12)
    assert _rel_err(vals[clat.PREC_FP32X], vals[clat.PREC_FP64]) < 1e-5
    if X * Y * Z <= 192 ** 3:
        rho = O.products(np.stack(w["states"]))
        Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
        want = O.compute_integrals([tuple(q) for q in quads], 2, rho, w["cells"], (X, Y, Z), Gh)
        assert _rel_err(vals[clat.PREC_FP64], want) < 1e-10


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
