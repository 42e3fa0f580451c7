"""GPU parity of the two-coset path (sites on x = y = z (mod 2) of an even
grid: zincblende on its a/4 grid) against the FP64 oracle and against the
engine's single-grid path on the same inputs.  Every call goes through the C
ABI; CLAT_COSETS=0 at clat_set_atoms keeps an engine on the single grid."""
import numpy as np
import pytest

import clat
import oracle as O
import zincblende as zb
from parity import check, check_fp32x_model, rel_err

pytestmark = pytest.mark.gpu

PRECS = [clat.PREC_FP64, clat.PREC_FP32X, clat.PREC_FP32]
PIDS = ["fp64", "fp32x", "fp32"]


def _engine(w, cells, prec, n_orb, states, cosets=True, monkeypatch=None):
    if monkeypatch is not None:
        if cosets:
            monkeypatch.delenv("CLAT_COSETS", raising=False)
        else:
            monkeypatch.setenv("CLAT_COSETS", "0")
    X, Y, Z = w["dims"]
    eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    eng.set_atoms(cells, n_orb)
    for c in states:
        eng.add_state(c)
    return eng


def _oracle(w, cells, states, quads, n_states):
    X, Y, Z = w["dims"]
    rho = O.products(np.stack(states))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    return O.compute_integrals([tuple(q) for q in quads], n_states, rho, cells, (X, Y, Z), Gh)


@pytest.mark.parametrize("prec", PRECS, ids=PIDS)
@pytest.mark.parametrize("box", [(16, 16, 16), (24, 20, 16)])
def test_cosets_match_oracle_and_single_grid(prec, box, monkeypatch):
    """64^3 and 96x80x64 zincblende grids (coarse 32^3 and 48x40x32, pairs
    8x4, 8x6, 8x5): integrals and per-atom potentials on the two-coset path
    against the oracle and against the single-grid path."""
    nh, ne = 2, 2
    w = zb.workload(*box, nh, ne, config=0, sigma_frac=0.2, n_orb=4)
    quads = clat.pattern_enumerate("****", nh + ne, nh)
    print("case cosets", box, prec, w["dims"], flush=True)
    a = _engine(w, w["cells"], prec, 4, w["states"], True, monkeypatch)
    b = _engine(w, w["cells"], prec, 4, w["states"], False, monkeypatch)
    assert {"cosets", "half_box"} <= a.kernels() and "cosets" not in b.kernels()
    ga, gb = a.compute_integrals(quads), b.compute_integrals(quads)
    pa, pb = a.potential_at_atoms(2), b.potential_at_atoms(2)
    a.close()
    b.close()
    want = _oracle(w, w["cells"], w["states"], quads, nh + ne)
    if prec == clat.PREC_FP64:
        check(ga, want, prec, quads)
        assert np.max(rel_err(ga, gb)) < 1e-12
        assert np.max(np.abs(pa - pb)) < 1e-12 * np.max(np.abs(pb))
    elif prec == clat.PREC_FP32X:
        check_fp32x_model(ga, want, quads, "cosets fp32x")
        assert np.max(np.abs(pa - pb)) < 1e-6 * np.max(np.abs(pb))
    else:
        check(ga, want, prec, quads)


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32X], ids=["fp64", "fp32x"])
def test_cosets_partial_occupancy(prec, monkeypatch):
    """Atoms inside an off-centre sphere, then every tenth removed: the two
    cosets occupy different coarse cells, so each holds positions without an
    atom (zero density there).  Integrals and potentials against the oracle."""
    nh, ne = 2, 1
    full = zb.workload(16, 16, 16, nh, ne, config=3, sigma_frac=0.2, n_orb=4)
    X, Y, Z = full["dims"]
    idx = full["idx"]
    c = np.array([0.45 * X, 0.55 * Y, 0.5 * Z])
    keep = np.sum((idx - c) ** 2, axis=1) < (0.33 * X) ** 2
    keep &= (np.arange(len(idx)) % 10) != 3
    cells = full["cells"][keep]
    states = [s[keep] for s in full["states"]]
    coset = (idx[keep, 0] % 2).astype(bool)
    u = (idx[keep] // 2)
    ucode = (u[:, 2] * (Y // 2) + u[:, 1]) * (X // 2) + u[:, 0]
    assert set(ucode[coset]) != set(ucode[~coset])  # positions occupied in one coset only
    quads = clat.pattern_enumerate("****", nh + ne, nh)
    print("case cosets partial", prec, len(cells), flush=True)
    a = _engine(full, cells, prec, 4, states, True, monkeypatch)
    assert "cosets" in a.kernels()
    got = a.compute_integrals(quads)
    pot = a.potential_at_atoms(1)
    pot_na = a.potential_at_atoms_noalloc(1)
    a.close()
    b = _engine(full, cells, prec, 4, states, False, monkeypatch)
    pot_b = b.potential_at_atoms(1)
    b.close()
    want = _oracle(full, cells, states, quads, nh + ne)
    if prec == clat.PREC_FP64:
        check(got, want, prec, quads)
        assert np.max(np.abs(pot - pot_b)) < 1e-12 * np.max(np.abs(pot_b))
    else:
        check_fp32x_model(got, want, quads, "cosets partial fp32x")
        assert np.max(np.abs(pot - pot_b)) < 1e-6 * np.max(np.abs(pot_b))
    assert np.array_equal(pot, pot_na)


def test_cosets_off_sublattice_falls_back(monkeypatch):
    """One atom moved off x = y = z (mod 2): the engine stays on the single
    grid (no 'cosets' path) and matches the oracle."""
    nh, ne = 1, 1
    w = zb.workload(16, 16, 16, nh, ne, config=4, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    idx = w["idx"].copy()
    free = set(w["cells"].tolist())
    k = len(idx) // 2
    moved = idx[k] + np.array([1, 0, 0])
    assert (moved[0] % 2) != (moved[1] % 2)
    idx[k] = moved
    cells = zb.cells_of(idx, (X, Y, Z))
    assert int(cells[k]) not in free
    quads = clat.pattern_enumerate("****", nh + ne, nh)
    eng = _engine(w, cells, clat.PREC_FP64, 4, w["states"], True, monkeypatch)
    assert "cosets" not in eng.kernels() and "line" in eng.kernels()
    got = eng.compute_integrals(quads)
    eng.close()
    check(got, _oracle(w, cells, w["states"], quads, nh + ne), clat.PREC_FP64, quads)


def test_cosets_batches_bit_identical(monkeypatch):
    """The coset planes of 1, 2, 4 and 8 prepares per launch (the measured
    batch) give bit-identical integrals, directly and replayed from graphs."""
    nh, ne = 3, 3
    w = zb.workload(16, 16, 16, nh, ne, config=5, sigma_frac=0.2, n_orb=4)
    quads = clat.pattern_enumerate("****", nh + ne, nh)
    ref = None
    for batch in ("1", "2", "4", "8", None):
        if batch is None:
            monkeypatch.delenv("CLAT_BATCH", raising=False)
        else:
            monkeypatch.setenv("CLAT_BATCH", batch)
        eng = _engine(w, w["cells"], clat.PREC_FP32X, 4, w["states"], True, monkeypatch)
        assert "cosets" in eng.kernels()
        runs = [eng.compute_integrals(quads) for _ in range(3)]
        eng.close()
        for r in runs[1:]:
            assert np.array_equal(r, runs[0]), batch
        if ref is None:
            ref = runs[0]
        assert np.array_equal(runs[0], ref), batch


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32X], ids=["fp64", "fp32x"])
def test_coset_table_routes_agree(prec, monkeypatch):
    """The coarse tables built from the transformed fine table (the dense
    protocol ran first, so the fine DCT-I exists when the atoms arrive: the
    eight half-box shifts) and from the real-space samples (an interaction
    set with the atoms in place, built at the next run: coarse DCT-I and
    DCT-II) give the same integrals; the dense protocol's value matches the
    oracle."""
    nh, ne = 2, 1
    w = zb.workload(16, 16, 16, nh, ne, config=6, sigma_frac=0.2, n_orb=4)
    quads = clat.pattern_enumerate("****", nh + ne, nh)
    monkeypatch.delenv("CLAT_COSETS", raising=False)
    X, Y, Z = w["dims"]
    eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    F = np.zeros((Z, Y, X), np.complex128)
    F[3, 5, 7], F[40, 22, 9], F[61, 60, 2] = 1.0 + 0.5j, -0.7, 0.25j
    clat._check(clat.lib().clat_set_input(eng.handle, F.view(np.float64).reshape(-1)))
    clat._check(clat.lib().clat_prepare(eng.handle))
    out = np.zeros(2)
    clat._check(clat.lib().clat_calculate(eng.handle, out))
    eng.set_atoms(w["cells"], 4)
    for c in w["states"]:
        eng.add_state(c)
    assert "cosets" in eng.kernels()
    shifted = eng.compute_integrals(quads)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    sampled = eng.compute_integrals(quads)
    eng.close()
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    dense_want = O.calculate(O.prepare(F, Gh), F)
    assert abs(complex(out[0], out[1]) - dense_want) < (1e-12 if prec == clat.PREC_FP64 else 1e-5) * abs(dense_want)
    want = _oracle(w, w["cells"], w["states"], quads, nh + ne)
    if prec == clat.PREC_FP64:
        assert np.max(rel_err(sampled, shifted)) < 1e-12
        check(sampled, want, prec, quads)
        check(shifted, want, prec, quads)
    else:
        assert np.max(rel_err(sampled, shifted)) < 1e-5
        check_fp32x_model(sampled, want, quads, "sampled tables fp32x")
        check_fp32x_model(shifted, want, quads, "shifted tables fp32x")


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32X], ids=["fp64", "fp32x"])
def test_cosets_full_box(prec, monkeypatch):
    """Both cosets, all 2 Xc x frequencies: an extra atom at fine (2, 0, 0)
    (coset 0, coarse sum odd) keeps the engine off the half box, and
    CLAT_HALFBOX=0 does the same for plain zincblende.  Integrals against the
    oracle and, for zincblende, against the half-box form."""
    nh, ne = 2, 1
    w = zb.workload(16, 16, 16, nh, ne, config=7, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    quads = clat.pattern_enumerate("****", nh + ne, nh)
    idx = np.concatenate([w["idx"], [[2, 0, 0]]])
    assert int(zb.cells_of(idx[-1:], (X, Y, Z))[0]) not in set(w["cells"].tolist())
    cells = zb.cells_of(idx, (X, Y, Z))
    rng = np.random.default_rng(11)
    states = [np.concatenate([s, (rng.standard_normal((1, 4)) + 1j * rng.standard_normal((1, 4))) * 0.05])
              for s in w["states"]]
    eng = _engine(w, cells, prec, 4, states, True, monkeypatch)
    assert "cosets" in eng.kernels() and "half_box" not in eng.kernels()
    got = eng.compute_integrals(quads)
    eng.close()
    want = _oracle(w, cells, states, quads, nh + ne)
    if prec == clat.PREC_FP64:
        check(got, want, prec, quads)
    else:
        check_fp32x_model(got, want, quads, "full box fp32x")
    monkeypatch.setenv("CLAT_HALFBOX", "0")
    full = _engine(w, w["cells"], prec, 4, w["states"], True, None)
    assert "cosets" in full.kernels() and "half_box" not in full.kernels()
    gf = full.compute_integrals(quads)
    full.close()
    monkeypatch.delenv("CLAT_HALFBOX")
    half = _engine(w, w["cells"], prec, 4, w["states"], True, None)
    assert "half_box" in half.kernels()
    gh = half.compute_integrals(quads)
    half.close()
    want = _oracle(w, w["cells"], w["states"], quads, nh + ne)
    if prec == clat.PREC_FP64:
        check(gf, want, prec, quads)
        assert np.max(rel_err(gh, gf)) < 1e-12
    else:
        check_fp32x_model(gf, want, quads, "full box fp32x")
        check_fp32x_model(gh, want, quads, "half box fp32x")


@pytest.mark.parametrize("prec", PRECS, ids=PIDS)
@pytest.mark.parametrize("box", [(32, 16, 16), (40, 16, 16)])
def test_half_lines_match_oracle(prec, box, monkeypatch):
    """128x64x64 and 160x64x64 zincblende grids: coarse lines of 64 and 80
    points, whose sites all have x = y + z (mod 2), carried as 32 and 40
    points (pairs 8x4 and 8x5) with the x parity as a phase.  Integrals
    against the oracle and against full coarse lines (CLAT_HALFLINES=0), the
    per-atom potentials against full lines too."""
    nh, ne = 2, 2
    w = zb.workload(*box, nh, ne, config=0, sigma_frac=0.2, n_orb=4)
    quads = clat.pattern_enumerate("****", nh + ne, nh)
    print("case half lines", box, prec, w["dims"], flush=True)
    monkeypatch.delenv("CLAT_HALFLINES", raising=False)
    a = _engine(w, w["cells"], prec, 4, w["states"], True, monkeypatch)
    monkeypatch.setenv("CLAT_HALFLINES", "0")
    b = _engine(w, w["cells"], prec, 4, w["states"], True, monkeypatch)
    monkeypatch.delenv("CLAT_HALFLINES")
    assert {"cosets", "half_box", "half_lines"} <= a.kernels()
    assert {"cosets", "half_box"} <= b.kernels() and "half_lines" not in b.kernels()
    ga, gb = a.compute_integrals(quads), b.compute_integrals(quads)
    pa, pb = a.potential_at_atoms(3), b.potential_at_atoms(3)
    a.close()
    b.close()
    want = _oracle(w, w["cells"], w["states"], quads, nh + ne)
    if prec == clat.PREC_FP64:
        check(ga, want, prec, quads)
        assert np.max(rel_err(ga, gb)) < 1e-12
        assert np.max(np.abs(pa - pb)) < 1e-12 * np.max(np.abs(pb))
    elif prec == clat.PREC_FP32X:
        check_fp32x_model(ga, want, quads, "half lines fp32x")
        check_fp32x_model(gb, want, quads, "full lines fp32x")
        assert np.max(np.abs(pa - pb)) < 1e-6 * np.max(np.abs(pb))
    else:
        check(ga, want, prec, quads)
        check(gb, want, prec, quads)



@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32X], ids=["fp64", "fp32x"])
def test_half_box_odd_radix(prec, monkeypatch):
    """144x64x64 zincblende grid: coarse x lines of 72 points (9 x 8, an odd
    first radix, so the half-box forward stores its frequencies below 72
    through out-of-range offsets instead of leaving outputs unformed, in a
    plane of 80 columns) and half-length lines of 36 points (6 x 6) on the
    x inverse.  Integrals against the oracle."""
    nh, ne = 2, 1
    w = zb.workload(36, 16, 16, nh, ne, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    assert (X, Y, Z) == (144, 64, 64)
    quads = clat.pattern_enumerate("****", nh + ne, nh)
    print("case half box odd radix", prec, w["dims"], flush=True)
    eng = _engine(w, w["cells"], prec, 4, w["states"], True, monkeypatch)
    assert {"cosets", "half_box", "half_lines"} <= eng.kernels(), eng.kernels()
    got = eng.compute_integrals(quads)
    eng.close()
    want = _oracle(w, w["cells"], w["states"], quads, nh + ne)
    if prec == clat.PREC_FP64:
        check(got, want, prec, quads)
    else:
        check_fp32x_model(got, want, quads, "odd radix fp32x")
