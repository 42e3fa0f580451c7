"""clat_compute_potential_at_atoms_noalloc: the allocation-free variant of
clat_compute_potential_at_atoms, whose self density, site potential and
per-atom staging live in the library's static device array.  Checked against
the allocating entry point on two calls in a row (the second on a smaller
input, so the array holds the first call's larger contents), in both
precisions, and against the oracle; sizes past the array are rejected."""
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


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32X], ids=["fp64", "fp32x"])
def test_noalloc_matches_allocating_entry(prec):
    print("case noalloc two calls", prec, flush=True)
    big = zb.workload(16, 16, 8, 1, 1, config=0, sigma_frac=0.2, n_orb=4)     # 64x64x32, 16384 atoms
    small = zb.workload(4, 4, 4, 1, 1, config=0, sigma_frac=0.2, n_orb=4)     # 16^3, 512 atoms
    outs = []
    for w, state in ((big, 2), (small, 1)):
        eng = _engine(w, prec, 4)
        want = eng.potential_at_atoms(state)
        got = eng.potential_at_atoms_noalloc(state)
        # same kernels on the same inputs: the same bits
        assert np.array_equal(got, want)
        # and the oracle's potential of that state's self density
        X, Y, Z = w["dims"]
        F = np.zeros(X * Y * Z, complex)
        F[w["cells"]] = np.sum(np.abs(w["states"][state - 1]) ** 2, axis=1)
        Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
        Vo = O.prepare(F.reshape(Z, Y, X), Gh).reshape(-1)[w["cells"]]
        bar = 1e-10 if prec == clat.PREC_FP64 else 1e-5
        assert np.max(np.abs(got - Vo)) / np.max(np.abs(Vo)) < bar
        outs.append(got)
        eng.close()
    assert len(outs[1]) < len(outs[0])


def test_noalloc_repeated_calls_reuse_out():
    """Back-to-back calls on one engine into one output array, alternating
    states: each call leaves exactly the allocating entry point's values."""
    print("case noalloc repeated", flush=True)
    w = zb.workload(4, 4, 4, 1, 2, config=0, sigma_frac=0.2, n_orb=4)
    eng = _engine(w, clat.PREC_FP32X, 4)
    want = [eng.potential_at_atoms(k) for k in (1, 2, 3)]
    out = np.zeros(len(w["cells"]), np.complex128)
    for k in (3, 1, 2, 1, 3):
        eng.potential_at_atoms_noalloc(k, out)
        assert np.array_equal(out, want[k - 1])
    eng.close()


def test_noalloc_rejects_sizes_past_the_array():
    """2 * sites + atoms above 4,194,304: CLAT_EUNSUP with the sizes named.
    Here 4.2M atoms share 64 cells (the last one of each cell counts)."""
    print("case noalloc size limit", flush=True)
    n_atoms = (1 << 22) + 16
    cells = (np.arange(n_atoms) % 64).astype(np.int32)
    eng = clat.Engine((8, 8, 8), clat.PREC_FP64, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionSimple((1.0, 1.0, 1.0), 0.0, 10.89))
    eng.set_atoms(cells, 1)
    eng.add_state(np.ones((n_atoms, 1), np.complex128))
    with pytest.raises(clat.ClatError) as ex:
        eng.potential_at_atoms_noalloc(1)
    assert ex.value.code == 5 and "4194304" in str(ex.value) and "scratch" in str(ex.value)
    # the allocating entry point still serves the same engine
    v = eng.potential_at_atoms(1)
    assert v.shape == (n_atoms,) and np.all(np.isfinite(v))
    eng.close()


def test_noalloc_through_c_abi_symbol():
    """The entry point as a C caller binds it: by name, with plain pointers."""
    print("case noalloc c abi", flush=True)
    L = clat.lib()
    fn = L.clat_compute_potential_at_atoms_noalloc
    w = zb.workload(2, 2, 2, 1, 0, config=0, sigma_frac=0.2, n_orb=2)
    eng = _engine(w, clat.PREC_FP64, 2)
    out = np.zeros(2 * len(w["cells"]))
    assert fn(eng.handle, 1, out) == 0
    got = out.view(np.complex128)
    assert np.array_equal(got, eng.potential_at_atoms(1))
    assert fn(eng.handle, 2, out) == 1  # state out of range: CLAT_EINVAL
    assert b"state index" in L.clat_last_error()
    eng.close()
