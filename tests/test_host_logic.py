"""Host-side integer logic of the C-ABI library (no GPU): grid rounding,
lattice detection, cell indexing, patterns, vertex cover and the planner.

Checked against oracle/_ref (the reference's own Round.hpp, Graph.hpp and
Pattern.hpp compiled from the reference tree) where that library exists, and
against the oracle's Python restatements otherwise."""
import numpy as np
import pytest

import clat
import oracle as O
import zincblende as zb

REF = O.ref_lib()


def test_round_up_tables():
    """run-tests.cpp:295-323."""
    assert [clat.round_up(n, 1, (2,)) for n in (0, 13, 16, 17, 34, 101)] == [1, 16, 16, 32, 64, 128]
    assert [clat.round_up(n, 1) for n in (0, 13, 16, 17, 34, 101)] == [1, 15, 16, 18, 36, 108]
    assert [clat.round_up(n, 4) for n in (0, 13, 16, 17, 34, 101)] == [4, 16, 16, 20, 36, 108]


def _two_level(m):
    """(R1, R2), R1 >= R2 the most balanced split of m into register sizes, or None."""
    regs = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 20}
    best = None
    for r2 in range(1, int(m ** 0.5) + 1):
        if m % r2 == 0 and r2 in regs and m // r2 in regs:
            best = (m // r2, r2)
    return best


def test_engine_length_keeps_reference_grids():
    """Every grid length the reference produces (Round::up<2,3,5>, as
    clat.round_up) from 32 to 400 is computed on as it is when it splits
    into two register sizes, and on the smallest length at or above it that
    splits otherwise; the result always splits, is never shorter, and for
    lengths that do not split lies no further than the next 2*3*5 length
    that does (brute force over all n up to 400)."""
    splits = [m for m in range(32, 401) if clat.round_up(m, 1) == m and _two_level(m)]
    for n in range(32, 401):
        e = clat.engine_length(n)
        assert e >= n and _two_level(e) is not None, (n, e)
        if clat.round_up(n, 1) == n and _two_level(n):
            assert e == n, (n, e)
        nxt = min(m for m in splits if m >= n) if any(m >= n for m in splits) else None
        if nxt is not None:
            assert e <= nxt, (n, e, nxt)
    assert [clat.engine_length(n) for n in (45, 100, 125, 250, 384, 49)] == [45, 100, 128, 256, 400, 50]
    assert clat.engine_length(401) == clat.round_up(401, 1, (2, 3, 5, 7))


@pytest.mark.skipif(REF is None, reason="reference tree absent: oracle/_ref not built")
def test_round_up_vs_reference_build():
    for c in (1, 2, 3, 4, 7):
        for n in range(0, 3000, 7):
            assert clat.round_up(n, c) == REF.ref_round_up_235(n, c)
            assert clat.round_up(n, c, (2,)) == REF.ref_round_up_2(n, c)


def _random_graph(rng):
    nv = int(rng.integers(1, 40))
    ne = int(rng.integers(0, 3 * nv))
    es = rng.integers(0, nv, (ne, 2)).astype(np.int32)
    loops = rng.random(nv) < 0.3
    if loops.any():
        lv = np.nonzero(loops)[0].astype(np.int32)
        es = np.concatenate([es, np.stack([lv, lv], 1)])
    return nv, es


def test_vertex_cover_vs_restatement():
    rng = np.random.default_rng(1)
    for _ in range(200):
        nv, es = _random_graph(rng)
        assert clat.vertex_cover(nv, es) == O.vertex_cover(nv, [tuple(e) for e in es])


@pytest.mark.skipif(REF is None, reason="reference tree absent: oracle/_ref not built")
def test_vertex_cover_vs_reference_build():
    rng = np.random.default_rng(2)
    for _ in range(300):
        nv, es = _random_graph(rng)
        cov = np.zeros(nv, np.int32)
        e = np.ascontiguousarray(es.reshape(-1), np.int32)
        k = REF.ref_vertex_cover(nv, e if len(e) else np.zeros(2, np.int32), len(es), cov)
        assert clat.vertex_cover(nv, es) == sorted(int(v) for v in cov[:k])


PATTERNS = ["****", "eeee,hhhh,ehhe,eheh", "ijji", "1*1*", "ijkl", "iijj", "*i3i", "hHhH", "aBcA,1212",
            "eh*e", "9999", "i*i*,j**j"]


@pytest.mark.parametrize("spec", PATTERNS)
def test_pattern_vs_restatement(spec):
    for n, nh in ((1, 0), (3, 2), (4, 1), (6, 3)):
        got = [tuple(q) for q in clat.pattern_enumerate(spec, n, nh)]
        assert got == O.pattern_enumerate(spec, n, nh)


@pytest.mark.skipif(REF is None, reason="reference tree absent: oracle/_ref not built")
@pytest.mark.parametrize("spec", PATTERNS)
def test_pattern_vs_reference_build(spec):
    for n, nh in ((1, 0), (3, 2), (4, 1), (6, 3)):
        cap = n ** 4
        buf = np.zeros(4 * cap, np.int32)
        k = REF.ref_pattern_enumerate(spec.encode(), n, nh, buf, cap, None, 0)
        ref = [tuple(int(v) for v in buf[4 * i:4 * i + 4]) for i in range(k)]
        assert [tuple(q) for q in clat.pattern_enumerate(spec, n, nh)] == ref


@pytest.mark.parametrize("bad", ["ee0e", "e-ee", "", "eeee,", ",eeee", "e e"])
def test_pattern_errors(bad):
    """Pattern.hpp:40-58: invalid character / invalid empty pattern."""
    with pytest.raises(clat.ClatError) as ex:
        clat.pattern_enumerate(bad, 2, 1)
    assert "pattern" in str(ex.value)


def _plan_key(rows):
    return sorted(tuple(int(v) for v in r) for r in rows)


@pytest.mark.parametrize("spec,n,nh", [("****", 4, 2), ("eeee,hhhh,ehhe,eheh", 8, 4), ("ijkl", 5, 2),
                                       ("*i3i", 6, 3), ("1*2*", 4, 1), ("ijji", 7, 3)])
def test_plan_vs_restatement(spec, n, nh):
    """MasterPlanner: same steps (id, left, lconj, right, rconj) and the same
    execution order up to ties of equal sort key (Planner.hpp:94-101)."""
    quads = clat.pattern_enumerate(spec, n, nh)
    got = clat.plan(quads, n)
    want = O.plan([tuple(q) for q in quads], n)
    assert _plan_key(got) == _plan_key([(k, L, lc, R, rc) for k, (L, lc), (R, rc) in want])
    key = lambda L, lc, R, rc: (L, R, 0 if lc == rc else 1)  # noqa: E731
    assert [key(*r[1:]) for r in got] == [key(L, lc, R, rc) for _, (L, lc), (R, rc) in want]


@pytest.mark.parametrize("n", [40, 320])
def test_plan_large_state_counts(n):
    """Many states: the product graph of 40 states goes through the edge
    bitmap and the radix-sorted steps (2 P^2 keys, more than 2^20 and than
    four per step); 320 states (P^2 past both limits) through the sorted
    edge list and the comparison sort.  Smaller key ranges, as in the
    restatement tests above, take one counting pass."""
    rng = np.random.default_rng(n)
    quads = rng.integers(1, n + 1, size=(3000, 4)).astype(np.int32)
    got = clat.plan(quads, n)
    want = O.plan([tuple(int(v) for v in q) for q in quads], n)
    assert _plan_key(got) == _plan_key([(k, L, lc, R, rc) for k, (L, lc), (R, rc) in want])
    key = lambda L, lc, R, rc: (L, R, 0 if lc == rc else 1)  # noqa: E731
    assert [key(*r[1:]) for r in got] == [key(L, lc, R, rc) for _, (L, lc), (R, rc) in want]


def test_plan_threaded_ranges():
    """200000 integrals: the steps, the edge bitmap and the cover swaps are
    built by several host threads over id ranges; same plan as the
    restatement, and a bad state index inside a later range is still an
    error."""
    rng = np.random.default_rng(12)
    quads = rng.integers(1, 13, size=(200000, 4)).astype(np.int32)
    got = clat.plan(quads, 12)
    want = O.plan([tuple(int(v) for v in q) for q in quads], 12)
    assert _plan_key(got) == _plan_key([(k, L, lc, R, rc) for k, (L, lc), (R, rc) in want])
    key = lambda L, lc, R, rc: (L, R, 0 if lc == rc else 1)  # noqa: E731
    assert [key(*r[1:]) for r in got] == [key(L, lc, R, rc) for _, (L, lc), (R, rc) in want]
    quads[170000, 1] = 13
    with pytest.raises(clat.ClatError):
        clat.plan(quads, 12)


def test_plan_product_indexing():
    """Planner.hpp:134-144: product(iL, iR) = iL(iL-1)/2 + iR - 1 with swap+conj."""
    rows = clat.plan([(1, 2, 3, 4)], 4)
    (_, L, lc, R, rc), = rows
    assert {(L, lc), (R, rc)} == {(4 * 3 // 2 + 1 - 1, 1), (3 * 2 // 2 + 2 - 1, 1)}


def test_plan_rejects_bad_state():
    with pytest.raises(clat.ClatError):
        clat.plan([(1, 2, 3, 9)], 4)


def test_lattice_vs_oracle_example(golden_dir):
    import os
    xyz = np.loadtxt(os.path.join(golden_dir, "example", "X.3d"), usecols=(0, 1, 2))
    a = clat.lattice_grid(xyz)
    b = O.lattice_grid(xyz)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(clat.lattice_cells(xyz, a[1], a[2], a[3]), O.lattice_cells(xyz, b[1], b[2], b[3]))


def test_lattice_zincblende_closed_form():
    w = zb.workload(5, 3, 2, 1, 0, n_orb=1)
    raw, pad, origin, step = clat.lattice_grid(w["pos"])
    assert list(raw) == [20, 12, 8] and list(pad) == [20, 12, 8]
    assert np.array_equal(clat.lattice_cells(w["pos"], pad, origin, step), w["cells"])


def test_lattice_padding_and_shift():
    """Raw 41 -> 45 (2*3*5 rounding); lrint on shifted, jittered coordinates."""
    rng = np.random.default_rng(4)
    idx = rng.integers(0, 41, (500, 3))
    idx[0] = 0
    idx[1] = 40
    step = np.array([1.1, 0.7, 2.3])
    xyz = idx * step + np.array([-3.0, 5.0, 0.25]) + rng.uniform(-1e-12, 1e-12, idx.shape)
    raw, pad, origin, st = clat.lattice_grid(xyz)
    assert list(raw) == [41, 41, 41] and list(pad) == [45, 45, 45]
    cells = clat.lattice_cells(xyz, pad, origin, st)
    assert np.array_equal(cells, (idx[:, 2] * 45 + idx[:, 1]) * 45 + idx[:, 0])


def test_lattice_error_message():
    xyz = np.array([[0, 0, 0], [1.0, 0, 0], [2.005, 0, 0]], float)
    with pytest.raises(clat.ClatError) as ex:
        clat.lattice_grid(xyz)
    assert "could not find common grid step" in str(ex.value)
