"""The engine's own planning: measured prepares per launch (the counterpart of
the reference's FFTW_MEASURE plans, CoulombCalculator.cpp:23,31) and the
device stage replayed as one HIP graph.  Neither may change a single bit of
the integrals."""
import numpy as np
import pytest

import clat
import zincblende as zb

pytestmark = pytest.mark.gpu


def _engine(w, prec, n_orb=4):
    X, Y, Z = w["dims"]
    eng = clat.Engine((X, Y, Z), prec, clat.SCHED_SHARED)
    eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    eng.set_atoms(w["cells"], n_orb)
    for s in w["states"]:
        eng.add_state(s)
    return eng


@pytest.mark.parametrize("prec", [clat.PREC_FP64, clat.PREC_FP32X], ids=["fp64", "fp32x"])
def test_graph_replays_match_direct_launches(prec, monkeypatch):
    """Runs 2.. of one plan replay the captured graph.  Between the replays
    the engine does eager work of its own (a potential, a run with another
    integral list, which re-captures, a new interaction table) and every
    replay must give exactly the values of direct launches."""
    print("case graph replay", prec, flush=True)
    w = zb.workload(8, 8, 8, 3, 2, config=0, sigma_frac=0.2, n_orb=4)
    quads = clat.pattern_enumerate("****", 5, 3)
    other = clat.pattern_enumerate("eeee,hhhh", 5, 3)
    monkeypatch.setenv("CLAT_GRAPH", "0")
    ref_eng = _engine(w, prec)
    ref = ref_eng.compute_integrals(quads)
    ref_other = ref_eng.compute_integrals(other)
    ref_pot = ref_eng.potential_at_atoms(2)
    assert ref_eng.tuning()["graph_captures"] == 0
    ref_eng.close()
    monkeypatch.delenv("CLAT_GRAPH")
    eng = _engine(w, prec)
    for k in range(4):
        assert np.array_equal(eng.compute_integrals(quads), ref), k
        assert np.array_equal(eng.potential_at_atoms(2), ref_pot)
        if k == 1:
            assert np.array_equal(eng.compute_integrals(other), ref_other)
        if k == 2:
            eng.initialize(clat.InteractionSimple(w["step"], 0.0, 10.89))
    t = eng.tuning()
    assert t["graph_captures"] >= 2 and t["graph_replays"] >= 2, t
    eng.close()


def test_kernel_timing_runs_launch_directly(monkeypatch):
    """Runs with kernel timing on launch every pass directly (events inside a
    graph cannot be timed on this runtime): the per-kind launch counts and
    timings are those of an engine without graphs, no replay happens while
    timing is on, and replays resume when it is off."""
    print("case graph timing", flush=True)
    w = zb.workload(8, 8, 8, 2, 2, config=0, sigma_frac=0.2, n_orb=4)
    quads = clat.pattern_enumerate("****", 4, 2)
    counts = {}
    for graph in ("0", "1"):
        monkeypatch.setenv("CLAT_GRAPH", graph)
        eng = _engine(w, clat.PREC_FP32X)
        eng.compute_integrals(quads)  # measures the batch outside the statistics
        eng.set_kernel_timing(True)
        eng.reset_kernel_stats()
        for _ in range(4):
            eng.compute_integrals(quads)
        st = eng.kernel_stats()
        counts[graph] = {k: v[0] for k, v in st.items()}
        assert all(v[1] > 0 for v in st.values() if v[0]), st
        assert eng.tuning()["graph_replays"] == 0
        if graph == "1":
            eng.set_kernel_timing(False)
            for _ in range(3):
                eng.compute_integrals(quads)
            # the graphs captured by the untimed first run replay again
            assert eng.tuning()["graph_replays"] == 3
        eng.close()
    assert counts["0"] == counts["1"] and counts["1"]["zconv"] > 0


def test_measured_batch_is_cached_and_bit_identical(monkeypatch):
    """The first run of a shape times the candidate batches; a second engine
    of the same shape takes the cached choice without trials; a rule-only
    engine and one prepare per launch give the same bits.  (The choice is
    cached per process, so this shape, a 32x32x64 grid, is used here only.)"""
    print("case autotuner", flush=True)
    w = zb.workload(8, 8, 16, 3, 2, config=0, sigma_frac=0.2, n_orb=4)
    quads = clat.pattern_enumerate("****", 5, 3)
    a = _engine(w, clat.PREC_FP32X)
    va = a.compute_integrals(quads)
    ta = a.tuning()
    assert ta["batch"] in ta["trial_ms"] and len(ta["trial_ms"]) >= 2, ta
    b = _engine(w, clat.PREC_FP32X)
    vb = b.compute_integrals(quads)
    tb = b.tuning()
    assert tb["batch"] == ta["batch"] and tb["trial_ms"] == {}, tb
    c = _engine(w, clat.PREC_FP32X)
    c.set_tuning(False)
    vc = c.compute_integrals(quads)
    # the fastest trial, or the fixed rule's count when within 2 % of it
    trials, rule = ta["trial_ms"], c.tuning()["batch"]
    fastest = min(trials, key=trials.get)
    assert ta["batch"] == fastest or (ta["batch"] == rule and trials[rule] <= 1.02 * trials[fastest]), (ta, rule)
    monkeypatch.setenv("CLAT_BATCH", "1")
    d = _engine(w, clat.PREC_FP32X)
    vd = d.compute_integrals(quads)
    assert d.tuning()["batch"] == 1
    for v in (vb, vc, vd):
        assert np.array_equal(v, va)
    for e in (a, b, c, d):
        e.close()


def test_group_matches_per_engine_runs(monkeypatch):
    """Three systems of different grids (32^3, 32x32x48, 64^3) in one group
    call: their device stages go out as one HIP graph, captured on the first
    call and replayed on later ones, and every member's values are
    bit-identical to its own compute_integrals.  Between the group calls the
    members do eager work (a potential, a solo run with another list) and the
    last group changes one member's integral list, which re-captures.  A
    group with a member off the line kernels (a 24x24x40 grid) launches each
    stage directly, with the same values."""
    print("case group graph", flush=True)
    ws = [zb.workload(8, 8, 8, 2, 2, config=0, sigma_frac=0.2, n_orb=4),
          zb.workload(8, 8, 12, 3, 2, config=0, sigma_frac=0.2, n_orb=4),
          zb.workload(16, 16, 16, 2, 3, config=0, sigma_frac=0.2, n_orb=4)]
    odd = zb.workload(6, 6, 10, 2, 2, config=0, sigma_frac=0.2, n_orb=4)
    lists = [clat.pattern_enumerate("****", 4, 2), clat.pattern_enumerate("****", 5, 3),
             clat.pattern_enumerate("****", 5, 3)]
    other = clat.pattern_enumerate("eeee,hhhh", 5, 3)
    for prec in (clat.PREC_FP64, clat.PREC_FP32X):
        solo = [_engine(w, prec) for w in ws]
        ref = [e.compute_integrals(q) for e, q in zip(solo, lists)]
        ref_other = solo[2].compute_integrals(other)
        ref_pot = solo[1].potential_at_atoms(2)
        # every member runs on the line kernels (the only path with graphs)
        assert all(e.tuning()["batch"] > 0 for e in solo), [e.tuning() for e in solo]
        for e in solo:
            e.close()
        grp = [_engine(w, prec) for w in ws]
        for k in range(4):
            got = clat.compute_integrals_group(grp, lists)
            for i in range(3):
                assert np.array_equal(got[i], ref[i]), (prec, k, i)
            assert np.array_equal(grp[1].potential_at_atoms(2), ref_pot)
            if k == 1:
                assert np.array_equal(grp[2].compute_integrals(other), ref_other)
        got = clat.compute_integrals_group(grp, lists[:2] + [other])
        assert np.array_equal(got[2], ref_other) and np.array_equal(got[0], ref[0])
        t = grp[0].tuning()
        assert t["graph_captures"] >= 2 and t["graph_replays"] >= 3, t
        o_solo = _engine(odd, prec)
        o_ref = o_solo.compute_integrals(lists[0])
        o_solo.close()
        o_eng = _engine(odd, prec)
        for _ in range(2):
            got = clat.compute_integrals_group([o_eng, grp[0]], lists[:1] + lists[:1])
            assert np.array_equal(got[1], ref[0]) and np.array_equal(got[0], o_ref)
        assert o_eng.tuning()["graph_captures"] == 0
        with pytest.raises(clat.ClatError):
            clat.compute_integrals_group([grp[0], grp[0]], lists[:2])
        for e in grp + [o_eng]:
            e.close()
    # members of different precisions in one graph
    mixed = [_engine(ws[0], clat.PREC_FP64), _engine(ws[1], clat.PREC_FP32X)]
    solo = [_engine(ws[0], clat.PREC_FP64), _engine(ws[1], clat.PREC_FP32X)]
    want = [e.compute_integrals(q) for e, q in zip(solo, lists[:2])]
    for _ in range(3):
        got = clat.compute_integrals_group(mixed, lists[:2])
        assert all(np.array_equal(g, w_) for g, w_ in zip(got, want))
    assert mixed[0].tuning()["graph_replays"] >= 2
    for e in mixed + solo:
        e.close()


@pytest.mark.parametrize("box", [(8, 8, 8), (16, 16, 16)])
def test_interaction_table_built_on_first_use(box):
    """The interaction table is enqueued by the first call that needs it.
    A second interaction set before any run replaces the first unseen; the
    table built inside a potential run, inside an integral run or after a
    user table gives the same bits as an engine that only ever saw the last
    interaction (32^3 grid on the single grid, 64^3 on the two-coset path)."""
    print("case deferred table", box, flush=True)
    w = zb.workload(*box, 2, 2, config=0, sigma_frac=0.2, n_orb=4)
    X, Y, Z = w["dims"]
    quads = clat.pattern_enumerate("****", 4, 2)
    first = clat.InteractionSimple(w["step"], 0.3, 4.0)
    last = clat.InteractionSimple(w["step"], 0.0, 10.89)

    def engine(inters):
        eng = clat.Engine((X, Y, Z), clat.PREC_FP64, clat.SCHED_SHARED)
        for i in inters:
            eng.initialize(i)
        eng.set_atoms(w["cells"], 4)
        for s in w["states"]:
            eng.add_state(s)
        return eng

    ref = engine([last])
    want_q, want_p = ref.compute_integrals(quads), ref.potential_at_atoms(3)
    on_cosets = "cosets" in ref.kernels()
    assert on_cosets == (box == (16, 16, 16))
    ref.close()
    # two interactions before any run: only the last is built
    a = engine([first, last])
    assert np.array_equal(a.compute_integrals(quads), want_q)
    a.close()
    # the table built by a potential run first
    b = engine([last])
    assert np.array_equal(b.potential_at_atoms(3), want_p)
    assert np.array_equal(b.compute_integrals(quads), want_q)
    # a new interaction after runs, then the old one again
    b.initialize(first)
    assert not np.allclose(b.compute_integrals(quads), want_q)
    b.initialize(last)
    assert np.array_equal(b.compute_integrals(quads), want_q)
    b.close()
