"""Pins the FP64 CPU oracle before it is trusted as the checker.

Each function of oracle.c / oracle.py is checked against the reference's own
known-answer values, the committed golden fixtures, a library routine or a
closed form -- never against a retyped copy of its own formula."""
import json
import os

import numpy as np
import pytest

import oracle as O
import zincblende as zb

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _mock_table(X, Y, Z):
    iz, iy, ix = np.meshgrid(np.arange(Z + 1), np.arange(Y + 1), np.arange(X + 1), indexing="ij")
    return np.exp(-ix - iy - iz).astype(np.float64)


def test_two_point_charges_kat():
    """run-tests.cpp:28-48 (CoulombCalculatorTest): 5x5x10, mock kernel
    exp(-ix-iy-iz): delta at (2,2,2) -> 1.0, delta at (2,2,7) -> e^-5, +-1e-12.
    Pins DCT-I, the 8 rounds, 1/(8N) and the reflected lookup together."""
    X, Y, Z = 5, 5, 10
    Gh = O.ghat(_mock_table(X, Y, Z))
    F = np.zeros((Z, Y, X), complex)
    F[2, 2, 2] = 1
    V = O.prepare(F, Gh)
    r = O.calculate(V, F)
    assert abs(r.real - 1.0) < 1e-12 and abs(r.imag) < 1e-12
    F2 = np.zeros((Z, Y, X), complex)
    F2[7, 2, 2] = 1
    r = O.calculate(V, F2)
    assert abs(r.real - np.exp(-5)) < 1e-12 and abs(r.imag) < 1e-12


def test_fft_sign_and_scale_kat():
    """run-tests.cpp:189-214 (FourierPlanTest): 1024-point transform of
    cos(2 pi 16 n/1024 + pi/4): |X16| = 512, arg = pi/4, others 0 (1e-12)."""
    n = 1024
    x = np.cos(2 * np.pi * 16 * np.arange(n) / n + np.pi / 4)
    Xf = O.fft(x, -1)[: n // 2 + 1]
    assert abs(abs(Xf[16]) - 512) < 1e-12
    assert abs(np.angle(Xf[16]) - np.pi / 4) < 1e-12
    assert np.max(np.abs(np.delete(Xf, 16))) < 1e-12


@pytest.mark.parametrize("n", [8, 12, 45, 64, 96, 120])
def test_fft_and_dct_vectors(n):
    """c2c FFT (both signs) and REDFT00 vs scipy.fft golden vectors."""
    v = np.load(os.path.join(GOLD, "dct_fft_vectors.npz"))
    z = v["fft_in_%d" % n]
    assert np.max(np.abs(O.fft(z, -1) - v["fft_out_%d" % n])) < 1e-12 * n
    assert np.max(np.abs(O.fft(z, +1) - v["ifft_out_%d" % n])) < 1e-12 * n
    x = v["dct_in_%d" % n]
    got = O.dct1_3d(x.reshape(1, 1, -1)).reshape(-1)
    assert np.max(np.abs(got - v["dct_out_%d" % n])) < 1e-12 * n


def test_dct3_vectors():
    v = np.load(os.path.join(GOLD, "dct_fft_vectors.npz"))
    got = O.dct1_3d(v["dct3_in"])
    assert np.max(np.abs(got - v["dct3_out"])) < 1e-11


def test_example_1111_direct_and_rounds():
    """example/run.sh: --atoms=X.3d --dielectric=11.4 --orbitals=10
    --skip-lines=1 e1.dat -> ' 1  1  1  1    0.090340071977'.  The 8-round
    oracle and the direct O(N^2) sum must both print that prefix, and agree
    with each other to 1e-12 relative (the method is exact)."""
    d = os.path.join(GOLD, "example")
    xyz = np.loadtxt(os.path.join(d, "X.3d"), usecols=(0, 1, 2))
    c = np.loadtxt(os.path.join(d, "e1.dat"), skiprows=1, usecols=(0, 1))
    c = (c[:, 0] + 1j * c[:, 1]).reshape(1, len(xyz), 10)
    raw, pad, origin, step = O.lattice_grid(xyz)
    assert list(raw) == [41, 41, 41] and list(pad) == [45, 45, 45]
    cells = O.lattice_cells(xyz, pad, origin, step)
    rho = O.products(c)[0]
    X, Y, Z = pad
    F = np.zeros(X * Y * Z, complex)
    F[cells] = rho
    G = O.gtable(X, Y, Z, step, 0.0, 11.4)
    V = O.prepare(F.reshape(Z, Y, X), O.ghat(G))
    v_rounds = O.calculate(V, F.reshape(Z, Y, X))
    v_direct = complex(np.sum(O.direct_potential(F.reshape(Z, Y, X), G, cells) * rho))
    for v in (v_rounds, v_direct):
        line = "%2d %2d %2d %2d   %17.14f %17.14f\n" % (1, 1, 1, 1, v.real, v.imag)
        assert line[:29] == " 1  1  1  1    0.090340071977"
    assert abs(v_rounds - v_direct) / abs(v_direct) < 1e-12


def _gold_cases():
    with open(os.path.join(GOLD, "zincblende_direct.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _gold_cases(), ids=lambda c: c["name"])
def test_oracle_main_loop_vs_direct_sum(case):
    """The 8-round oracle driven by the restated planner/main loop
    (conj bookkeeping, cover swap) reproduces the direct-sum fixtures."""
    w = zb.workload(*case["cells"], case["holes"], case["electrons"], config=case["config"],
                    sigma_frac=case["sigma_frac"], n_orb=case["orbitals"])
    import hashlib
    h = hashlib.sha256()
    for s in w["states"]:
        h.update(np.ascontiguousarray(s, np.complex128).tobytes())
    assert h.hexdigest() == case["coeff_sha256"], "generator drift"
    X, Y, Z = w["dims"]
    n = case["holes"] + case["electrons"]
    quads = [tuple(q) for q in case["quads"]]
    assert quads == O.pattern_enumerate(case["integrals"], n, case["holes"])
    rho = O.products(np.stack(w["states"]))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], case["onsite"], case["dielectric"]))
    got = O.compute_integrals(quads, n, rho, w["cells"], (X, Y, Z), Gh)
    want = np.array([complex(a, b) for a, b in case["values"]])
    assert np.max(np.abs(got - want) / np.abs(want)) < 1e-12


def test_integral_symmetries():
    """Invariants of the Coulomb form on the golden values: Hermitian
    I(1234) = conj(I(4321)); G even: I(1234) = I(2143); J = I(ijji) real."""
    case = _gold_cases()[0]
    vals = {tuple(q): complex(a, b) for q, (a, b) in zip(case["quads"], case["values"])}
    scale = max(abs(v) for v in vals.values())
    for (a, b, c, d), v in vals.items():
        assert abs(v - np.conj(vals[(d, c, b, a)])) < 1e-13 * scale
        assert abs(v - vals[(b, a, d, c)]) < 1e-13 * scale
        if a == d and b == c:
            assert abs(v.imag) < 1e-13 * scale


def test_padding_independence():
    """Appendix A of the method: any padding gives the same potential; an
    oracle run on a larger grid must agree on the original cells."""
    rng = np.random.default_rng(5)
    X, Y, Z = 6, 5, 4
    F = rng.standard_normal((Z, Y, X)) + 1j * rng.standard_normal((Z, Y, X))
    step = (1.0, 1.2, 0.8)
    V = O.prepare(F, O.ghat(O.gtable(X, Y, Z, step, 0.3, 2.0)))
    Fp = np.zeros((Z + 3, Y + 2, X + 5), complex)
    Fp[:Z, :Y, :X] = F
    Vp = O.prepare(Fp, O.ghat(O.gtable(X + 5, Y + 2, Z + 3, step, 0.3, 2.0)))
    assert np.max(np.abs(Vp[:Z, :Y, :X] - V)) < 1e-12 * np.max(np.abs(V))


def test_prepare_vs_direct_random():
    """8-round potential == direct O(N^2) sum on a random dense input."""
    rng = np.random.default_rng(11)
    X, Y, Z = 7, 6, 5
    F = rng.standard_normal((Z, Y, X)) + 1j * rng.standard_normal((Z, Y, X))
    G = O.gtable(X, Y, Z, (1.3, 0.7, 1.1), 0.25, 1.7)
    V = O.prepare(F, O.ghat(G)).reshape(-1)
    Vd = O.direct_potential(F, G, np.arange(X * Y * Z, dtype=np.int32))
    assert np.max(np.abs(V - Vd)) < 1e-12 * np.max(np.abs(Vd))


def _tf_eps_profile(dielectric, a, step, n):
    """eps(r) = E2_4PE0 / (r G(r)) read back from the oracle's Thomas-Fermi
    table along the x axis, r = k * step for k = 1 .. n."""
    G = O.gtable(n, 0, 0, (step, 1.0, 1.0), 0.0, dielectric, tf_lattice=a).reshape(-1)
    r = np.arange(1, n + 1) * step
    return r, O.E2_4PE0 / (r * G[1:])


@pytest.mark.parametrize("dielectric,a", [(10.89, 5.6533), (11.4, 5.431), (3.0, 4.0)])
def test_thomas_fermi_screening_properties(dielectric, a):
    """Interaction.cpp:18-39, pinned by what its mathematics fixes instead of
    by its formula.  With x the positive root of sinh(x) = eps x and
    q rTF = x, the Resta screening eps(r) must
      * tend to 1 as r -> 0 (no screening at contact: eps x / sinh(x) = 1),
        with slope q tanh(x/2) there (first order of the sinh form),
      * rise monotonically on (0, rTF) and stay within [1, eps],
      * reach the bulk eps continuously at rTF and equal it beyond (there the
        table is bit-identical to the unscreened one),
    and the crossover radius must be the root x over the Thomas-Fermi wave
    vector q = 2/sqrt(pi) (96 pi^2)^(1/3) / a that Interaction.cpp:21 states.
    A sign slip inside the sinh, a dropped eps factor or a wrong root fails
    one of these."""
    from scipy.optimize import brentq
    x = brentq(lambda t: np.sinh(t) - dielectric * t, 1e-3, 50.0)
    q = 2.0 / np.sqrt(np.pi) * np.cbrt(96.0 * np.pi ** 2) / a
    rTF = x / q
    step = rTF / 2000.5  # no grid point lands on rTF itself
    r, eps = _tf_eps_profile(dielectric, a, step, 3000)
    inner = r < rTF
    assert inner.sum() == 2000
    # contact limit and initial slope, extrapolated from two points at
    # r = h, 2h with h = 1e-7 rTF (the O(r^2) terms are then ~1e-13)
    h = rTF * 1e-7
    _, e0 = _tf_eps_profile(dielectric, a, h, 2)
    slope = (e0[1] - e0[0]) / h
    assert abs((2 * e0[0] - e0[1]) - 1.0) < 1e-10, e0
    assert abs(slope - q * np.tanh(x / 2)) < 1e-6 * q, (slope, q * np.tanh(x / 2))
    # monotone, bounded
    assert np.all(np.diff(eps[inner]) > 0)
    assert np.all(eps[inner] >= 1.0) and np.all(eps[inner] < dielectric)
    # continuity at rTF: the last inner point is within one step of slope of eps
    slope_max = np.max(np.diff(eps[inner]) / step)
    assert dielectric - eps[inner][-1] <= slope_max * step * 1.01
    # beyond rTF: bulk screening, the unscreened table itself
    assert np.all(eps[~inner] == pytest.approx(dielectric, rel=1e-14))
    Gs = O.gtable(3000, 0, 0, (step, 1.0, 1.0), 0.0, dielectric).reshape(-1)[1:]
    Gt = O.gtable(3000, 0, 0, (step, 1.0, 1.0), 0.0, dielectric, tf_lattice=a).reshape(-1)[1:]
    assert np.array_equal(Gt[~inner], Gs[~inner])


def test_thomas_fermi_table_on_3d_grid():
    """On a 3-D grid the screened table depends on |r| only: every point at
    the same distance (permuted axes with equal steps) holds the same value,
    points beyond rTF equal the unscreened table, and G(0) is the on-site
    value."""
    step = (0.7, 0.7, 0.7)
    G = O.gtable(6, 6, 6, step, 0.25, 10.0, tf_lattice=5.6533)
    Gs = O.gtable(6, 6, 6, step, 0.25, 10.0)
    assert G[0, 0, 0] == 0.25
    # (r is summed in another order after a permutation: equal to rounding)
    for perm in ((1, 0, 2), (2, 1, 0), (0, 2, 1)):
        assert np.allclose(G, np.transpose(G, perm), rtol=1e-14, atol=0)
    r = 0.7 * np.sqrt(np.add.outer(np.add.outer(np.arange(7) ** 2, np.arange(7) ** 2), np.arange(7) ** 2))
    q = 2 / np.sqrt(np.pi) * np.cbrt(96 * np.pi ** 2) / 5.6533
    from scipy.optimize import brentq
    rTF = brentq(lambda t: np.sinh(t) - 10.0 * t, 1e-3, 50.0) / q
    far = r > rTF * (1 + 1e-12)
    assert far.any() and (~far).sum() > 1
    assert np.array_equal(G[far], Gs[far])
    near = (~far) & (r > 0)
    assert np.all(G[near] > Gs[near])  # less screening at short range


def test_complex64_density_rounding_floor():
    """The error model of any complex64 FFT path, shown on the oracle alone:
    on the 64^3 multi-system box (16^3 cells, 4 holes + 4 electrons, 20
    orbitals) the exchange-type integral I(4,3,4,3) cancels to 7e-8 of the
    largest integral.  Rounding its pair density conj(f4) f3 to complex64
    once, with everything after it exact FP64, moves it by 1.3e-5 of itself
    (9e-13 of the largest integral), while the same rounding moves the
    largest integral by 3e-10 of itself.  So below ~1e-6 of the largest no
    complex64 path can promise 1e-5 relative; an absolute bound of 1e-11 of
    the largest is what the storage format allows."""
    w = zb.workload(16, 16, 16, 4, 4, config=4, sigma_frac=0.15, n_orb=20)
    X, Y, Z = w["dims"]
    st, cells = w["states"], w["cells"]
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))

    def integral(rl, rr):
        F = np.zeros(X * Y * Z, complex)
        F[cells] = rl
        return np.sum(O.prepare(F.reshape(Z, Y, X), Gh).reshape(-1)[cells] * rr)

    def c64(r):
        return r.astype(np.complex64).astype(np.complex128)

    r11 = np.sum(np.abs(st[0]) ** 2, axis=1)
    r43 = np.sum(np.conj(st[3]) * st[2], axis=1)   # left product of (4,3,4,3): conj(f4) f3
    big = integral(r11, r11)                        # I(1,1,1,1), the largest
    small = integral(r43, np.conj(r43))             # I(4,3,4,3) = <r43|G|r43> >= 0
    assert small.real > 0 and abs(small) / abs(big) < 1e-7
    d_small = abs(integral(c64(r43), np.conj(r43)) - small)
    d_big = abs(integral(c64(r11), r11) - big)
    assert d_small > 1e-5 * abs(small), d_small / abs(small)
    assert d_small < 1e-11 * abs(big), d_small / abs(big)
    assert d_big < 1e-8 * abs(big), d_big / abs(big)


def test_export_rows_hand_fixture():
    """coulombo.cpp:213-262 worked by hand for 2 holes (states 1, 2) and one
    electron (state 3): per-type numbers are n = spec - 2 for the electron
    and n = 3 - spec for holes (reversed: hole state 1 is 'hole 2'), file
    name letters h/e per position, rows sorted by the per-type numbers,
    format '%2d %2d %2d %2d   %17.14f %17.14f'."""
    quads = [(1, 2, 2, 1), (2, 2, 2, 2), (1, 1, 1, 1), (3, 1, 1, 3), (1, 3, 3, 1), (3, 3, 3, 3),
             (2, 3, 1, 3), (2, 1, 2, 1)]
    values = [0.25 - 0.5j, 1.0, -2.0, 0.125j, 3.5, -0.75 + 0.0625j, 1e-3, 2.0 + 2.0j]
    files = O.export_rows(quads, values, 2, 1)
    assert files == {
        # (2,2,2,2) -> 1 1 1 1; (2,1,2,1) -> 1 2 1 2; (1,2,2,1) -> 2 1 1 2; (1,1,1,1) -> 2 2 2 2
        "hhhh": (" 1  1  1  1    1.00000000000000  0.00000000000000\n"
                 " 1  2  1  2    2.00000000000000  2.00000000000000\n"
                 " 2  1  1  2    0.25000000000000 -0.50000000000000\n"
                 " 2  2  2  2   -2.00000000000000  0.00000000000000\n"),
        "ehhe": " 1  2  2  1    0.00000000000000  0.12500000000000\n",   # (3,1,1,3)
        "heeh": " 2  1  1  2    3.50000000000000  0.00000000000000\n",   # (1,3,3,1)
        "eeee": " 1  1  1  1   -0.75000000000000  0.06250000000000\n",   # (3,3,3,3)
        "hehe": " 1  1  2  1    0.00100000000000  0.00000000000000\n",   # (2,3,1,3)
    }


def test_round_tables():
    """run-tests.cpp:295-323 (RoundTest)."""
    assert [O.round_up(n, 1, (2,)) for n in (0, 13, 16, 17, 34, 101)] == [1, 16, 16, 32, 64, 128]
    assert [O.round_up(n, 1) for n in (0, 13, 16, 17, 34, 101)] == [1, 15, 16, 18, 36, 108]
    assert [O.round_up(n, 4) for n in (0, 13, 16, 17, 34, 101)] == [4, 16, 16, 20, 36, 108]


def test_round_brute_force():
    def smooth(m):
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        return m == 1
    for n in range(1, 400):
        want = next(m for m in range(n, 4 * n + 2) if smooth(m))
        assert O.round_up(n, 1) == want


def test_graph_cases():
    """run-tests.cpp:230-281 (GraphTest)."""
    assert len(O.vertex_cover(4, [(i, j) for j in range(4) for i in range(j)])) == 3
    assert O.vertex_cover(3, []) == []
    assert len(O.vertex_cover(3, [(0, 0), (1, 1), (2, 2)])) == 3
    assert O.vertex_cover(3, [(0, 1), (1, 2)]) == [1]
    c = O.vertex_cover(3, [(0, 0), (0, 1), (1, 1), (1, 2), (2, 0)])
    assert len(c) == 2 and 0 in c and 1 in c


def test_vertex_cover_is_cover_random():
    rng = np.random.default_rng(3)
    for _ in range(50):
        nv = int(rng.integers(1, 12))
        es = [tuple(int(v) for v in rng.integers(0, nv, 2)) for _ in range(int(rng.integers(0, 20)))]
        c = set(O.vertex_cover(nv, es))
        assert all(u in c or v in c for u, v in es)


def test_readme_pattern_examples():
    """README.md:136-147 pattern examples (2 states: h1 then e1 -> 1=h, 2=e)."""
    assert O.pattern_enumerate("ijji", 2, 1) == [(1, 1, 1, 1), (1, 2, 2, 1), (2, 1, 1, 2), (2, 2, 2, 2)]
    assert O.pattern_enumerate("1*1*", 2, 1) == [(1, 1, 1, 1), (1, 1, 1, 2), (1, 2, 1, 1), (1, 2, 1, 2)]
    assert O.pattern_enumerate("1212", 2, 1) == [(1, 2, 1, 2)]
    # ehhe with h1 h2 e1 (states 1,2 holes; 3 electron)
    assert O.pattern_enumerate("ehhe", 3, 2) == [(3, 1, 1, 3), (3, 1, 2, 3), (3, 2, 1, 3), (3, 2, 2, 3)]


def test_zincblende_lattice_closed_form():
    """Broadcaster.cpp:24-43 on zincblende: step a/4, 4n points per axis,
    origin 0, cell index (iz*Y + iy)*X + ix of the generator's grid index."""
    w = zb.workload(3, 2, 4, 1, 0, n_orb=1)
    raw, pad, origin, step = O.lattice_grid(w["pos"])
    assert list(raw) == [12, 8, 16]
    assert list(pad) == [12, 8, 16]
    assert np.allclose(step, 5.6533 / 4, rtol=1e-12) and np.allclose(origin, 0)
    assert np.array_equal(O.lattice_cells(w["pos"], pad, origin, step), w["cells"])


def test_lattice_rejects_non_lattice():
    """Broadcaster.cpp:38: 'could not find common grid step'."""
    xyz = np.array([[0, 0, 0], [1.0, 0, 0], [2.005, 0, 0]], float)
    with pytest.raises(ValueError):
        O.lattice_grid(xyz)
