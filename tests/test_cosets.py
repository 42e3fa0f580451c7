"""The mathematics of the engine's two-coset path, checked on the CPU against
the oracle (no GPU).

When every site of an even grid lies on x = y = z (mod 2), the sites split
into the cosets 2m (p = 0) and 2m + (1,1,1) (p = 1) of the coarse grid with
half the points per axis, and the open-boundary convolution at the sites is

    V_0 = G_0 * rho_0 + G_-t * rho_1,    V_1 = G_t * rho_0 + G_0 * rho_1,

with G_p(d) = G(2d + p) convolved over the coarse grid.  The engine builds the
coarse transforms of G_0 and G_(1,1,1) in one of two ways: from the fine DCT-I
table Gs (the octant of DCT-I(G) / (8XYZ), CoulombCalculator.cpp:57-103) by
summing its eight half-box shifts (kernels.hpp: k_coarse_g0, k_coarse_gt), or
from the real-space samples through a coarse DCT-I and DCT-II (k_coset_gtable or k_coset_samples,
k_dct12, k_coarse_q).  These tests pin both against the transforms of the
sampled kernels, and the whole decomposition against the oracle's eight-round
prepare (CoulombCalculator.cpp:105-163)."""
import numpy as np
import pytest

import oracle as O


def coarse_tables(Gs):
    """Restatement of k_coarse_g0 / k_coarse_gt on the full coarse box:
    G0(q) = sum_s Gs(q + s X),  Gt(q) = e^{i pi (qx/X + qy/Y + qz/Z)} sum_s (-1)^{|s|} Gs(q + s X),
    with Gs(k) for k in [X, 2X) read at the octant entry 2X - k."""
    Z1, Y1, X1 = Gs.shape
    X, Y, Z = X1 - 1, Y1 - 1, Z1 - 1
    qz, qy, qx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    g0 = np.zeros((Z, Y, X))
    st = np.zeros((Z, Y, X))
    for c in (0, 1):
        for b in (0, 1):
            for a in (0, 1):
                v = Gs[Z - qz if c else qz, Y - qy if b else qy, X - qx if a else qx]
                g0 += v
                st += v * (-1.0) ** (a + b + c)
    gt = np.exp(1j * np.pi * (qx / X + qy / Y + qz / Z)) * st
    return g0, gt


def sampled_transform(G, p):
    """Transform of the coarse kernel G_p(d) = G(2d + p), d on the circular
    coarse box (X points per axis: d in [-X/2, X/2)), scaled by 1 / (8 N_coarse)
    like the engine's tables; G is the fine real-space octant (evenness gives
    the other octants)."""
    Z1, Y1, X1 = G.shape
    X, Y, Z = X1 - 1, Y1 - 1, Z1 - 1
    d = [np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n) for n in (Z, Y, X)]
    dz, dy, dx = np.meshgrid(*d, indexing="ij")
    K = G[np.abs(2 * dz + p), np.abs(2 * dy + p), np.abs(2 * dx + p)]
    return np.fft.fftn(K) / (X * Y * Z)   # 8 N_coarse = 8 (X/2)(Y/2)(Z/2)


@pytest.mark.parametrize("dims,tf", [((12, 8, 10), None), ((8, 14, 6), None), ((10, 10, 16), 5.65)])
def test_coset_tables_are_coarse_transforms(dims, tf):
    X, Y, Z = dims
    G = O.gtable(X, Y, Z, (0.7, 0.9, 1.1), 0.31, 2.5, tf_lattice=tf)
    g0, gt = coarse_tables(O.ghat(G))
    want0 = sampled_transform(G, 0)
    want1 = sampled_transform(G, 1)
    scale = np.max(np.abs(want0))
    assert np.max(np.abs(g0 - want0)) < 1e-13 * scale
    assert np.max(np.abs(gt - want1)) < 1e-13 * scale
    # the symmetries the tiled tables rely on: G0 even and St odd about X/2 per axis
    st = gt * np.exp(-1j * np.pi * np.add.outer(np.add.outer(np.arange(Z) / Z, np.arange(Y) / Y), np.arange(X) / X))
    assert np.max(np.abs(g0[:, :, 1:] - g0[:, :, :0:-1])) < 1e-13 * scale
    assert np.max(np.abs(st[:, :, 1:] + st[:, :, :0:-1])) < 1e-13 * scale
    assert np.max(np.abs(st[:, 1:, :] + st[:, :0:-1, :])) < 1e-13 * scale


@pytest.mark.parametrize("dims", [(12, 8, 10), (8, 14, 6)])
def test_coset_tables_from_real_space_samples(dims):
    """The second route to the same tables, from the real-space fine table
    before its DCT (kernels.hpp: k_coset_gtable, k_dct12, k_coarse_q): the
    within-coset octant is DCT-I(G(2d)) / (8 N_coarse) on d in [0, X/2]^3, and
    the cross-coset one e^{i pi q.(1/X)} DCT-II(G(2d + 1)) / (8 N_coarse) on d in
    [0, X/2)^3, odd about X/2."""
    import scipy.fft as sf
    X, Y, Z = dims
    Xc, Yc, Zc = X // 2, Y // 2, Z // 2
    G = O.gtable(X, Y, Z, (0.75, 0.95, 1.05), 0.2, 4.0)
    g0, gt = coarse_tables(O.ghat(G))
    n8 = 8.0 * Xc * Yc * Zc
    g0_route = O.dct1_3d(G[0::2, 0::2, 0::2]) / n8
    st_route = sf.dctn(G[1:Z:2, 1:Y:2, 1:X:2], type=2) / n8
    qz, qy, qx = np.meshgrid(np.arange(Zc), np.arange(Yc), np.arange(Xc), indexing="ij")
    gt_route = np.exp(1j * np.pi * (qx / X + qy / Y + qz / Z)) * st_route
    scale = np.max(np.abs(g0_route))
    assert np.max(np.abs(g0[:Zc + 1, :Yc + 1, :Xc + 1] - g0_route)) < 1e-13 * scale
    assert np.max(np.abs(gt[:Zc, :Yc, :Xc] - gt_route)) < 1e-13 * scale


@pytest.mark.parametrize("dims", [(12, 8, 10), (16, 6, 8)])
def test_two_coset_convolution_matches_oracle(dims):
    """Random complex density on x = y = z (mod 2), some sites empty: the
    coupled coarse convolution equals the oracle's eight-round prepare at every
    site of both cosets.  A dropped cross term, a conjugate on the wrong side
    or a coset shift of the wrong sign fails it."""
    X, Y, Z = dims
    rng = np.random.default_rng(X * 100 + Y + Z)
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    on = ((xx % 2) == (yy % 2)) & ((yy % 2) == (zz % 2)) & (rng.random((Z, Y, X)) < 0.7)
    F = np.zeros((Z, Y, X), complex)
    F[on] = rng.standard_normal(on.sum()) + 1j * rng.standard_normal(on.sum())
    Gs = O.ghat(O.gtable(X, Y, Z, (0.8, 1.0, 1.2), 0.5, 3.0))
    Vo = O.prepare(F, Gs)
    g0, gt = coarse_tables(Gs)
    Xc, Yc, Zc = X // 2, Y // 2, Z // 2

    def fwd(r):
        P = np.zeros((Z, Y, X), complex)   # coarse box: 2 Xc = X points per axis
        P[:Zc, :Yc, :Xc] = r
        return np.fft.fftn(P)

    def inv(P):  # unnormalised inverse (the tables carry the 1 / (8 N_coarse))
        return np.fft.ifftn(P)[:Zc, :Yc, :Xc] * (X * Y * Z)

    r0, r1 = fwd(F[0::2, 0::2, 0::2]), fwd(F[1::2, 1::2, 1::2])
    v0 = inv(g0 * r0 + np.conj(gt) * r1)
    v1 = inv(gt * r0 + g0 * r1)
    scale = np.max(np.abs(Vo))
    assert np.max(np.abs(v0 - Vo[0::2, 0::2, 0::2])) < 1e-12 * scale
    assert np.max(np.abs(v1 - Vo[1::2, 1::2, 1::2])) < 1e-12 * scale
    # the cross terms carry real weight here: without them coset 1 is far off
    v1_alone = inv(g0 * r1)
    assert np.max(np.abs(v1_alone - Vo[1::2, 1::2, 1::2])) > 1e-3 * scale


@pytest.mark.parametrize("dims", [(16, 12, 20), (8, 8, 12)])
def test_half_box_matches_oracle(dims):
    """Sites on the coarse grid's fcc sublattice (mx + my + mz even, in both
    cosets): the density's coarse transform is periodic under the half-box
    shift s = (Xc, Yc, Zc), so carrying only kx < Xc against the kernels summed
    over the shift, K(k) + K(k + s) (kernels.hpp: k_gconv_tiled_half,
    k_coarse_q_half), gives the potential at every site.  One site with an
    odd coarse sum breaks it (why the engine checks every position)."""
    X, Y, Z = dims
    Xc, Yc, Zc = X // 2, Y // 2, Z // 2
    rng = np.random.default_rng(X + 7 * Y + 13 * Z)
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    p = xx % 2
    fcc = ((xx - p) // 2 + (yy - p) // 2 + (zz - p) // 2) % 2 == 0
    on = ((xx % 2) == (yy % 2)) & ((yy % 2) == (zz % 2)) & fcc & (rng.random((Z, Y, X)) < 0.8)
    Gs = O.ghat(O.gtable(X, Y, Z, (0.7, 0.9, 1.1), 0.3, 2.0))
    g0, gt = coarse_tables(Gs)

    def per(T):
        return T + np.roll(T, (-Zc, -Yc, -Xc), axis=(0, 1, 2))

    def half_box(F):
        def fwd(r):
            P = np.zeros((Z, Y, X), complex)
            P[:Zc, :Yc, :Xc] = r
            return np.fft.fftn(P)

        def inv(P):
            Q = np.zeros((Z, Y, X), complex)
            Q[:, :, :Xc] = P[:, :, :Xc]
            return np.fft.ifftn(Q)[:Zc, :Yc, :Xc] * (X * Y * Z)

        r0, r1 = fwd(F[0::2, 0::2, 0::2]), fwd(F[1::2, 1::2, 1::2])
        return inv(per(g0) * r0 + per(np.conj(gt)) * r1), inv(per(gt) * r0 + per(g0) * r1)

    F = np.zeros((Z, Y, X), complex)
    F[on] = rng.standard_normal(on.sum()) + 1j * rng.standard_normal(on.sum())
    Vo = O.prepare(F, Gs)
    v0, v1 = half_box(F)
    m0, m1 = on[0::2, 0::2, 0::2], on[1::2, 1::2, 1::2]
    scale = np.max(np.abs(Vo))
    assert np.max(np.abs(v0 - Vo[0::2, 0::2, 0::2])[m0]) < 1e-12 * scale
    assert np.max(np.abs(v1 - Vo[1::2, 1::2, 1::2])[m1]) < 1e-12 * scale
    # one extra site at an odd coarse sum (coarse (1, 0, 0) of coset 0)
    F[0, 0, 2] = 1.0
    Vo = O.prepare(F, Gs)
    v0, _ = half_box(F)
    assert np.max(np.abs(v0 - Vo[0::2, 0::2, 0::2])[m0]) > 1e-6 * scale


def test_half_length_lines():
    """The x passes on half-length lines (linepass.hpp, HL): on the coarse
    grid's fcc sublattice every site of a line (y, z) has x = 2j + p with
    p = (y + z) mod 2, so the stored frequencies f < X of the line's 2X-point
    transform are W_2X^{f p} times the 2(X/2)-point transform of the X/2
    points j, and the inverse of frequencies f < X at x = 2j + p is the
    X/2-point line's inverse of the spectrum times conj(W_2X^{f p}).  Checked
    on random lines of every parity, and the parity rule on zincblende's own
    coarse positions; the other parity's phase fails."""
    import zincblende as zb
    rng = np.random.default_rng(5)
    for X in (16, 20, 64):
        H = X // 2
        W2X = np.exp(-2j * np.pi * np.arange(X) / (2 * X))   # W_2X^f, f < X
        f = np.arange(X)
        for p in (0, 1):
            u = np.zeros(X, complex)
            u[p::2] = rng.standard_normal(H) + 1j * rng.standard_normal(H)
            full = np.fft.fft(np.concatenate([u, np.zeros(X)]))[:X]           # 2X-point, f < X
            half = W2X ** p * np.fft.fft(np.concatenate([u[p::2], np.zeros(H)]))  # 2H = X points
            assert np.max(np.abs(half - full)) < 1e-12 * np.max(np.abs(full))
            wrong = W2X ** (1 - p) * np.fft.fft(np.concatenate([u[p::2], np.zeros(H)]))
            assert np.max(np.abs(wrong - full)) > 1e-3 * np.max(np.abs(full))
            S = rng.standard_normal(X) + 1j * rng.standard_normal(X)            # frequencies f < X
            inv_full = np.fft.ifft(np.concatenate([S, np.zeros(X)])) * (2 * X)  # unnormalised, 2X points
            inv_half = np.fft.ifft(S * np.conj(W2X ** p)) * X                   # X = 2H points, first H
            assert np.max(np.abs(inv_half[:H] - inv_full[p:X:2])) < 1e-12 * np.max(np.abs(inv_full))
        assert np.allclose(W2X[f] ** 2, np.exp(-2j * np.pi * f / X))
    # the parity rule on zincblende's coarse positions: both cosets' sites
    # (2m and 2m + (1,1,1)) have m_x = m_y + m_z (mod 2)
    idx = zb.grid_indices(4, 6, 8)
    assert np.all((idx[:, 0] % 2 == idx[:, 1] % 2) & (idx[:, 1] % 2 == idx[:, 2] % 2))
    m = idx // 2
    assert np.all(m[:, 0] % 2 == (m[:, 1] + m[:, 2]) % 2)
