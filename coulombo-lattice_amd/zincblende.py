"""Synthetic zincblende workloads for the integral path.

Atoms sit at (4c + b) * a/4 for conventional cells c in [0,nx)x[0,ny)x[0,nz)
and the eight basis offsets b (fcc + fcc shifted by (1/4,1/4,1/4) a), so the
lattice grid has step a/4 and 4n points per axis.  Positions are sorted by
grid index (iz, iy, ix), the order a multi-rank reference run requires
(src/Broadcaster.cpp:81-86).

States: c[a, o] = (N(0,1) + i N(0,1)) * env(r_a), normalised to sum |c|^2 = 1,
drawn from numpy PCG64 with seed = 1000 * config + state number (1-based).
Electrons carry an s-like Gaussian envelope, holes a p-like one
((x - x0)/sigma_x times the Gaussian).  sigma = sigma_frac * box per axis; the
centre is the box centre moved by a seeded uniform offset of +-0.05 box.

Run as a script to write an --atoms file and h*/e* LCAO files:
    python zincblende.py OUTDIR NX NY NZ NHOLES NELEC [CONFIG] [SIGMA_FRAC]
"""
from __future__ import annotations

import os
import sys

import numpy as np

A_GAAS = 5.6533
BASIS = np.array([(0, 0, 0), (0, 2, 2), (2, 0, 2), (2, 2, 0),
                  (1, 1, 1), (1, 3, 3), (3, 1, 3), (3, 3, 1)], np.int64)


def grid_indices(nx: int, ny: int, nz: int) -> np.ndarray:
    """(n_atoms, 3) integer grid indices (ix, iy, iz), sorted by (iz, iy, ix)."""
    c = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    idx = (4 * c[:, None, :] + BASIS[None]).reshape(-1, 3)
    order = np.lexsort((idx[:, 0], idx[:, 1], idx[:, 2]))
    return idx[order]


def positions(nx: int, ny: int, nz: int, a: float = A_GAAS) -> np.ndarray:
    return grid_indices(nx, ny, nz) * (a / 4.0)


def cells_of(idx: np.ndarray, dims) -> np.ndarray:
    X, Y, _ = dims
    return ((idx[:, 2] * Y + idx[:, 1]) * X + idx[:, 0]).astype(np.int32)


BLOCK = 1 << 20          # atoms per independently seeded block (large boxes)
BIG = 2 * BLOCK          # boxes above this many atoms draw their normals in blocks


def _envelope(pos, r0, sig, hole):
    d = (pos - r0) / sig
    env = np.exp(-0.5 * np.sum(d * d, axis=1))
    return env * d[:, 0] if hole else env


def state(pos: np.ndarray, box, hole: bool, seed: int, sigma_frac: float, n_orb: int = 20) -> np.ndarray:
    """One state's coefficients, shape (n_atoms, n_orb), complex128.

    Up to BIG atoms one PCG64 stream (seed) draws the centre offset, then all
    real parts, then all imaginary parts.  Larger boxes (the 8M-atom grid)
    draw the offset from that stream and the normals in blocks of BLOCK atoms,
    block b from PCG64 seeded with (seed, b), on parallel threads; the values
    do not depend on the thread count."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, np.float64)
    r0 = box / 2 + rng.uniform(-0.05, 0.05, 3) * box
    sig = sigma_frac * box
    n = len(pos)
    if n <= BIG:
        env = _envelope(pos, r0, sig, hole)
        re = rng.standard_normal((n, n_orb))
        im = rng.standard_normal((n, n_orb))
        c = (re + 1j * im) * env[:, None]
        c /= np.sqrt(np.sum(np.abs(c) ** 2))
        return c
    from concurrent.futures import ThreadPoolExecutor
    c = np.empty((n, n_orb), np.complex128)
    nblk = (n + BLOCK - 1) // BLOCK

    def fill(b):
        lo, hi = b * BLOCK, min(n, (b + 1) * BLOCK)
        g = np.random.default_rng([seed, b])
        v = c[lo:hi]
        v.real = g.standard_normal((hi - lo, n_orb))
        v.imag = g.standard_normal((hi - lo, n_orb))
        v *= _envelope(pos[lo:hi], r0, sig, hole)[:, None]
        return float(np.sum(v.real ** 2) + np.sum(v.imag ** 2))

    workers = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(workers) as ex:
        norm2 = sum(ex.map(fill, range(nblk)))  # summed in block order
    scale = 1.0 / np.sqrt(norm2)

    def rescale(b):
        c[b * BLOCK:(b + 1) * BLOCK] *= scale

    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(rescale, range(nblk)))
    return c


def workload(nx, ny, nz, n_holes, n_elec, config=0, sigma_frac=0.2, n_orb=20, a=A_GAAS):
    """-> dict(idx, pos, dims, step, cells, states (list, holes first), names)"""
    idx = grid_indices(nx, ny, nz)
    pos = idx * (a / 4.0)
    dims = (4 * nx, 4 * ny, 4 * nz)
    box = (nx * a, ny * a, nz * a)
    states, names = [], []
    for k in range(n_holes + n_elec):
        hole = k < n_holes
        states.append(state(pos, box, hole, 1000 * config + k + 1, sigma_frac, n_orb))
        names.append(("h%d" % (k + 1)) if hole else ("e%d" % (k + 1 - n_holes)))
    return dict(idx=idx, pos=pos, dims=dims, step=(a / 4.0,) * 3, cells=cells_of(idx, dims),
                states=states, names=names)


def write_files(outdir: str, w: dict) -> list:
    os.makedirs(outdir, exist_ok=True)
    np.savetxt(os.path.join(outdir, "zb.3d"), w["pos"], fmt="%.10f")
    paths = []
    for name, c in zip(w["names"], w["states"]):
        p = os.path.join(outdir, name + ".dat")
        flat = c.reshape(-1)
        np.savetxt(p, np.stack([flat.real, flat.imag], 1), fmt="%.17g")
        paths.append(p)
    return paths


if __name__ == "__main__":
    if len(sys.argv) < 7:
        sys.stderr.write(__doc__)
        sys.exit(2)
    out, nx, ny, nz, nh, ne = sys.argv[1], *map(int, sys.argv[2:7])
    cfg = int(sys.argv[7]) if len(sys.argv) > 7 else 0
    sf = float(sys.argv[8]) if len(sys.argv) > 8 else 0.2
    write_files(out, workload(nx, ny, nz, nh, ne, cfg, sf))
