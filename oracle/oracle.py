"""TEST INFRASTRUCTURE ONLY -- the FP64 CPU oracle of the FFT Coulomb-integral path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product library never calls it.

Numerics live in oracle.c (built into oracle/liboracle.so); this module wraps
them with numpy and restates the host-side integer logic in plain Python:

    pattern_enumerate   src/Pattern.hpp:13-101 + coulombo.cpp:135-148
    vertex_cover        src/Graph.hpp:49-91
    plan                src/Planner.hpp:65-145 (MasterPlanner)
    compute_integrals   coulombo.cpp:150-209 (main loop, conj bookkeeping)
    export_rows         coulombo.cpp:213-262 (file names, numbering, format)

Each function's header names the reference lines it restates.  What pins each
one (golden vectors, the reference's own headers compiled into oracle/_ref,
closed forms, brute force) is listed in the tests that use it.
"""
from __future__ import annotations

import ctypes
import functools
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
E2_4PE0 = 14.39963737103201  # src/base.hpp:28

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so (and oracle/_ref where the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


@functools.lru_cache(maxsize=None)
def lib() -> ctypes.CDLL:
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.orc_round_up.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int]
    L.orc_round_up.restype = ctypes.c_int
    L.orc_lattice_grid.argtypes = [_dp, ctypes.c_int64, _ip, _ip, _dp, _dp]
    L.orc_lattice_grid.restype = ctypes.c_int
    L.orc_lattice_cells.argtypes = [_dp, ctypes.c_int64, _ip, _dp, _dp, _ip]
    L.orc_gtable.argtypes = [ctypes.c_int] * 3 + [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                  ctypes.c_double, _dp]
    L.orc_dct1_3d.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_fft.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
    L.orc_prepare.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
    L.orc_calculate.argtypes = [_dp, _dp, ctypes.c_int64, _dp]
    L.orc_products.argtypes = [_dp, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _dp]
    L.orc_direct_potential.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _ip,
                                       ctypes.c_int64, _dp]
    L.orc_num_threads.restype = ctypes.c_int
    return L


@functools.lru_cache(maxsize=None)
def ref_lib():
    """oracle/_ref/libref.so built from the reference's own headers, or None."""
    path = os.path.join(HERE, "_ref", "libref.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.ref_round_up_2.argtypes = [ctypes.c_int, ctypes.c_int]
    L.ref_round_up_235.argtypes = [ctypes.c_int, ctypes.c_int]
    L.ref_vertex_cover.argtypes = [ctypes.c_int, _ip, ctypes.c_int64, _ip]
    L.ref_pattern_enumerate.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int64,
                                        ctypes.c_char_p, ctypes.c_int]
    L.ref_pattern_enumerate.restype = ctypes.c_int64
    return L


def _c(a: np.ndarray) -> np.ndarray:
    """complex128 array -> contiguous interleaved float64 view"""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64).reshape(-1)


# ------------------------------------------------------------------ lattice
def round_up(n: int, c: int = 1, primes=(2, 3, 5)) -> int:
    """Round::up<primes...>(n, c) -- src/Round.hpp:20-27,38-49"""
    return lib().orc_round_up(n, c, np.array(primes, np.int32), len(primes))


def lattice_grid(xyz: np.ndarray):
    """Broadcaster.cpp:8-43,53-56,136-143 -> (raw, pad, origin, step)"""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    raw = np.zeros(3, np.int32); pad = np.zeros(3, np.int32)
    origin = np.zeros(3); step = np.zeros(3)
    if lib().orc_lattice_grid(xyz, len(xyz) // 3, raw, pad, origin, step):
        raise ValueError("could not find common grid step")
    return raw, pad, origin, step


def lattice_cells(xyz, dim, origin, step) -> np.ndarray:
    """Broadcaster.cpp:76-88 at one rank"""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    out = np.zeros(len(xyz) // 3, np.int32)
    lib().orc_lattice_cells(xyz, len(out), np.asarray(dim, np.int32), np.asarray(origin, np.float64),
                            np.asarray(step, np.float64), out)
    return out


# -------------------------------------------------------------- interaction
def gtable(X, Y, Z, step, onsite=0.0, dielectric=1.0, tf_lattice=None) -> np.ndarray:
    """InteractionBase::map on the (X+1)(Y+1)(Z+1) first octant, shape (Z+1, Y+1, X+1)."""
    G = np.zeros((Z + 1) * (Y + 1) * (X + 1))
    model = 0 if tf_lattice is None else 1
    lib().orc_gtable(X + 1, Y + 1, Z + 1, np.asarray(step, np.float64), onsite, dielectric, model,
                     0.0 if tf_lattice is None else float(tf_lattice), G)
    return G.reshape(Z + 1, Y + 1, X + 1)


def dct1_3d(G: np.ndarray) -> np.ndarray:
    """REDFT00 on every axis of a (Z1, Y1, X1) table (unnormalised)."""
    out = np.ascontiguousarray(G, np.float64).copy()
    Z1, Y1, X1 = out.shape
    lib().orc_dct1_3d(out.reshape(-1), X1, Y1, Z1)
    return out


def ghat(G: np.ndarray) -> np.ndarray:
    """CoulombCalculator::initialize at one rank: DCT-I / (8 X Y Z)  (CoulombCalculator.cpp:57-103)"""
    Z1, Y1, X1 = G.shape
    return dct1_3d(G) / (8.0 * (X1 - 1) * (Y1 - 1) * (Z1 - 1))


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """1-D unnormalised DFT, FFTW sign convention (sign -1 forward)."""
    a = np.ascontiguousarray(x, np.complex128).copy()
    lib().orc_fft(a.view(np.float64).reshape(-1), len(a), sign)
    return a


def prepare(F: np.ndarray, Gh: np.ndarray) -> np.ndarray:
    """CoulombCalculator::prepare (CoulombCalculator.cpp:105-163); F shape (Z, Y, X)."""
    Z, Y, X = F.shape
    V = np.zeros(F.shape, np.complex128)
    lib().orc_prepare(_c(F), X, Y, Z, np.ascontiguousarray(Gh, np.float64).reshape(-1),
                      V.view(np.float64).reshape(-1))
    return V


def calculate(V: np.ndarray, F: np.ndarray) -> complex:
    """CoulombCalculator::calculate (CoulombCalculator.cpp:165-172), compensated."""
    out = np.zeros(2)
    lib().orc_calculate(_c(V), _c(F), V.size, out)
    return complex(out[0], out[1])


def products(coeffs: np.ndarray) -> np.ndarray:
    """All pair densities, shape (P, n_atoms); coeffs shape (n_states, n_atoms, n_orb)."""
    n, na, no = coeffs.shape
    P = n * (n + 1) // 2
    out = np.zeros((P, na), np.complex128)
    lib().orc_products(_c(coeffs), n, na, no, out.view(np.float64).reshape(-1))
    return out


def direct_potential(F: np.ndarray, G: np.ndarray, cells: np.ndarray) -> np.ndarray:
    """V at the listed cells by the direct O(N^2) lattice sum (G untransformed)."""
    Z, Y, X = F.shape
    cells = np.ascontiguousarray(cells, np.int32)
    out = np.zeros(len(cells), np.complex128)
    lib().orc_direct_potential(_c(F), X, Y, Z, np.ascontiguousarray(G, np.float64).reshape(-1), cells,
                               len(cells), out.view(np.float64).reshape(-1))
    return out


def num_threads() -> int:
    return lib().orc_num_threads()


# ------------------------------------------------------------ host logic
def pattern_enumerate(spec: str, n_states: int, n_holes: int) -> list:
    """Pattern (src/Pattern.hpp:13-101) + the n^4 loop of coulombo.cpp:135-148."""
    pats = spec.split(",")
    for p in pats:
        if not p:
            raise ValueError("invalid empty pattern")
        for ch in p:
            if not (ch == "*" or (ch.isascii() and ch.isalnum() and ch != "0")):
                raise ValueError("invalid character '%s' in pattern" % ch)

    def letter(c, i, assign):
        if c == "*":
            return True
        if "1" <= c <= "9":
            return i == ord(c) - 48
        if c == "h":
            return i <= n_holes
        if c == "e":
            return i > n_holes
        if c in assign:
            return i == assign[c]
        assign[c] = i
        return True

    def match(p, q):
        assign = {}
        for k in range(4):
            c = p[k] if k < len(p) else "\0"
            if not letter(c, q[k], assign):
                return False
        return True

    out = []
    r = range(1, n_states + 1)
    for i1 in r:
        for i2 in r:
            for i3 in r:
                for i4 in r:
                    q = (i1, i2, i3, i4)
                    if any(match(p, q) for p in pats):
                        out.append(q)
    return out


def vertex_cover(nv: int, edges) -> list:
    """Graph::computeVertexCover (src/Graph.hpp:49-91)."""
    es = sorted(set((int(u), int(v)) for u, v in edges))
    cover = {u for u, v in es if u == v}
    nb = [set() for _ in range(nv)]
    for u, v in es:
        if u != v and u not in cover and v not in cover:
            nb[u].add(v)
            nb[v].add(u)
    key = lambda v: (len(nb[v]), v)  # noqa: E731
    queue = {v for v in range(nv) if nb[v]}
    while queue:
        first = min(queue, key=key)
        v = next(iter(nb[first])) if len(nb[first]) == 1 else max(queue, key=key)
        queue.discard(v)
        cover.add(v)
        for w in sorted(nb[v]):
            if w in queue:
                queue.discard(w)
                nb[w].discard(v)
                if nb[w]:
                    queue.add(w)
    return sorted(cover)


def get_product(iL: int, iR: int):
    """Planner.hpp:134-144 (1-based states)"""
    conj = 0
    if iL < iR:
        conj = 1
        iL, iR = iR, iL
    return (iL * (iL - 1) // 2 + iR - 1, conj)


def plan(quads, n_states: int) -> list:
    """MasterPlanner::addIntegral/computePlan -> [(id, (L, lc), (R, rc))] in execution order."""
    steps = [(k, get_product(q[0], q[3]), get_product(q[1], q[2])) for k, q in enumerate(quads)]
    P = n_states * (n_states + 1) // 2
    cover = set(vertex_cover(P, [(s[1][0], s[2][0]) for s in steps]))
    steps = [(k, l, r) if l[0] in cover else (k, r, l) for k, l, r in steps]
    # sort key of Planner.hpp:94-101: (left, right, conj-class) with rc==0 first
    steps.sort(key=lambda s: (s[1][0], s[2][0], 0 if s[1][1] == s[2][1] else 1))
    return steps


def compute_integrals(quads, n_states, rho_sites, sites_cells, dims, Gh, stale=False):
    """coulombo.cpp:186-209 on the 8-round oracle.

    rho_sites: (P, n_sites) pair densities at effective sites; sites_cells:
    linear cells of those sites on the (X, Y, Z) grid.  Returns complex values
    in quad order.
    """
    X, Y, Z = dims
    out = np.zeros(len(quads), np.complex128)
    lastL, lastR, lastRc = -1, -1, 0
    value = 0j
    V = None
    for k, (L, lc), (R, rcj) in plan(quads, n_states):
        if L != lastL:
            F = np.zeros(X * Y * Z, np.complex128)
            F[sites_cells] = rho_sites[L]
            V = prepare(F.reshape(Z, Y, X), Gh).reshape(-1)[sites_cells]
            lastL = L
            if not stale:
                lastR = -1
        rc = 1 if lc != rcj else 0
        if R != lastR or rc != lastRc:
            r = np.conj(rho_sites[R]) if rc else rho_sites[R]
            value = complex(np.sum(V * r))
            lastR, lastRc = R, rc
        out[k] = np.conj(value) if lc else value
    return out


def export_rows(quads, values, n_holes: int, n_elec: int) -> dict:
    """coulombo.cpp:213-262: {'eeee': text, ...} for files holding >= 1 integral."""
    vals = {tuple(int(x) for x in q): v for q, v in zip(quads, values)}
    files = {}
    for ti in (0, 1):
        for tj in (0, 1):
            for tk in (0, 1):
                for tl in (0, 1):
                    core = "".join("e" if t else "h" for t in (ti, tj, tk, tl))
                    counts = [n_elec if t else n_holes for t in (ti, tj, tk, tl)]
                    lines = []
                    for ni in range(1, counts[0] + 1):
                        for nj in range(1, counts[1] + 1):
                            for nk in range(1, counts[2] + 1):
                                for nl in range(1, counts[3] + 1):
                                    spec = tuple((n_holes + n) if t else (n_holes + 1 - n)
                                                 for t, n in zip((ti, tj, tk, tl), (ni, nj, nk, nl)))
                                    if spec in vals:
                                        v = vals[spec]
                                        lines.append("%2d %2d %2d %2d   %17.14f %17.14f\n"
                                                     % (ni, nj, nk, nl, v.real, v.imag))
                    if lines:
                        files[core] = "".join(lines)
    return files
