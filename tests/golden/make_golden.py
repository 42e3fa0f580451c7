"""Writes the committed golden fixtures of this directory.

zincblende_direct.json -- Coulomb integrals of small seeded zincblende boxes
by the direct O(N^2) lattice sum in FP64 (no FFT), from the definition
    I(i1,i2,i3,i4) = sum_a sum_b conj(f_i1) f_i4 (a) G(|r_a - r_b|) conj(f_i2) f_i3 (b)
(README.md:136-138), with G(r) = 14.39963737103201/(eps r), G(0) = onsite.
The inputs are regenerated from their seeds by zincblende.workload(); a
checksum of the coefficients guards against generator drift.

dct_fft_vectors.npz -- DCT-I (REDFT00) and c2c FFT vectors from scipy.fft
at sizes 8, 12, 45, 64, 96, 120 on seeded inputs.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "coulombo-lattice_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
import zincblende as zb  # noqa: E402

CASES = [
    # name, cells, holes, electrons, config seed, sigma_frac, orbitals, dielectric, onsite, patterns
    ("zb2_2h2e", (2, 2, 2), 2, 2, 0, 0.2, 20, 10.89, 0.0, "****"),
    ("zb4_2h2e_onsite", (4, 4, 4), 2, 2, 7, 0.2, 20, 10.89, 1.5, "eeee,hhhh,ehhe,eheh,1234,4321,1*2*"),
    ("config0_8cells_2h2e", (8, 8, 8), 2, 2, 0, 0.2, 20, 10.89, 0.0, "eeee,hhhh,ehhe,eheh"),
]


def coeff_digest(states):
    h = hashlib.sha256()
    for s in states:
        h.update(np.ascontiguousarray(s, np.complex128).tobytes())
    return h.hexdigest()


def direct_integrals(w, quads, dielectric, onsite):
    X, Y, Z = w["dims"]
    G = O.gtable(X, Y, Z, w["step"], onsite, dielectric)
    cells = w["cells"]
    st = w["states"]
    cache = {}

    def V_of(i, j):  # potential of conj(f_i) f_j at the atom cells
        if (i, j) not in cache:
            F = np.zeros(X * Y * Z, np.complex128)
            F[cells] = np.sum(np.conj(st[i - 1]) * st[j - 1], axis=1)
            cache[(i, j)] = O.direct_potential(F.reshape(Z, Y, X), G, cells)
        return cache[(i, j)]

    out = []
    for (i1, i2, i3, i4) in quads:
        rR = np.sum(np.conj(st[i2 - 1]) * st[i3 - 1], axis=1)
        out.append(complex(np.sum(V_of(i1, i4) * rR)))
    return out


def main():
    cases = []
    for name, cells, nh, ne, cfg, sf, norb, eps, onsite, pat in CASES:
        w = zb.workload(*cells, nh, ne, config=cfg, sigma_frac=sf, n_orb=norb)
        quads = O.pattern_enumerate(pat, nh + ne, nh)
        vals = direct_integrals(w, quads, eps, onsite)
        cases.append(dict(name=name, cells=list(cells), holes=nh, electrons=ne, config=cfg, sigma_frac=sf,
                          orbitals=norb, dielectric=eps, onsite=onsite, integrals=pat,
                          dims=list(w["dims"]), coeff_sha256=coeff_digest(w["states"]),
                          quads=[list(q) for q in quads], values=[[v.real, v.imag] for v in vals]))
        print(name, len(quads), "integrals", flush=True)
    with open(os.path.join(os.path.dirname(__file__), "zincblende_direct.json"), "w") as f:
        json.dump(dict(source="direct O(N^2) FP64 lattice sum (oracle.c orc_direct_potential)",
                       cases=cases), f, indent=1)

    import scipy.fft as sf
    rng = np.random.default_rng(2024)
    vec = {}
    for n in (8, 12, 45, 64, 96, 120):
        x = rng.standard_normal(n + 1)
        vec["dct_in_%d" % n] = x
        vec["dct_out_%d" % n] = sf.dct(x, type=1, norm=None)
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        vec["fft_in_%d" % n] = z
        vec["fft_out_%d" % n] = sf.fft(z)
        vec["ifft_out_%d" % n] = sf.ifft(z) * n
    g = rng.standard_normal((6, 5, 9))
    vec["dct3_in"] = g
    vec["dct3_out"] = sf.dctn(g, type=1, norm=None)
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "dct_fft_vectors.npz"), **vec)
    print("dct/fft vectors written")


if __name__ == "__main__":
    main()
