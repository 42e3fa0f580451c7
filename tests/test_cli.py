"""The coulombo / potentials drivers: option handling, file contract and error
messages of the reference (coulombo.cpp:30-315, potentials.cpp:30-207).

The CPU tests stop before any GPU work (every failure they provoke happens
while reading options or files); the GPU tests run the example and a seeded
zincblende case end to end and compare the output files."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import zincblende as zb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "coulombo-lattice_amd", "bin")
EXAMPLE = os.path.join(ROOT, "tests", "golden", "example")


def run(args, cwd, tool="coulombo"):
    p = subprocess.run([os.path.join(BIN, tool)] + args, cwd=cwd, capture_output=True, text=True, timeout=600)
    return p.returncode, p.stdout, p.stderr


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    d = tmp_path_factory.mktemp("zb")
    w = zb.workload(2, 2, 2, 2, 2, config=0, sigma_frac=0.2, n_orb=4)
    zb.write_files(str(d), w)
    return str(d), w


def test_usage_without_inputs(tmp_path):
    rc, _, err = run([], str(tmp_path))
    assert rc == 1 and "USAGE:" in err and "--atoms=PATH" in err


@pytest.mark.parametrize("args,msg", [
    (["--bogus=1", "e1.dat"], "ERROR: unknown value 'bogus'"),
    (["--atoms", "e1.dat"], "ERROR: missing value for parameter 'atoms'"),
    (["--flag", "e1.dat"], "ERROR: unknown flag 'flag'"),
    (["-x", "e1.dat"], "ERROR: unknown flag 'x'"),
    (["e1.dat"], "ERROR: parameter 'atoms' not specified"),
    (["--atoms=zb.3d", "--orbitals=0", "e1.dat"], "ERROR: invalid value for orbitals"),
    (["--atoms=zb.3d", "--skip-lines=0", "e1.dat"], "ERROR: invalid value for skip-lines"),
    (["--atoms=zb.3d", "--threads-per-node=0", "e1.dat"], "ERROR: invalid value for threads-per-node"),
    (["--atoms=missing.3d", "--orbitals=4", "e1.dat"], "ERROR: could not open file: missing.3d"),
    (["--atoms=zb.3d", "--orbitals=4", "e1.dat", "h1.dat"], "ERROR: hole states must appear before electron states"),
    (["--atoms=zb.3d", "--orbitals=4", "x1.dat"], "ERROR: invalid state file name: x1.dat"),
    (["--atoms=zb.3d", "--orbitals=5", "e1.dat"], "ERROR: file is truncated: e1.dat"),
    # a file before a bad name fails first, as in the reference's sequential read
    (["--atoms=zb.3d", "--orbitals=5", "h1.dat", "x1.dat"], "ERROR: file is truncated: h1.dat"),
    (["--atoms=zb.3d", "--orbitals=4", "h1.dat", "e1.dat", "h2.dat"],
     "ERROR: hole states must appear before electron states"),
    (["--atoms=zb.3d", "--orbitals=4", "--integrals=e0ee", "e1.dat"], "ERROR: invalid character '0' in pattern"),
    (["--atoms=zb.3d", "--orbitals=4", "--integrals=eeee,", "e1.dat"], "ERROR: invalid empty pattern"),
])
def test_error_messages(small, args, msg):
    d, _ = small
    if "x1.dat" in args:
        os.symlink(os.path.join(d, "e1.dat"), os.path.join(d, "x1.dat")) if not os.path.exists(
            os.path.join(d, "x1.dat")) else None
    rc, _, err = run(args, d)
    assert rc == 1
    assert msg in err, err


def test_non_lattice_geometry(tmp_path):
    (tmp_path / "bad.3d").write_text("0 0 0\n1 0 0\n2.005 0 0\n")
    (tmp_path / "e1.dat").write_text("1 0\n" * 3)
    rc, _, err = run(["--atoms=bad.3d", "--orbitals=1", "e1.dat"], str(tmp_path))
    assert rc == 1 and "ERROR: could not find common grid step" in err


@pytest.mark.parametrize("trailing_newline", [True, False])
def test_large_file_line_count(tmp_path, trailing_newline):
    """Multi-megabyte inputs are split into lines by several threads: the
    line count must come out exact (one line short is 'truncated', the exact
    count is accepted), with or without a final newline."""
    w = zb.workload(16, 16, 8, 0, 1, config=0, sigma_frac=0.2, n_orb=4)
    zb.write_files(str(tmp_path), w)
    lines = (tmp_path / "e1.dat").read_text().splitlines()
    assert len(lines) == 4 * len(w["cells"]) > 60000
    body = "\n".join(lines[:-1]) + ("\n" if trailing_newline else "")
    (tmp_path / "e2.dat").write_text(body)
    rc, _, err = run(["--atoms=zb.3d", "--orbitals=4", "--threads-per-node=8", "e2.dat"], str(tmp_path))
    assert rc == 1 and "ERROR: file is truncated: e2.dat" in err, err
    (tmp_path / "e3.dat").write_text("\n".join(lines) + ("\n" if trailing_newline else ""))
    rc, _, err = run(["--atoms=zb.3d", "--orbitals=4", "--threads-per-node=8", "e3.dat"], str(tmp_path))
    assert "Reading wavefunctions... (" in err and "truncated" not in err, err


def test_stage_messages(small):
    """Timer lines in the reference's format: '<stage>... (x.xxx s)'."""
    d, _ = small
    rc, _, err = run(["--atoms=zb.3d", "--orbitals=4", "--integrals=e0ee", "h1.dat", "h2.dat", "e1.dat"], d)
    assert "Reading atom positions... (" in err and "Reading wavefunctions... (" in err


@pytest.mark.gpu
def test_example_end_to_end(tmp_path):
    """example/run.sh through the drop-in driver."""
    print("case cli example", flush=True)
    rc, _, err = run(["--atoms=" + os.path.join(EXAMPLE, "X.3d"), "--dielectric=11.4", "--orbitals=10",
                      "--skip-lines=1", os.path.join(EXAMPLE, "e1.dat")], str(tmp_path))
    assert rc == 0, err
    text = (tmp_path / "eeee.txt").read_text()
    assert text[:29] == " 1  1  1  1    0.090340071977"
    assert "Computing requested integrals... (" in err
    assert sorted(os.listdir(tmp_path)) == ["eeee.txt"]


# %17.14f prints 14 decimals: a printed value is within half a unit of the
# 14th decimal of the computed one
PRINT_HALF_ULP = 0.5e-14 * 1.0001


def _rows(path):
    """(ni, nj, nk, nl, value) of every row of an output file."""
    out = []
    for line in path.read_text().splitlines():
        f = line.split()
        out.append((int(f[0]), int(f[1]), int(f[2]), int(f[3]), complex(float(f[4]), float(f[5]))))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32x"])
def test_files_match_oracle_export(small, tmp_path, precision):
    """All 16 output files of --integrals=**** on 2 holes + 2 electrons:
    same file set, same row layout (per-type numbering, holes reversed), and
    every value within the precision's bar of the oracle's (relative to the
    value itself, plus the half unit of the printed 14th decimal)."""
    print("case cli zincblende", precision, flush=True)
    d, w = small
    out = tmp_path / "out"
    out.mkdir()
    rc, _, err = run(["--atoms=" + os.path.join(d, "zb.3d"), "--orbitals=4", "--dielectric=10.89",
                      "--integrals=****", "--precision=" + precision, "--output-dir=" + str(out),
                      "h1.dat", "h2.dat", "e1.dat", "e2.dat"], d)
    assert rc == 0, err
    X, Y, Z = w["dims"]
    quads = O.pattern_enumerate("****", 4, 2)
    rho = O.products(np.stack(w["states"]))
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    vals = O.compute_integrals(quads, 4, rho, w["cells"], (X, Y, Z), Gh)
    want = O.export_rows(quads, vals, 2, 2)
    assert sorted(p[:-4] for p in os.listdir(out)) == sorted(want)
    tol = 1e-10 if precision == "fp64" else 1e-5
    for core, text in want.items():
        got = (out / (core + ".txt")).read_text().splitlines()
        exp = text.splitlines()
        assert len(got) == len(exp)
        for g, e in zip(got, exp):
            assert g[:11] == e[:11]  # the four state numbers
            gv = complex(float(g.split()[4]), float(g.split()[5]))
            ev = complex(float(e.split()[4]), float(e.split()[5]))
            assert abs(gv - ev) <= tol * abs(ev) + 2 * PRINT_HALF_ULP, (core, g, e)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32x"])
def test_export_against_direct_sum_fixture(tmp_path, precision, golden_dir):
    """The file contract pinned without the oracle's export restatement: the
    2-hole + 2-electron, 20-orbital zincblende fixture (zincblende_direct.json
    'zb2_2h2e', every integral by the direct O(N^2) lattice sum) through
    bin/coulombo with --integrals=****.  Every row of the 16 files is mapped
    back to its 1-based state numbers by the rule of coulombo.cpp:244-247
    (file letter e: n_holes + n; letter h: n_holes + 1 - n, holes reversed)
    and compared with the fixture's value for those states; every fixture
    integral appears exactly once.  In particular hhhh.txt's ' 1  1  1  1'
    row is the self-energy of the last h file on the command line."""
    import json
    print("case cli export vs direct sums", precision, flush=True)
    with open(os.path.join(golden_dir, "zincblende_direct.json")) as f:
        case = [c for c in json.load(f)["cases"] if c["name"] == "zb2_2h2e"][0]
    nh, ne = case["holes"], case["electrons"]
    assert (nh, ne, case["integrals"]) == (2, 2, "****")
    w = zb.workload(*case["cells"], nh, ne, config=case["config"], sigma_frac=case["sigma_frac"],
                    n_orb=case["orbitals"])
    zb.write_files(str(tmp_path), w)
    fixture = {tuple(q): complex(a, b) for q, (a, b) in zip(case["quads"], case["values"])}
    out = tmp_path / "out"
    out.mkdir()
    rc, _, err = run(["--atoms=zb.3d", "--orbitals=%d" % case["orbitals"], "--dielectric=%r" % case["dielectric"],
                      "--integrals=****", "--precision=" + precision, "--output-dir=" + str(out),
                      "h1.dat", "h2.dat", "e1.dat", "e2.dat"], str(tmp_path))
    assert rc == 0, err
    tol = 1e-10 if precision == "fp64" else 1e-5
    seen = set()
    for name in sorted(os.listdir(out)):
        letters = name[:-4]
        assert len(letters) == 4 and set(letters) <= {"h", "e"} and name.endswith(".txt")
        for row in _rows(out / name):
            spec = tuple(nh + n if t == "e" else nh + 1 - n for t, n in zip(letters, row[:4]))
            assert spec not in seen
            seen.add(spec)
            want = fixture[spec]
            assert abs(row[4] - want) <= tol * abs(want) + PRINT_HALF_ULP, (name, row, spec, want)
    assert seen == set(fixture)
    first = _rows(out / "hhhh.txt")[0]
    assert first[:4] == (1, 1, 1, 1)
    assert abs(first[4] - fixture[(2, 2, 2, 2)]) <= tol * abs(fixture[(2, 2, 2, 2)]) + PRINT_HALF_ULP


@pytest.mark.gpu
def test_potentials_tool(small, tmp_path):
    """potentials: one 'potential-<name>' per input, '%.12le' real part per atom."""
    print("case cli potentials", flush=True)
    d, w = small
    rc, _, err = run(["--atoms=" + os.path.join(d, "zb.3d"), "--orbitals=4", "--dielectric=10.89",
                      "--output-dir=" + str(tmp_path), "h1.dat", "e2.dat"], d, tool="potentials")
    assert rc == 0, err
    X, Y, Z = w["dims"]
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    for name, s in (("h1.dat", w["states"][0]), ("e2.dat", w["states"][3])):
        lines = (tmp_path / ("potential-" + name)).read_text().splitlines()
        assert len(lines) == len(w["cells"])
        F = np.zeros(X * Y * Z, complex)
        F[w["cells"]] = np.sum(np.abs(s) ** 2, axis=1)
        Vo = O.prepare(F.reshape(Z, Y, X), Gh).reshape(-1)[w["cells"]].real
        got = np.array([float(x) for x in lines])
        assert lines[0] == "%.12le" % got[0]
        assert np.max(np.abs(got - Vo)) <= 1e-10 * np.max(np.abs(Vo)) + 1e-12
