"""Per-integral parity bars shared by the GPU tests.

The bars are relative to each integral on its own, with no floor:

* FP64 path: every integral within 1e-10 of the FP64 oracle / direct sum;
* FP32X path (complex64 passes, FP64 last pass, FP64 overlap sums): every
  integral within 1e-5 of the FP64 result.

The pure-FP32 mode (every pass in FP32, the last one included) is not held to
the per-integral bar: its rounding error is absolute, of the order of 1e-7 of
the potential's magnitude, so an integral that cancels down to 1e-4 of the
run's largest can lose its relative accuracy.  Its stated guarantee is 1e-5
relative to max(|I|, 1e-3 max|I|) over the run, and that is what it is
checked against.
"""
import numpy as np

import clat

TOL = {clat.PREC_FP64: 1e-10, clat.PREC_FP32: 1e-5, clat.PREC_FP32X: 1e-5}


def rel_err(got, want):
    """|got - want| / |want| per integral; an exact zero in want counts as
    an error unless got is exactly zero there too."""
    got = np.asarray(got, np.complex128)
    want = np.asarray(want, np.complex128)
    diff = np.abs(got - want)
    den = np.abs(want)
    out = np.full(diff.shape, np.inf)
    nz = den > 0
    out[nz] = diff[nz] / den[nz]
    out[~nz & (diff == 0)] = 0.0
    return out


def floored_err(got, want, floor=1e-3):
    want = np.asarray(want, np.complex128)
    scale = np.max(np.abs(want)) if want.size else 0.0
    return np.abs(np.asarray(got) - want) / np.maximum(np.abs(want), floor * scale)


def worst(got, want, quads=None, k=5):
    """Text naming the k worst integrals: quad, value, |I|/max|I|, error."""
    e = rel_err(got, want)
    want = np.asarray(want)
    scale = np.max(np.abs(want)) if want.size else 1.0
    lines = []
    for i in np.argsort(-e)[:k]:
        q = tuple(int(v) for v in quads[i]) if quads is not None else i
        lines.append("%s: want %.6e%+.6ej (|I|/max %.2e) got %.6e%+.6ej rel %.2e"
                     % (q, want[i].real, want[i].imag, abs(want[i]) / scale, got[i].real, got[i].imag, e[i]))
    return "\n".join(lines)


def check(got, want, prec, quads=None, tol=None):
    """Assert the bar of `prec` (see the module text); returns the max error."""
    tol = TOL[prec] if tol is None else tol
    if prec == clat.PREC_FP32:
        err = float(np.max(floored_err(got, want))) if len(want) else 0.0
    else:
        err = float(np.max(rel_err(got, want))) if len(want) else 0.0
    assert err < tol, "max rel err %.3e >= %.1e\n%s" % (err, tol, worst(got, want, quads))
    return err


# Complex64 storage of the pair densities bounds what any complex64 FFT path
# can reach: rounding a density to complex64 once, everything else exact,
# moves an integral that cancels to ~1e-8 of the run's largest by more than
# 1e-5 of itself (tests/test_oracle.py::test_complex64_density_rounding_floor).
# For the FP32X path on general workloads the bar is therefore per integral
#   |dI| <= 1e-5 |I|                      for |I| >= 1e-6 max|I|,
#   |dI| <= 1e-11 max|I|  (and listed)    for |I| <  1e-6 max|I|.
SMALL = 1e-6
SMALL_ABS = 1e-11


def check_fp32x_model(got, want, quads=None, label=""):
    """The FP32X bar on a general workload (see above).  Prints every integral
    that the absolute clause admits past 1e-5 relative, by name, with its
    magnitude and error.  Returns (max rel err over |I| >= SMALL max, count
    of integrals admitted by the absolute clause)."""
    got = np.asarray(got, np.complex128)
    want = np.asarray(want, np.complex128)
    scale = np.max(np.abs(want))
    mag = np.abs(want) / scale
    d = np.abs(got - want)
    e = rel_err(got, want)
    big = mag >= SMALL
    err_big = float(np.max(e[big])) if big.any() else 0.0
    assert err_big < 1e-5, "%s: max rel err %.3e over |I| >= %.0e max\n%s" % (
        label, err_big, SMALL, worst(got[big], want[big], None if quads is None else np.asarray(quads)[big]))
    small = ~big
    assert np.all(d[small] <= SMALL_ABS * scale), "%s: absolute error %.3e of max on a small integral" % (
        label, float(np.max(d[small]) / scale))
    past = np.nonzero(small & (e >= 1e-5))[0]
    for i in past:
        q = tuple(int(v) for v in quads[i]) if quads is not None else int(i)
        print("%s: %s |I| = %.3e (%.2e of max) rel err %.2e, abs err %.2e of max (complex64 density rounding)"
              % (label, q, abs(want[i]), mag[i], e[i], d[i] / scale))
    return err_big, len(past)
