"""Python binding of the engine's C ABI (include/clat.h) through ctypes.

It mirrors the reference's own objects for the integral path so that callers
and tests read like the reference's code:

    CoulombCalculator   src/CoulombCalculator.hpp:28-94 (ctor, initialize,
                        input, prepare, calculate, potential)
    InteractionSimple / InteractionThomasFermi / InteractionTable
                        src/Interaction.hpp:12-46 (the `map` of a kernel)
    Engine              FunctionCollection + Product + the coulombo main loop
                        (coulombo.cpp:150-209), batched on the GPU

The library is libclat.so next to this file; every call goes to the HIP
engine.  There is no CPU fallback: a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclat.so")

PREC_FP64, PREC_FP32, PREC_FP32X = 0, 1, 2
SCHED_SHARED, SCHED_ROUNDS = 0, 1
FLAG_STALE_VALUES = 1
# bits of clat_engine_kernels
KERNEL_FLAGS = {"two_level": 1, "column": 2, "line": 4, "conv_tiled": 8, "b_strips": 16, "ad_strips": 32,
                "cosets": 64, "half_box": 128, "half_lines": 256}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64

# every symbol include/clat.h declares, for the export check
SYMBOLS = [
    "clat_last_error", "clat_version", "clat_round_up", "clat_engine_length", "clat_lattice_grid",
    "clat_lattice_cells",
    "clat_pattern_enumerate", "clat_vertex_cover", "clat_plan", "clat_create", "clat_destroy",
    "clat_set_interaction_simple", "clat_set_interaction_thomas_fermi", "clat_set_interaction_table",
    "clat_set_input", "clat_prepare", "clat_calculate", "clat_get_potential", "clat_set_atoms",
    "clat_add_state", "clat_clear_states", "clat_compute_integrals", "clat_compute_integrals_group",
    "clat_compute_potential_at_atoms",
    "clat_compute_potential_at_atoms_noalloc",
    "clat_last_stage_ms", "clat_set_kernel_timing", "clat_set_kernel_timing_kinds", "clat_kernel_kind_count",
    "clat_kernel_kind_name",
    "clat_kernel_stats", "clat_reset_kernel_stats", "clat_engine_info", "clat_engine_kernels", "clat_set_tuning",
    "clat_tuning_info",
]


class ClatError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_LIB = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ClatError(-1, "libclat.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.clat_last_error.restype = ctypes.c_char_p
    L.clat_version.restype = ctypes.c_char_p
    L.clat_round_up.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.clat_engine_length.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.clat_lattice_grid.argtypes = [_dp, _i64, _ip, _ip, _dp, _dp]
    L.clat_lattice_cells.argtypes = [_dp, _i64, _ip, _dp, _dp, _ip]
    L.clat_pattern_enumerate.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, vp, _i64,
                                         ctypes.POINTER(_i64)]
    L.clat_vertex_cover.argtypes = [ctypes.c_int, vp, _i64, _ip, ctypes.POINTER(ctypes.c_int)]
    L.clat_plan.argtypes = [vp, _i64, ctypes.c_int, _ip]
    L.clat_create.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(vp)]
    L.clat_destroy.argtypes = [vp]
    L.clat_set_interaction_simple.argtypes = [vp, _dp, ctypes.c_double, ctypes.c_double]
    L.clat_set_interaction_thomas_fermi.argtypes = [vp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.clat_set_interaction_table.argtypes = [vp, _dp]
    L.clat_set_input.argtypes = [vp, _dp]
    L.clat_prepare.argtypes = [vp]
    L.clat_calculate.argtypes = [vp, _dp]
    L.clat_get_potential.argtypes = [vp, _dp]
    L.clat_set_atoms.argtypes = [vp, _ip, _i64, ctypes.c_int]
    L.clat_add_state.argtypes = [vp, _dp]
    L.clat_clear_states.argtypes = [vp]
    L.clat_compute_integrals.argtypes = [vp, vp, _i64, ctypes.c_int, vp]
    L.clat_compute_integrals_group.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]
    L.clat_compute_potential_at_atoms.argtypes = [vp, ctypes.c_int, _dp]
    L.clat_compute_potential_at_atoms_noalloc.argtypes = [vp, ctypes.c_int, _dp]
    L.clat_last_stage_ms.argtypes = [vp, _dp]
    L.clat_set_kernel_timing.argtypes = [vp, ctypes.c_int]
    L.clat_set_kernel_timing_kinds.argtypes = [vp, ctypes.c_uint32]
    L.clat_kernel_kind_count.restype = ctypes.c_int
    L.clat_kernel_kind_name.argtypes = [ctypes.c_int]
    L.clat_kernel_kind_name.restype = ctypes.c_char_p
    L.clat_kernel_stats.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]
    L.clat_reset_kernel_stats.argtypes = [vp]
    L.clat_engine_info.argtypes = [vp, _ip, ctypes.POINTER(_i64)]
    L.clat_engine_kernels.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
    L.clat_set_tuning.argtypes = [vp, ctypes.c_int]
    L.clat_tuning_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int), _dp, ctypes.POINTER(_i64),
                                   ctypes.POINTER(_i64)]
    _LIB = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise ClatError(rc, lib().clat_last_error().decode())


def _c(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64).reshape(-1)


# ------------------------------------------------------------ host helpers
def round_up(n: int, c: int = 1, primes=(2, 3, 5)) -> int:
    out = ctypes.c_int()
    _check(lib().clat_round_up(n, c, np.asarray(primes, np.int32), len(primes), ctypes.byref(out)))
    return out.value


def engine_length(n: int) -> int:
    """The axis length an engine computes on for a user length n."""
    out = ctypes.c_int()
    _check(lib().clat_engine_length(n, ctypes.byref(out)))
    return out.value


def lattice_grid(xyz):
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    raw = np.zeros(3, np.int32); pad = np.zeros(3, np.int32)
    origin = np.zeros(3); step = np.zeros(3)
    _check(lib().clat_lattice_grid(xyz, len(xyz) // 3, raw, pad, origin, step))
    return raw, pad, origin, step


def lattice_cells(xyz, dim, origin, step) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    out = np.zeros(len(xyz) // 3, np.int32)
    _check(lib().clat_lattice_cells(xyz, len(out), np.asarray(dim, np.int32), np.asarray(origin, np.float64),
                                    np.asarray(step, np.float64), out))
    return out


def pattern_enumerate(spec: str, n_states: int, n_holes: int) -> np.ndarray:
    cnt = _i64()
    _check(lib().clat_pattern_enumerate(spec.encode(), n_states, n_holes, None, 0, ctypes.byref(cnt)))
    out = np.zeros((max(cnt.value, 1), 4), np.int32)
    _check(lib().clat_pattern_enumerate(spec.encode(), n_states, n_holes, out.ctypes.data, cnt.value,
                                        ctypes.byref(cnt)))
    return out[:cnt.value]


def vertex_cover(nv: int, edges) -> list:
    e = np.ascontiguousarray(np.asarray(edges, np.int32).reshape(-1, 2))
    cover = np.zeros(max(nv, 1), np.int32)
    n = ctypes.c_int()
    _check(lib().clat_vertex_cover(nv, e.ctypes.data if len(e) else None, len(e), cover, ctypes.byref(n)))
    return [int(v) for v in cover[:n.value]]


def plan(quads, n_states: int) -> np.ndarray:
    q = np.ascontiguousarray(np.asarray(quads, np.int32).reshape(-1, 4))
    out = np.zeros((max(len(q), 1), 5), np.int32)
    _check(lib().clat_plan(q.ctypes.data if len(q) else None, len(q), n_states, out.reshape(-1)))
    return out[:len(q)]


def kernel_kinds() -> list:
    L = lib()
    return [L.clat_kernel_kind_name(k).decode() for k in range(L.clat_kernel_kind_count())]


# ------------------------------------------------------------ interactions
class InteractionSimple:
    """src/Interaction.cpp:8-14 -- epsilon constant."""

    def __init__(self, step, onsite: float, dielectric: float):
        self.step = np.asarray(step, np.float64)
        self.onsite = float(onsite)
        self.dielectric = float(dielectric)

    def _apply(self, h):
        _check(lib().clat_set_interaction_simple(h, self.step, self.onsite, self.dielectric))


class InteractionThomasFermi:
    """src/Interaction.cpp:18-39 -- Thomas-Fermi-Resta screening."""

    def __init__(self, step, onsite: float, dielectric: float, lattice_constant: float):
        self.step = np.asarray(step, np.float64)
        self.onsite = float(onsite)
        self.dielectric = float(dielectric)
        self.lattice_constant = float(lattice_constant)

    def _apply(self, h):
        _check(lib().clat_set_interaction_thomas_fermi(h, self.step, self.onsite, self.dielectric,
                                                       self.lattice_constant))


class InteractionTable:
    """An arbitrary Interaction::map (Interaction.hpp:14): fill(G) receives the
    (Z+1, Y+1, X+1) first-octant table to write, as the reference's
    CoulombCalculatorTest_Interaction does (run-tests.cpp:13-21)."""

    def __init__(self, fill):
        self.fill = fill

    def _apply(self, h, dims):
        X, Y, Z = dims
        G = np.zeros((Z + 1, Y + 1, X + 1))
        self.fill(G)
        _check(lib().clat_set_interaction_table(h, np.ascontiguousarray(G).reshape(-1)))


# ------------------------------------------------------------ calculator
class _Handle:
    def __init__(self, X, Y, Z, precision, schedule):
        h = ctypes.c_void_p()
        _check(lib().clat_create(X, Y, Z, precision, schedule, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().clat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CoulombCalculator:
    """CoulombCalculator (src/CoulombCalculator.hpp:28-94) on the GPU.

    Protocol: ctor(dims) -> initialize(interaction) -> { write input ->
    prepare() -> { write input -> calculate() } }.  `input` is a (Z, Y, X)
    complex128 numpy array (x fastest), zeroed at construction.
    """

    def __init__(self, dimension, precision=PREC_FP64, schedule=SCHED_SHARED):
        X, Y, Z = (int(v) for v in dimension)
        self.dims = (X, Y, Z)
        self._h = _Handle(X, Y, Z, precision, schedule)
        self.input = np.zeros((Z, Y, X), np.complex128)

    @property
    def handle(self):
        return self._h.h

    def kernels(self) -> set:
        """Names of the kernel paths the passes run on (KERNEL_FLAGS)."""
        return _kernels(self.handle)

    def initialize(self, interaction) -> None:
        if isinstance(interaction, InteractionTable):
            interaction._apply(self.handle, self.dims)
        else:
            interaction._apply(self.handle)

    def _upload(self):
        _check(lib().clat_set_input(self.handle, _c(self.input)))

    def prepare(self) -> None:
        self._upload()
        _check(lib().clat_prepare(self.handle))

    def calculate(self) -> complex:
        self._upload()
        out = np.zeros(2)
        _check(lib().clat_calculate(self.handle, out))
        return complex(out[0], out[1])

    def potential(self) -> np.ndarray:
        X, Y, Z = self.dims
        V = np.zeros((Z, Y, X), np.complex128)
        _check(lib().clat_get_potential(self.handle, V.view(np.float64).reshape(-1)))
        return V

    def close(self):
        self._h.close()


class Engine:
    """The atom-site integral path: atoms -> states -> integrals on one GPU."""

    def __init__(self, dimension, precision=PREC_FP64, schedule=SCHED_SHARED):
        X, Y, Z = (int(v) for v in dimension)
        self.dims = (X, Y, Z)
        self._h = _Handle(X, Y, Z, precision, schedule)
        self.n_atoms = 0
        self.n_orb = 0

    @property
    def handle(self):
        return self._h.h

    def initialize(self, interaction) -> None:
        if isinstance(interaction, InteractionTable):
            interaction._apply(self.handle, self.dims)
        else:
            interaction._apply(self.handle)

    def set_atoms(self, cells, n_orbitals: int) -> None:
        cells = np.ascontiguousarray(cells, np.int32)
        _check(lib().clat_set_atoms(self.handle, cells, len(cells), n_orbitals))
        self.n_atoms = len(cells)
        self.n_orb = n_orbitals

    def add_state(self, coeffs) -> None:
        c = np.ascontiguousarray(coeffs, np.complex128)
        if c.size != self.n_atoms * self.n_orb:
            raise ValueError("state has %d coefficients, expected %d" % (c.size, self.n_atoms * self.n_orb))
        _check(lib().clat_add_state(self.handle, _c(c)))

    def clear_states(self) -> None:
        _check(lib().clat_clear_states(self.handle))

    def compute_integrals(self, quads, stale_values: bool = False) -> np.ndarray:
        q = np.ascontiguousarray(np.asarray(quads, np.int32).reshape(-1, 4))
        out = np.zeros(len(q), np.complex128)
        _check(lib().clat_compute_integrals(self.handle, q.ctypes.data if len(q) else None, len(q),
                                            FLAG_STALE_VALUES if stale_values else 0,
                                            out.ctypes.data if len(q) else None))
        return out

    def potential_at_atoms(self, state: int) -> np.ndarray:
        out = np.zeros(self.n_atoms, np.complex128)
        _check(lib().clat_compute_potential_at_atoms(self.handle, state, out.view(np.float64).reshape(-1)))
        return out

    def potential_at_atoms_noalloc(self, state: int, out: np.ndarray | None = None) -> np.ndarray:
        """The same values through the allocation-free entry point (scratch
        in the library's static device array); out, if given, is reused."""
        if out is None:
            out = np.zeros(self.n_atoms, np.complex128)
        if out.dtype != np.complex128 or out.shape != (self.n_atoms,) or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous complex128 array of n_atoms values")
        _check(lib().clat_compute_potential_at_atoms_noalloc(self.handle, state, out.view(np.float64).reshape(-1)))
        return out

    def stage_ms(self) -> np.ndarray:
        out = np.zeros(5)
        _check(lib().clat_last_stage_ms(self.handle, out))
        return out

    def set_kernel_timing(self, on: bool) -> None:
        _check(lib().clat_set_kernel_timing(self.handle, 1 if on else 0))

    def set_kernel_timing_kinds(self, names) -> None:
        """Time only the kernel kinds named (the names of kernel_kinds())."""
        known = kernel_kinds()
        mask = 0
        for n in names:
            if n not in known:
                raise ValueError("unknown kernel kind %r" % n)
            mask |= 1 << known.index(n)
        _check(lib().clat_set_kernel_timing_kinds(self.handle, mask))

    def reset_kernel_stats(self) -> None:
        _check(lib().clat_reset_kernel_stats(self.handle))

    def kernel_stats(self) -> dict:
        out = {}
        for k, name in enumerate(kernel_kinds()):
            n = _i64(); ms = ctypes.c_double(); b = ctypes.c_double()
            _check(lib().clat_kernel_stats(self.handle, k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)))
            out[name] = (n.value, ms.value, b.value)
        return out

    def set_tuning(self, measure: bool) -> None:
        _check(lib().clat_set_tuning(self.handle, 1 if measure else 0))

    def tuning(self) -> dict:
        """Prepares per launch in use, measured ms per tried count, graph counts."""
        b = ctypes.c_int()
        ms = np.zeros(8)
        rep = _i64()
        cap = _i64()
        _check(lib().clat_tuning_info(self.handle, ctypes.byref(b), ms, ctypes.byref(rep), ctypes.byref(cap)))
        return {"batch": b.value, "trial_ms": {i + 1: float(ms[i]) for i in range(8) if ms[i] > 0},
                "graph_replays": rep.value, "graph_captures": cap.value}

    def info(self):
        dim = np.zeros(3, np.int32)
        nb = _i64()
        _check(lib().clat_engine_info(self.handle, dim, ctypes.byref(nb)))
        return tuple(int(v) for v in dim), nb.value

    def kernels(self) -> set:
        """Names of the kernel paths the engine's passes run on (KERNEL_FLAGS)."""
        return _kernels(self.handle)

    def close(self):
        self._h.close()


def _kernels(handle) -> set:
    f = ctypes.c_uint32()
    _check(lib().clat_engine_kernels(handle, ctypes.byref(f)))
    return {k for k, bit in KERNEL_FLAGS.items() if f.value & bit}


def compute_integrals_group(engines, quad_lists, stale_values: bool = False) -> list:
    """Integrals of several systems (one engine each, any grid sizes) with
    their device stages launched as one HIP graph; returns one complex array
    per engine, bit-identical to each engine's compute_integrals."""
    engines = list(engines)
    qs = [np.ascontiguousarray(np.asarray(q, np.int32).reshape(-1, 4)) for q in quad_lists]
    if len(qs) != len(engines) or not engines:
        raise ValueError("one integral list per engine, at least one engine")
    outs = [np.zeros(len(q), np.complex128) for q in qs]
    k = len(engines)
    hs = (ctypes.c_void_p * k)(*[e.handle for e in engines])
    qp = (ctypes.c_void_p * k)(*[q.ctypes.data if len(q) else None for q in qs])
    ns = (_i64 * k)(*[len(q) for q in qs])
    op = (ctypes.c_void_p * k)(*[o.ctypes.data if len(o) else None for o in outs])
    _check(lib().clat_compute_integrals_group(ctypes.cast(hs, ctypes.c_void_p), k, ctypes.cast(qp, ctypes.c_void_p),
                                              ctypes.cast(ns, ctypes.c_void_p),
                                              FLAG_STALE_VALUES if stale_values else 0,
                                              ctypes.cast(op, ctypes.c_void_p)))
    return outs


def run_concurrent(jobs) -> list:
    """Run several jobs at once, one host thread each, and return their results
    in order.  Each job is a callable without arguments that uses engines of
    its own (an engine handle is single-threaded).  Every engine launches on
    its own stream and shares only the tuner's choices (under a lock) and the
    allocation-free entry point's array with other engines, so the GPU overlaps
    the engines' kernels: the passes of small grids fill the device while a
    large grid streams.  ctypes releases the interpreter lock during each
    library call."""
    from concurrent.futures import ThreadPoolExecutor
    lib()  # load the library before the threads start
    jobs = list(jobs)
    with ThreadPoolExecutor(max_workers=max(1, len(jobs))) as ex:
        futures = [ex.submit(j) for j in jobs]
        return [f.result() for f in futures]
