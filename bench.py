"""Headline benchmark: integrals/s on the 1M-atom zincblende grid.

Workload (synthetic, generated here with the zincblende module): 64x64x32
conventional cells = 1,048,576 atoms on a 256x256x128 lattice grid, 20
orbitals per atom, 8 states (4 p-like holes, 4 s-like electrons, sigma = 0.15
box), --integrals=eeee,hhhh,ehhe,eheh (1024 integrals), --dielectric=10.89.

One step = the compute stage of one coulombo run with the states already in
HBM: interaction table + DCT-I, pair densities, the 36 prepares, every
overlap, the results back on the host.  value = integrals / second of the
complex64-storage / FP64-arithmetic path (CLAT_PREC_FP32X).  The FP64 path,
the plain FP64 path (eight independent rounds) and pure FP32 run the same
inputs for the accuracy and speed comparison; every FP32X integral's error
against the FP64 path is summarised by magnitude (error_distribution) and,
with --error-report FILE, written one line per integral.

Beside the headline, on the same JSON line (omitted with --headline-only):
  batched_form  32^3 cells = 262,144 atoms on a 128^3 grid, 32 states (16
                holes, 16 electrons), the full ijkl set (1,048,576 integrals);
  multi_system  six boxes of 16^3, 24^3, 30^3, 40^3, 48^3 and 100^3 cells
                (grids 64^3 ... 400^3, up to 8M atoms), 8 states and 20
                orbitals each, eeee,hhhh,ehhe,eheh; one engine per box, all
                boxes' steps run concurrently (one host thread and stream each).
Each carries ms/step, integrals/s, its dominant kernel's roofline and the
FP32X-vs-FP64 error summary.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells NX NY NZ]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "coulombo-lattice_amd"))

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
FFT_KINDS = ("xfwd", "yfwd", "zconv", "yinv", "xinv")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def reference_deps():
    """Whether this host could build and run the reference coulombo
    (Makefile:1-3: FFTW3 with its MPI and OpenMP parts, armadillo, an MPI
    compiler and launcher), probed by library and program name only."""
    import ctypes.util
    import shutil
    libs = {name: ctypes.util.find_library(name) for name in
            ("fftw3", "fftw3_mpi", "fftw3_omp", "armadillo")}
    progs = {name: shutil.which(name) for name in ("mpirun", "mpic++")}
    hdr = [h for h in ("fftw3.h", "fftw3-mpi.h", "armadillo")
           if any(os.path.exists(os.path.join(d, h)) for d in ("/usr/include", "/usr/local/include"))]
    ok = all(libs.values()) and all(progs.values()) and len(hdr) == 3
    return {"buildable": ok, "libraries": libs, "programs": progs, "headers_found": hdr}


def host_cores():
    n_logical = os.cpu_count() or 1
    try:
        n_avail = len(os.sched_getaffinity(0))
    except AttributeError:
        n_avail = n_logical
    phys = set()
    try:
        with open("/proc/cpuinfo") as f:
            pid = core = None
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    core = line.split(":")[1].strip()
                    phys.add((pid, core))
    except OSError:
        pass
    return {"logical": n_logical, "available_to_process": n_avail, "physical": len(phys) or None}


def cpu_baseline(w, quads, budget_s=12.0):
    """Oracle (FP64 C restatement, OpenMP) on a bounded sample of the same run:
    the interaction table, all pair densities, then whole prepares of the left
    products in plan order with their overlaps until about budget_s seconds
    of prepares have run; the per-prepare and per-overlap means are scaled to
    the full plan."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    import clat

    X, Y, Z = w["dims"]
    n = len(w["states"])
    steps = clat.plan(quads, n)
    lefts = sorted(set(int(s[1]) for s in steps))
    t0 = time.time()
    Gh = O.ghat(O.gtable(X, Y, Z, w["step"], 0.0, 10.89))
    t_init = time.time() - t0
    t0 = time.time()
    rho = O.products(np.stack(w["states"]))
    t_rho = time.time() - t0
    t_prep, t_ovl, n_prep, n_ovl = 0.0, 0.0, 0, 0
    for L in lefts:
        t0 = time.time()
        F = np.zeros(X * Y * Z, np.complex128)
        F[w["cells"]] = rho[L]
        V = O.prepare(F.reshape(Z, Y, X), Gh).reshape(-1)[w["cells"]]
        t_prep += time.time() - t0
        n_prep += 1
        t0 = time.time()
        for s in steps:
            if int(s[1]) == L:
                _ = np.sum(V * rho[int(s[3])])
                n_ovl += 1
        t_ovl += time.time() - t0
        if t_prep >= budget_s:
            break
    total = t_init + t_rho + len(lefts) * t_prep / n_prep + len(steps) * t_ovl / max(1, n_ovl)
    sampled = t_init + t_rho + t_prep + t_ovl
    deps = reference_deps()
    return {"value": len(quads) / total, "unit": "integrals/s", "cores": O.num_threads(), "kind": "port",
            "sample": "FP64 CPU restatement of the reference path (oracle/oracle.c: the same eight-round "
                      "method with its own recursive mixed-radix FFT, OpenMP; not the reference binary, "
                      "whose FFTW3/armadillo/MPI are %s on this host, so it understates an FFTW run) on "
                      "the same workload: interaction table + DCT-I, all pair densities, %d of %d prepares "
                      "and their %d overlaps (%.1f s of CPU work), scaled to the %d-integral run"
                      % ("present" if deps["buildable"] else "absent", n_prep, len(lefts), n_ovl, sampled,
                         len(quads)),
            "host_cores": host_cores(), "reference_deps": deps}


def traffic_of(kind, default_workload):
    """HBM bytes per launch of a pass kind from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the plain-copy calibration) of
    the default workload (1M atoms, 8 states), or None for another workload or
    when no summary exists for the kind."""
    if not default_workload:
        return None
    path = os.path.join(ROOT, "profiles", "traffic_1M_atoms.json")
    try:
        with open(path) as f:
            t = json.load(f)["kernels"]
    except (OSError, KeyError, ValueError):
        return None
    if kind in t:
        return t[kind]["hbm_bytes_per_launch"]
    # kinds timed over several kernels or keyed by kernel name in the summary
    parts = {"products": ("k_products",), "overlap": ("k_overlap_partial", "k_overlap_reduce")}.get(kind)
    if not parts:
        return None
    tot = 0.0
    for p in parts:
        hits = [v["hbm_bytes_per_launch"] for k, v in t.items() if k.startswith(p)]
        if len(hits) != 1:
            return None
        tot += hits[0]
    return tot


def error_summary(np, got, ref, quads, small=1e-6):
    """Per-integral relative error of got against ref (no floor), by decade
    of |I| / max|I|, with the largest errors named."""
    e = np.abs(got - ref) / np.abs(ref)
    scale = float(np.max(np.abs(ref)))
    mag = np.abs(ref) / scale
    bins = []
    for lo in range(0, -14, -1):
        sel = (mag <= 10.0 ** lo) & (mag > 10.0 ** (lo - 1))
        if sel.any():
            bins.append({"mag_range": "(1e%d, 1e%d]" % (lo - 1, lo), "integrals": int(sel.sum()),
                         "max_rel_err": float(np.max(e[sel])), "median_rel_err": float(np.median(e[sel]))})
    big = mag >= small
    worst = []
    for i in np.argsort(-e)[:5]:
        worst.append({"quad": [int(v) for v in quads[i]], "mag": float(mag[i]), "rel_err": float(e[i]),
                      "abs_err_over_max": float(abs(got[i] - ref[i]) / scale)})
    return {"integrals": int(len(e)), "max_rel_err": float(np.max(e)), "median_rel_err": float(np.median(e)),
            "n_over_1e-5": int(np.sum(e > 1e-5)),
            "max_rel_err_at_least_1e-6_of_max": float(np.max(e[big])) if big.any() else 0.0,
            "max_abs_err_over_max_below_1e-6_of_max":
                float(np.max(np.abs(got - ref)[~big]) / scale) if (~big).any() else 0.0,
            "by_magnitude": bins, "worst": worst}


def make_engine(clat, w, prec, sched, n_orb):
    X, Y, Z = w["dims"]
    eng = clat.Engine((X, Y, Z), prec, sched)
    eng.set_atoms(w["cells"], n_orb)
    for s in w["states"]:
        eng.add_state(s)
    return eng


def timed(eng, step, steps, warmup, dominant=False, breakdown=False):
    """warmup untimed steps, then `steps` timed ones.  With dominant=True one
    untimed step with every kernel timed first names the dominant kernel and
    the timed steps carry HIP events around its launches only (events around
    every launch would cost ~0.25 ms per step); breakdown=True adds two more
    untimed steps with every kernel timed (per-kernel ms per step)."""
    for _ in range(warmup):
        step(eng)
    dom = None
    if dominant:
        eng.set_kernel_timing(True)
        eng.reset_kernel_stats()
        step(eng)
        pre = eng.kernel_stats()
        dom = max(pre, key=lambda k: pre[k][1])
        eng.set_kernel_timing_kinds([dom])
    eng.reset_kernel_stats()
    vals = None
    t_all = time.perf_counter()
    for _ in range(steps):
        vals = step(eng)
    t_all = time.perf_counter() - t_all
    stats = eng.kernel_stats()
    t_graph = None
    if dominant:
        # the same steps with kernel timing off: the device stage replays as
        # HIP graphs (runs with timing events launch every pass directly)
        eng.set_kernel_timing(False)
        step(eng)
        t_graph = time.perf_counter()
        for _ in range(steps):
            step(eng)
        t_graph = time.perf_counter() - t_graph
    bd = None
    if breakdown:
        eng.set_kernel_timing(True)
        eng.reset_kernel_stats()
        for _ in range(2):
            step(eng)
        bd = {k: (v[0] / 2, v[1] / 2, v[2] / 2) for k, v in eng.kernel_stats().items()}
    eng.set_kernel_timing(False)
    return dict(vals=vals, total=t_all, steps=steps, dom=dom, stats=stats, breakdown=bd,
                stage=eng.stage_ms().tolist(), total_graph=t_graph, tuning=eng.tuning())


def roofline_of(r, flops_overlap=None, traffic=None):
    """Roofline of the dominant kernel from the events around its launches:
    HBM bytes for the passes, FP64 MFMA flops when the overlaps dominate."""
    dom = r["dom"]
    n_l, t_ms, byts = r["stats"][dom]
    avg_ms = t_ms / max(1, n_l)
    if dom == "overlap" and flops_overlap:
        ach = flops_overlap / (avg_ms * 1e-3) / 1e12
        return {"bound": "mfma", "kernel": dom, "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_FP64_TFLOPS, "traffic": traffic, "avg_launch_ms": avg_ms,
                "algorithmic_flops_per_launch": flops_overlap}
    ach = (byts / max(1, n_l)) / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": ach / PEAK_HBM_GBS, "traffic": traffic, "avg_launch_ms": avg_ms,
            "algorithmic_bytes_per_launch": byts / max(1, n_l)}


def batched_form(args, clat, zb, np):
    """32^3 cells, 16 holes + 16 electrons, 20 orbitals, ijkl: the batched form."""
    t0 = time.time()
    w = zb.workload(32, 32, 32, 16, 16, config=3, sigma_frac=0.15)
    quads = clat.pattern_enumerate("ijkl", 32, 16)
    X, Y, Z = w["dims"]
    log("batched form: %d atoms, grid %dx%dx%d, 32 states, %d integrals (%.1f s)"
        % (len(w["cells"]), X, Y, Z, len(quads), time.time() - t0))
    inter = clat.InteractionSimple(w["step"], 0.0, 10.89)

    def step(eng):
        eng.initialize(inter)
        return eng.compute_integrals(quads)

    steps = max(1, args.steps // 2)
    eng = make_engine(clat, w, clat.PREC_FP32X, clat.SCHED_SHARED, 20)
    r = timed(eng, step, steps, max(1, args.warmup // 2), dominant=True)
    eng.close()
    eng = make_engine(clat, w, clat.PREC_FP64, clat.SCHED_SHARED, 20)
    r64 = timed(eng, step, 1, 1)
    eng.close()
    n_left = len(np.unique(clat.plan(quads, 32)[:, 1]))
    P = 32 * 33 // 2
    ms = 1e3 * r["total"] / steps
    log("batched form: FP32X %.2f ms per step, FP64 %.2f" % (ms, 1e3 * r64["total"]))
    out = {"workload": "zincblende 32x32x32 cells = %d atoms, grid %dx%dx%d, 32 states, ijkl"
                       % (len(w["cells"]), X, Y, Z),
           "integrals": int(len(quads)), "steps": steps, "ms_per_step": ms,
           "integrals_per_s": len(quads) / (r["total"] / steps),
           "graph_replay_ms_per_step": 1e3 * r["total_graph"] / steps,
           "fp64_ms_per_step": 1e3 * r64["total"],
           "stage_ms": dict(zip(("table", "pair_densities", "prepares", "overlaps", "call"), r["stage"])),
           # every product is a left slot here, so the overlap table is symmetric:
           # its algorithmic flops are those of the P (P + 1) / 2 distinct pairs
           "roofline": roofline_of(r, 8.0 * (P * (P + 1) // 2 if n_left == P else n_left * P) * len(w["cells"])),
           "fp32x_vs_fp64": error_summary(np, r["vals"], r64["vals"], quads)}
    return out


def multi_system_batch(args, clat, zb, np, sizes=(16, 24, 30, 40, 48, 100)):
    """Six boxes, 4 holes + 4 electrons, 20 orbitals, eeee,hhhh,ehhe,eheh; one
    engine per box, every box's step run concurrently (run_concurrent), as one
    group (interaction builds concurrently, then every box's device stage in
    one HIP graph launch), and one after another for comparison."""
    quads = clat.pattern_enumerate("eeee,hhhh,ehhe,eheh", 8, 4)
    systems = []
    for n in sizes:
        t0 = time.time()
        w = zb.workload(n, n, n, 4, 4, config=4, sigma_frac=0.15, n_orb=20)
        systems.append(w)
        log("multi-system box %d^3 cells: grid %dx%dx%d, %d atoms (%.1f s)"
            % (n, *w["dims"], len(w["cells"]), time.time() - t0))

    def job_of(eng, w):
        inter = clat.InteractionSimple(w["step"], 0.0, 10.89)

        def job():
            eng.initialize(inter)
            return eng.compute_integrals(quads)
        return job

    engs = [make_engine(clat, w, clat.PREC_FP32X, clat.SCHED_SHARED, 20) for w in systems]
    jobs = [job_of(e, w) for e, w in zip(engs, systems)]
    steps = max(1, args.steps // 2)
    for _ in range(max(1, args.warmup // 2)):
        clat.run_concurrent(jobs)
    # the largest box names the dominant kernel (it takes most of the step)
    big = engs[-1]
    big.set_kernel_timing(True)
    big.reset_kernel_stats()
    jobs[-1]()
    pre = big.kernel_stats()
    dom = max(pre, key=lambda k: pre[k][1])
    big.set_kernel_timing_kinds([dom])
    big.reset_kernel_stats()
    t = time.perf_counter()
    for _ in range(steps):
        con = clat.run_concurrent(jobs)
    t_con = (time.perf_counter() - t) / steps
    stats = big.kernel_stats()
    big.set_kernel_timing(False)
    inters = [clat.InteractionSimple(w["step"], 0.0, 10.89) for w in systems]
    inits = [(lambda e=e, it=it: e.initialize(it)) for e, it in zip(engs, inters)]

    def group_step():
        clat.run_concurrent(inits)
        return clat.compute_integrals_group(engs, [quads] * len(engs))
    group_step()  # captures the group graph
    t = time.perf_counter()
    for _ in range(steps):
        grp = group_step()
    t_grp = (time.perf_counter() - t) / steps
    t = time.perf_counter()
    seq = [j() for j in jobs]
    t_seq = time.perf_counter() - t
    per_box = []
    for e, w, j in zip(engs, systems, jobs):
        t = time.perf_counter()
        j()
        per_box.append(1e3 * (time.perf_counter() - t))
    for e in engs:
        e.close()
    same = all(np.array_equal(a, b) for a, b in zip(seq, con))
    same_grp = all(np.array_equal(a, b) for a, b in zip(grp, con))
    errs = []
    for w, v in zip(systems, con):
        e64 = make_engine(clat, w, clat.PREC_FP64, clat.SCHED_SHARED, 20)
        ref = job_of(e64, w)()
        e64.close()
        s = error_summary(np, v, ref, quads)
        errs.append({"grid": "%dx%dx%d" % tuple(w["dims"]), "atoms": int(len(w["cells"])),
                     "max_rel_err": s["max_rel_err"], "n_over_1e-5": s["n_over_1e-5"],
                     "max_rel_err_at_least_1e-6_of_max": s["max_rel_err_at_least_1e-6_of_max"],
                     "max_abs_err_over_max_below_1e-6_of_max": s["max_abs_err_over_max_below_1e-6_of_max"],
                     "worst": s["worst"][:2]})
    n_int = len(quads) * len(systems)
    log("multi-system: concurrent %.2f ms per step, one group launch %.2f, one after another %.2f"
        % (1e3 * t_con, 1e3 * t_grp, 1e3 * t_seq))
    r = {"dom": dom, "stats": stats}
    return {"workload": "zincblende boxes of %s cells, 8 states and 20 orbitals each, eeee,hhhh,ehhe,eheh"
                        % "/".join("%d^3" % n for n in sizes),
            "integrals": n_int, "steps": steps, "ms_per_step": 1e3 * t_con, "integrals_per_s": n_int / t_con,
            "group_ms_per_step": 1e3 * t_grp, "group_integrals_per_s": n_int / t_grp,
            "sequential_ms_per_step": 1e3 * t_seq, "box_ms": per_box,
            "concurrent_equals_sequential": bool(same), "group_equals_concurrent": bool(same_grp),
            "roofline_largest_box": roofline_of(r), "fp32x_vs_fp64": errs}


def multi_system(args, clat, zb, np):
    """Several zincblende systems of uneven size, one engine each (FP32X): the
    step is every system's interaction build and integrals, run one after
    another and then concurrently (one host thread and one stream per engine)."""
    sizes = [int(v) for v in args.systems.split(",")]
    nh, ne = args.states
    n_orb = 20
    systems = []
    for k, n in enumerate(sizes):
        t0 = time.time()
        w = zb.workload(n, n, n, nh, ne, config=4, sigma_frac=0.15, n_orb=n_orb)
        X, Y, Z = w["dims"]
        eng = make_engine(clat, w, clat.PREC_FP32X, clat.SCHED_SHARED, n_orb)
        quads = clat.pattern_enumerate(args.integrals, nh + ne, nh)
        inter = clat.InteractionSimple(w["step"], 0.0, 10.89)
        systems.append((eng, inter, quads, (X, Y, Z), len(w["cells"])))
        del w
        log("system %d: %d^3 cells, grid %dx%dx%d, %d atoms, %d integrals (%.1f s)"
            % (k, n, X, Y, Z, systems[-1][4], len(quads), time.time() - t0))

    def job(sys_):
        eng, inter, quads = sys_[0], sys_[1], sys_[2]
        eng.initialize(inter)
        return eng.compute_integrals(quads)

    for _ in range(max(1, args.warmup)):
        for sy in systems:
            job(sy)
    t = time.perf_counter()
    for _ in range(args.steps):
        seq = [job(sy) for sy in systems]
    t_seq = (time.perf_counter() - t) / args.steps
    t = time.perf_counter()
    for _ in range(args.steps):
        con = clat.run_concurrent([(lambda sy=sy: job(sy)) for sy in systems])
    t_con = (time.perf_counter() - t) / args.steps
    same = all(np.array_equal(a, b) for a, b in zip(seq, con))
    n_int = sum(len(sy[2]) for sy in systems)
    out = {
        "metric": "integrals/s, multi-system batch (one engine per system, concurrent)",
        "value": n_int / t_con, "unit": "integrals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_con, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (FP32 FFT passes; FP64 last pass, kernel table, pair densities and overlap sums)",
        "data": "synthetic (seeded zincblende, %d orbitals per atom)" % n_orb,
        "config": {"workload": "zincblende boxes of %s cells, %d states each, %s"
                               % ("/".join("%d^3" % n for n in sizes), nh + ne, args.integrals),
                   "grids": ["%dx%dx%d" % sy[3] for sy in systems], "atoms": [sy[4] for sy in systems]},
        "sequential_ms_per_step": 1e3 * t_seq, "concurrent_speedup": t_seq / t_con,
        "concurrent_equals_sequential": bool(same),
    }
    for sy in systems:
        sy[0].close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, nargs=3, default=[64, 64, 32])
    ap.add_argument("--states", type=int, nargs=2, default=[4, 4])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the batched-form and multi-system entries of the JSON line")
    ap.add_argument("--error-report", default="", metavar="FILE",
                    help="write every headline integral with its FP32X and FP64 values and relative error")
    ap.add_argument("--schedule", choices=["shared", "rounds"], default="shared")
    ap.add_argument("--integrals", default="eeee,hhhh,ehhe,eheh",
                    help="integral pattern list (ijkl with --cells 32 32 32 --states 16 16: the batched form)")
    ap.add_argument("--systems", default="",
                    help="comma-separated cubic box sizes in cells (e.g. 16,24,30,40,48,100): only the "
                         "multi-system workload, one engine per system, its own JSON line")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 3 copies of %d MiB" % args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported "
              "(the integral path needs no collectives; one engine serves one GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import clat
    import zincblende as zb

    if args.systems:
        multi_system(args, clat, zb, np)
        return

    nx, ny, nz = args.cells
    nh, ne = args.states
    t0 = time.time()
    w = zb.workload(nx, ny, nz, nh, ne, config=1, sigma_frac=0.15)
    X, Y, Z = w["dims"]
    quads = clat.pattern_enumerate(args.integrals, nh + ne, nh)
    log("workload: %d atoms, grid %dx%dx%d, %d states, %d integrals (%.1f s)"
        % (len(w["cells"]), X, Y, Z, nh + ne, len(quads), time.time() - t0))
    sched = clat.SCHED_SHARED if args.schedule == "shared" else clat.SCHED_ROUNDS
    inter = clat.InteractionSimple(w["step"], 0.0, 10.89)

    def step(eng):
        eng.initialize(inter)
        return eng.compute_integrals(quads)

    # headline: complex64 storage + FP64 arithmetic (CLAT_PREC_FP32X), shared rounds.
    # Beside it, on the same inputs: the FP64 path (accuracy reference), the
    # plain FP64 path (the reference's eight independent rounds) and pure FP32.
    runs = [("fp32x", clat.PREC_FP32X, sched), ("fp64", clat.PREC_FP64, sched),
            ("fp64_rounds", clat.PREC_FP64, clat.SCHED_ROUNDS), ("fp32", clat.PREC_FP32, sched)]
    if args.no_fp64:
        runs = runs[:1]
    res = {}
    for name, prec, sch in runs:
        eng = make_engine(clat, w, prec, sch, 20)
        head_run = name == "fp32x"
        steps = args.steps if name in ("fp32x", "fp64") else max(1, args.steps // 3)
        res[name] = timed(eng, step, steps, args.warmup, dominant=head_run, breakdown=head_run)
        log("%s: mean step %.2f ms  stages(ms) table %.2f rho %.2f prepares %.2f overlaps %.2f"
            % (name, 1e3 * res[name]["total"] / steps, *res[name]["stage"][:4]))
        eng.close()

    head = res["fp32x"]
    default_workload = (list(args.cells) == [64, 64, 32] and list(args.states) == [4, 4]
                        and args.integrals == "eeee,hhhh,ehhe,eheh" and args.schedule == "shared")
    ms = 1e3 * head["total"] / args.steps
    value = len(quads) / (head["total"] / args.steps)
    roof = roofline_of(head, traffic=traffic_of(head["dom"], default_workload))
    out = {
        "metric": "integrals/s, 1M-atom grid, FP32+FP64-acc vs FP64 at 1e-5 rel; FFT HBM GB/s",
        "value": value, "unit": "integrals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (FP32 FFT passes; FP64 last pass, kernel table, pair densities and overlap sums)",
        "data": "synthetic (seeded zincblende, see zincblende.py)",
        "config": {"workload": "zincblende %dx%dx%d cells = %d atoms, grid %dx%dx%d, %d states, %s"
                               % (nx, ny, nz, len(w["cells"]), X, Y, Z, nh + ne, args.integrals),
                   "integrals": len(quads), "schedule": args.schedule},
        "roofline": roof,
        "kernel_ms_per_step": {k: v[1] for k, v in head["breakdown"].items() if v[0]},
        "graph_replay_ms_per_step": 1e3 * head["total_graph"] / args.steps,
        "tuning": head["tuning"],
        "fft_pass_gbs": {k: (v[2] / max(1e-12, v[1] * 1e-3)) / 1e9 for k, v in head["breakdown"].items()
                         if v[0] and k in FFT_KINDS},
    }
    # FFT HBM GB/s: measured HBM bytes per launch of each pass (the committed
    # FETCH_SIZE + WRITE_SIZE summary of this workload) times launches per
    # step, over the five passes' event-timed duration per step
    fft_tr = {k: traffic_of(k, default_workload) for k in FFT_KINDS}
    if all(fft_tr[k] for k in FFT_KINDS) and all(head["breakdown"][k][0] for k in FFT_KINDS):
        b = head["breakdown"]
        hbm = sum(fft_tr[k] * b[k][0] for k in FFT_KINDS)
        t_s = sum(b[k][1] for k in FFT_KINDS) * 1e-3
        out["fft_hbm_gbs"] = {"value": hbm / t_s / 1e9, "peak": PEAK_HBM_GBS,
                              "frac": hbm / t_s / 1e9 / PEAK_HBM_GBS,
                              "bytes_per_step": hbm, "pass_ms_per_step": t_s * 1e3}
    if "fp64" in res:
        v64 = res["fp64"]["vals"]
        for name in ("fp32x", "fp32", "fp64_rounds"):
            if name in res:
                rel = np.abs(res[name]["vals"] - v64) / np.abs(v64)
                out["max_rel_err_%s_vs_fp64" % name] = float(np.max(rel))
                out["%s_ms_per_step" % name] = 1e3 * res[name]["total"] / res[name]["steps"]
        out["fp64_ms_per_step"] = 1e3 * res["fp64"]["total"] / res["fp64"]["steps"]
        out["fp64_integrals_per_s"] = len(quads) / (res["fp64"]["total"] / res["fp64"]["steps"])
        out["speedup_vs_fp64_shared"] = out["fp64_ms_per_step"] / ms
        if "fp64_rounds" in res:
            out["speedup_vs_plain_fp64_rounds"] = out["fp64_rounds_ms_per_step"] / ms
        out["error_distribution"] = error_summary(np, head["vals"], v64, quads)
        if args.error_report:
            with open(args.error_report, "w") as f:
                f.write("# i1 i2 i3 i4  fp32x_re fp32x_im  fp64_re fp64_im  rel_err  |I|/max\n")
                scale = np.max(np.abs(v64))
                for q, a, b in zip(quads, head["vals"], v64):
                    f.write("%2d %2d %2d %2d  %.17g %.17g  %.17g %.17g  %.3e  %.3e\n"
                            % (*q, a.real, a.imag, b.real, b.imag, abs(a - b) / abs(b), abs(b) / scale))
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(w, quads)
        except Exception as ex:  # reported, never fatal to the GPU number
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    del w
    if not args.headline_only:
        for key, fn in (("batched_form", batched_form), ("multi_system", multi_system_batch)):
            try:
                out[key] = fn(args, clat, zb, np)
            except Exception as ex:  # reported, never fatal to the headline
                out[key] = {"error": "%s: %s" % (type(ex).__name__, ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
