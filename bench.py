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
inputs for the accuracy and speed comparison.

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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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
    return {"value": len(quads) / total, "unit": "integrals/s", "cores": O.num_threads(), "kind": "port",
            "sample": "FP64 oracle on the same workload: interaction table + DCT-I, all pair densities, "
                      "%d of %d prepares and their %d overlaps (%.1f s of CPU work), scaled to the %d-integral "
                      "run" % (n_prep, len(lefts), n_ovl, sampled, len(quads))}


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
            t = json.load(f)
        return t["kernels"][kind]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def multi_system(args, clat, zb, np):
    """Several zincblende systems of uneven size, one engine each (FP32X): the
    step is every system's interaction build and integrals, run one after
    another and then concurrently (one host thread and one stream per engine).
    Four orbitals per atom keep the host-side generation of the 8M-atom
    coefficients short; only the pair-density kernel reads them."""
    sizes = [int(v) for v in args.systems.split(",")]
    nh, ne = args.states
    n_orb = 4
    systems = []
    for k, n in enumerate(sizes):
        t0 = time.time()
        w = zb.workload(n, n, n, nh, ne, config=4, sigma_frac=0.15, n_orb=n_orb)
        X, Y, Z = w["dims"]
        eng = clat.Engine((X, Y, Z), clat.PREC_FP32X, clat.SCHED_SHARED)
        eng.set_atoms(w["cells"], n_orb)
        for st in w["states"]:
            eng.add_state(st)
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
    log("one after another: %.2f ms per step" % (1e3 * t_seq))
    t = time.perf_counter()
    for _ in range(args.steps):
        con = clat.run_concurrent([(lambda sy=sy: job(sy)) for sy in systems])
    t_con = (time.perf_counter() - t) / args.steps
    log("concurrent: %.2f ms per step" % (1e3 * t_con))
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
    ap.add_argument("--schedule", choices=["shared", "rounds"], default="shared")
    ap.add_argument("--integrals", default="eeee,hhhh,ehhe,eheh",
                    help="integral pattern list (ijkl with --cells 32 32 32 --states 16 16: the batched form)")
    ap.add_argument("--systems", default="",
                    help="comma-separated cubic box sizes in cells (e.g. 16,24,30,40,48,100): the multi-system "
                         "workload, one engine per system, timed one after another and concurrently")
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

    def make(prec):
        eng = clat.Engine((X, Y, Z), prec, sched)
        eng.set_atoms(w["cells"], 20)
        for s in w["states"]:
            eng.add_state(s)
        return eng

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
        eng = clat.Engine((X, Y, Z), prec, sch)
        eng.set_atoms(w["cells"], 20)
        for s in w["states"]:
            eng.add_state(s)
        for _ in range(args.warmup):
            step(eng)
        head_run = name == "fp32x"
        dom = None
        if head_run:
            # one untimed step with every kernel timed names the dominant
            # kernel; the timed steps then carry events around its launches
            # only (events around every launch cost ~0.25 ms per step)
            eng.set_kernel_timing(True)
            eng.reset_kernel_stats()
            step(eng)
            pre = eng.kernel_stats()
            dom = max(pre, key=lambda k: pre[k][1])
            eng.set_kernel_timing_kinds([dom])
        eng.reset_kernel_stats()
        steps = args.steps if name in ("fp32x", "fp64") else max(1, args.steps // 3)
        times = []
        vals = None
        t_all = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            vals = step(eng)
            times.append(time.perf_counter() - t)
        t_all = time.perf_counter() - t_all
        stats = eng.kernel_stats()
        breakdown = None
        if head_run:
            # per-kernel breakdown from two more (untimed) steps
            eng.set_kernel_timing(True)
            eng.reset_kernel_stats()
            for _ in range(2):
                step(eng)
            breakdown = {k: (v[0] / 2, v[1] / 2, v[2] / 2) for k, v in eng.kernel_stats().items()}
            eng.set_kernel_timing(False)
        res[name] = dict(times=times, vals=vals, stats=stats, dom=dom, breakdown=breakdown,
                         stage=eng.stage_ms().tolist(), total=t_all, steps=steps)
        log("%s: mean step %.2f ms  stages(ms) table %.2f rho %.2f prepares %.2f overlaps %.2f"
            % (name, 1e3 * t_all / steps, *res[name]["stage"][:4]))
        eng.close()

    head = res["fp32x"]
    default_workload = (list(args.cells) == [64, 64, 32] and list(args.states) == [4, 4]
                        and args.integrals == "eeee,hhhh,ehhe,eheh" and args.schedule == "shared")
    ms = 1e3 * head["total"] / args.steps
    value = len(quads) / (head["total"] / args.steps)
    # roofline of the dominant kernel, from HIP events around each of its
    # launches in the timed steps
    stats = head["stats"]
    dom = head["dom"]
    n_l, t_ms, byts = stats[dom]
    avg_ms = t_ms / max(1, n_l)
    achieved = (byts / max(1, n_l)) / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "integrals/s, 1M-atom grid, FP32+FP64-acc vs FP64 at 1e-5 rel; FFT HBM GB/s",
        "value": value, "unit": "integrals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (FP32 FFT passes; FP64 last pass, kernel table, pair densities and overlap sums)",
        "data": "synthetic (seeded zincblende, see zincblende.py)",
        "config": {"workload": "zincblende %dx%dx%d cells = %d atoms, grid %dx%dx%d, %d states, %s"
                               % (nx, ny, nz, len(w["cells"]), X, Y, Z, nh + ne, args.integrals),
                   "integrals": len(quads), "schedule": args.schedule},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_of(dom, default_workload),
                     "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": byts / max(1, n_l)},
        "kernel_ms_per_step": {k: v[1] for k, v in head["breakdown"].items() if v[0]},
        "fft_pass_gbs": {k: (v[2] / max(1e-12, v[1] * 1e-3)) / 1e9 for k, v in head["breakdown"].items()
                         if v[0] and k in ("xfwd", "yfwd", "zconv", "yinv", "xinv")},
    }
    # FFT HBM GB/s: measured HBM bytes per launch of each pass (the committed
    # FETCH_SIZE + WRITE_SIZE summary of this workload) times launches per
    # step, over the five passes' event-timed duration per step
    fft_kinds = ("xfwd", "yfwd", "zconv", "yinv", "xinv")
    fft_tr = {k: traffic_of(k, default_workload) for k in fft_kinds}
    if all(fft_tr[k] for k in fft_kinds) and all(head["breakdown"][k][0] for k in fft_kinds):
        b = head["breakdown"]
        hbm = sum(fft_tr[k] * b[k][0] for k in fft_kinds)
        t_s = sum(b[k][1] for k in fft_kinds) * 1e-3
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
        out["speedup_vs_fp64_shared"] = out["fp64_ms_per_step"] / ms
        if "fp64_rounds" in res:
            out["speedup_vs_plain_fp64_rounds"] = out["fp64_rounds_ms_per_step"] / ms
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(w, quads)
        except Exception as ex:  # reported, never fatal to the GPU number
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
