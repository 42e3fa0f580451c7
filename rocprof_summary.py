"""Condense rocprofv3 outputs of the integral workload into short text summaries.

Reads the rocpd databases (``<dir>/**/*_results.db``) that

    rocprofv3 --kernel-trace --stats -d <dir> -o run -- python3 bench.py ...
    rocprofv3 --pmc FETCH_SIZE -d <dir> -o run -- python3 bench.py ...
    rocprofv3 --pmc WRITE_SIZE -d <dir> -o run -- python3 bench.py ...

write, plus the same two counter runs over ``bench.py --copy-probe MIB`` (a
plain device copy of known size), and writes

* a Markdown table of per-kernel call counts and average durations, and
* ``--traffic-json``: per pass kind, FETCH_SIZE and WRITE_SIZE in KiB per
  dispatch and the HBM bytes per launch after scaling each counter by
  (bytes the copy moved) / (what the counter read for it).

Pass kinds follow the kernel template arguments: mode 0/1/2 = forward / conv /
inverse, line = x pass, column = y or z pass.
"""
import argparse
import glob
import json
import os
import re
import sqlite3
from collections import defaultdict


def kind_of(name):
    m = re.search(r"k_pass2<(\w+), (\w+), (\w+), (\d), (true|false), (\d+), (\d+)>", name)
    if m:
        mode, line = int(m.group(4)), m.group(5) == "true"
        return {(0, True): "xfwd", (0, False): "yfwd", (1, False): "zconv",
                (2, False): "yinv", (2, True): "xinv"}[(mode, line)]
    m = re.search(r"k_colpass<(\w+), (\w+), (\w+), (\d),", name)
    if m:
        return {0: "yfwd", 1: "zconv", 2: "yinv"}[int(m.group(4))]
    if "k_colconv2<" in name:
        return "zconv"
    m = re.search(r"k_linepass<(\w+), (\w+), (\w+), (\d),", name)
    if m:
        return {0: "xfwd", 2: "xinv"}[int(m.group(4))]
    for key in ("k_products_direct", "k_products", "k_overlap_partial", "k_overlap_reduce", "k_dct1",
                "k_gconv", "k_gtable", "k_gpad", "k_d2f", "k_xfwd", "k_xinv", "k_col", "k_grid_dot",
                "k_atoms_from_sites", "k_coarse_g0", "k_coarse_gt", "k_coarse_q", "k_coset_gtable",
                "k_coset_samples", "k_coeffs_to_sites", "copyBuffer"):
        if key in name:
            return key[2:] if key.startswith("k_") else key
    return name.split("(")[0][:60]


def short(name):
    m = re.search(r"(k_(?:pass2|colpass|linepass|colconv2)<[^>]*>)", name)
    return m.group(1) if m else name.split("(")[0].replace("void ", "")[:70]


def db_of(d):
    hits = sorted(glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True))
    if not hits:
        raise SystemExit("no rocpd database under %s" % d)
    return hits[0]


def trace_stats(d):
    con = sqlite3.connect(db_of(d))
    rows = con.execute("select name, total_calls, total_duration, average from top_kernels").fetchall()
    return [(n, int(c), float(t), float(a)) for n, c, t, a in rows]


def counters(d, counter):
    con = sqlite3.connect(db_of(d))
    rows = con.execute("select kernel_name, counter_name, value from counters_collection").fetchall()
    acc = defaultdict(list)
    for k, c, v in rows:
        if c == counter:
            acc[k].append(float(v))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", help="output dir of the --kernel-trace --stats run")
    ap.add_argument("--sq", nargs="*", default=[],
                    help="output dirs of SQ counter runs (SQ_WAVES, SQ_WAVE_CYCLES, SQ_WAIT_ANY, "
                         "SQ_INSTS_VALU, SQ_LDS_BANK_CONFLICT, SQ_LDS_IDX_ACTIVE, ...)")
    ap.add_argument("--fetch", help="output dir of the FETCH_SIZE run")
    ap.add_argument("--write", help="output dir of the WRITE_SIZE run")
    ap.add_argument("--copy-fetch", help="FETCH_SIZE run of bench.py --copy-probe")
    ap.add_argument("--copy-write", help="WRITE_SIZE run of bench.py --copy-probe")
    ap.add_argument("--rdreq", help="run with TCC_EA0_RDREQ_{32B,64B,128B}_sum (read requests by size)")
    ap.add_argument("--copy-mib", type=float, default=1024.0)
    ap.add_argument("--title", default="kernel trace")
    ap.add_argument("--commands", default="", help="text block naming the commands that made the inputs")
    ap.add_argument("--md", required=True)
    ap.add_argument("--traffic-json")
    a = ap.parse_args()

    out = ["# %s" % a.title, "", "Commands:", "", "```", a.commands.strip(), "```", ""]
    if a.trace:
        stats = trace_stats(a.trace)
        total = sum(t for _, _, t, _ in stats)
        out += ["| kernel | kind | calls | avg (us) | total (ms) | share |", "|---|---|---:|---:|---:|---:|"]
        for n, c, t, avg in stats:
            out.append("| `%s` | %s | %d | %.1f | %.2f | %.1f%% |" % (short(n), kind_of(n), c, avg, t / 1e3,
                                                                       100.0 * t / total))
    if a.sq:
        # per-dispatch averages of the SQ counters, merged over the runs
        cs = {}
        for d in a.sq:
            con = sqlite3.connect(db_of(d))
            for c in {r[0] for r in con.execute("select distinct counter_name from counters_collection")}:
                for k, v in counters(d, c).items():
                    cs.setdefault(k, {})[c] = v
        out += ["SQ counters per dispatch (averages over the run's dispatches). Waiting share = SQ_WAIT_ANY / "
                "SQ_WAVE_CYCLES (wave-cycles waiting on anything); LDS bank-conflict share = "
                "SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE.", "",
                "| kernel | waves | VALU instr per wave | LDS instr per wave | waiting share | LDS bank-conflict share |",
                "|---|---:|---:|---:|---:|---:|"]
        for k in sorted(cs, key=lambda n: kind_of(n)):
            v = cs[k]
            if "SQ_WAVES" not in v or not v["SQ_WAVES"]:
                continue
            w = v["SQ_WAVES"]
            out.append("| `%s` | %.0f | %.0f | %s | %s | %s |" % (
                short(k), w, v.get("SQ_INSTS_VALU", 0) / w,
                "%.0f" % (v["SQ_INSTS_LDS"] / w) if "SQ_INSTS_LDS" in v else "—",
                "%.2f" % (v["SQ_WAIT_ANY"] / v["SQ_WAVE_CYCLES"]) if v.get("SQ_WAVE_CYCLES") else "—",
                "%.2f" % (v["SQ_LDS_BANK_CONFLICT"] / v["SQ_LDS_IDX_ACTIVE"]) if v.get("SQ_LDS_IDX_ACTIVE") else "—"))
    traffic = None
    if a.fetch and a.write and a.copy_fetch and a.copy_write:
        copy_bytes = a.copy_mib * (1 << 20)
        cf = counters(a.copy_fetch, "FETCH_SIZE")
        cw = counters(a.copy_write, "WRITE_SIZE")
        ck = [k for k in cf if "copyBuffer" in k]
        if not ck:
            raise SystemExit("copy kernel not found in the calibration run")
        f_factor = copy_bytes / (cf[ck[0]] * 1024.0)
        w_factor = copy_bytes / (cw[ck[0]] * 1024.0)
        fetch, write = counters(a.fetch, "FETCH_SIZE"), counters(a.write, "WRITE_SIZE")
        per_kind = {}
        for k in fetch:
            kd = kind_of(k) if ("k_pass2<" in k or "k_colpass<" in k or "k_linepass<" in k or "k_colconv2<" in k) \
                else short(k).replace("clat::", "")
            fk, wk = fetch[k], write.get(k, 0.0)
            per_kind[kd] = {"kernel": short(k), "fetch_kib": fk, "write_kib": wk,
                            "hbm_bytes_per_launch": fk * 1024.0 * f_factor + wk * 1024.0 * w_factor}
        traffic = {"copy_probe_mib": a.copy_mib, "fetch_factor": f_factor, "write_factor": w_factor,
                   "kernels": per_kind}
        out += ["", "## HBM traffic per dispatch", "",
                "Copy calibration: %.0f MiB copy read as FETCH_SIZE %.0f KiB (factor %.3f), "
                "WRITE_SIZE %.0f KiB (factor %.3f)." % (a.copy_mib, cf[ck[0]], f_factor, cw[ck[0]], w_factor),
                "", "| kind | kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM MB (scaled) |",
                "|---|---|---:|---:|---:|"]
        for kd, v in sorted(per_kind.items()):
            out.append("| %s | `%s` | %.0f | %.0f | %.1f |" % (kd, v["kernel"], v["fetch_kib"], v["write_kib"],
                                                               v["hbm_bytes_per_launch"] / 1e6))
        if a.rdreq:
            # read bytes from the request-size counters: 32, 64 and 128-byte requests
            r32, r64, r128 = (counters(a.rdreq, "TCC_EA0_RDREQ_%s_sum" % b) for b in ("32B", "64B", "128B"))
            out += ["", "Cross-check: HBM read bytes from the request-size counters (32·RDREQ_32B + "
                    "64·RDREQ_64B + 128·RDREQ_128B) against the scaled FETCH_SIZE.", "",
                    "| kind | 128-B requests | 64-B | 32-B | read MB (requests) | read MB (FETCH_SIZE x factor) |",
                    "|---|---:|---:|---:|---:|---:|"]
            for k in r128:
                kd = kind_of(k) if "k_pass2<" in k or "k_colpass<" in k or "k_linepass<" in k \
                    else short(k).replace("clat::", "")
                rb = 32 * r32.get(k, 0) + 64 * r64.get(k, 0) + 128 * r128[k]
                fb = fetch.get(k, 0.0) * 1024.0 * f_factor
                out.append("| %s | %.0f | %.0f | %.0f | %.1f | %.1f |" % (kd, r128[k], r64.get(k, 0), r32.get(k, 0),
                                                                        rb / 1e6, fb / 1e6))
                if kd in per_kind:
                    per_kind[kd]["read_bytes_by_request_size"] = rb
    with open(a.md, "w") as f:
        f.write("\n".join(out) + "\n")
    if traffic is not None and a.traffic_json:
        with open(a.traffic_json, "w") as f:
            json.dump(traffic, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
