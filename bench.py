"""Benchmark of the matcher on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config s22|s18|small|batched|diff]
                    [--no-cpu-baseline] [--no-pruned] [--no-profile]

A step is one full match of the workload: prep (cdist and edge-cost statistics, src/main.cu:107-148) plus pn
iterations of the loop (src/main.cu:160-245) plus the output write, with the graphs already resident in HBM.
Throughput is the reference's unit of work, one (data edge, pattern edge) pair per iteration:
    value = de * pe * pn / t_loop   (edge-pair updates per second; t_loop = device time of the iterations and the
                                     output write, the reference's timed region without its prep).
The default workload is R-MAT scale 22 (64-dim features, planted 16-vertex / 24-edge pattern).  --config batched
scores 256 uneven patterns against R-MAT 20 (batched against sequential runs); --config diff times the adjoint
(forward+backward against forward, 64 patterns on R-MAT 18).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# Arithmetic of the path: values stored in FP32 (costs, maxima, VR rows) and per-term exponentials in FP32; every
# difference, maximum and log-sum-exp merge in FP64 registers.
DTYPE = "f32 storage, f64 accumulation"


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(dn, de, pn, pe):
    """Per iteration.  iteration: the fixed figure 4*pe*de + 8*de + (40*pe + 12*pn)*dn (FP32 values, int32 ids);
    edge_walk: the edge stream (costs + ids) plus one read of VR and one write of FMax per vertex."""
    return {
        "iteration": 4 * pe * de + 8 * de + (40 * pe + 12 * pn) * dn,
        "edge_walk": 4 * pe * de + 8 * de + 8 * pe * dn,
        "vertex_a": (24 * pe + 4 * pn) * dn,
        "vertex_c": (8 * pe + 4 * pn) * dn,
    }


def cpu_model():
    """The host CPU's model name (lscpu "Model name", else /proc/cpuinfo), or None."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return None


def _cpu_leg(scale, threads, precision):
    from application_classification_amd import workloads as W
    from oracle import oracle as O

    w = W.rmat(scale, 64, 16, 24, seed=0)
    d, p = w.data, w.pattern
    D = O.Graph64.of(d.vfeat, d.efeat, d.src, d.dst)
    P = O.Graph64.of(p.vfeat, p.efeat, p.src, p.dst)
    prep_s, loop_s, _ = O.timed_run(D, P, threads=threads, precision=precision)
    de, pe, pn = d.num_edges, p.num_edges, p.num_nodes
    return {"value": de * pe * pn / loop_s, "threads": threads, "precision": precision,
            "sample": f"R-MAT scale {scale} ({d.num_nodes} vertices, {de} edges), 64-dim features, planted 16/24 "
                      f"pattern, {pn} iterations, {'FP64' if precision == 'f64' else 'FP32'}; loop {loop_s:.2f} s, prep {prep_s:.2f} s"}


def cpu_baseline(threads, legs=True):
    """The C++ restatement of the reference (oracle/: the same stages and materialised pattern x data buffers as
    src/main.cu, OpenMP loops in place of Thrust), timed on this host over the reference's loop region, on bounded
    samples of the benchmark workload's shape (64-dim features, planted 16/24 pattern).  The headline leg is FP64
    (the reference's precision) on every host thread at R-MAT scale 18; the other legs are FP32 on every thread and
    FP64 on one thread (at scale 16, so that it stays within seconds)."""
    main = _cpu_leg(18, threads, "f64")
    out = {"value": main["value"], "unit": "edge-pair updates/s", "cores": threads, "kind": "port",
           "sample": main["sample"], "model": cpu_model()}
    if legs:
        out["legs"] = [_cpu_leg(18, threads, "f32"), _cpu_leg(16, 1, "f64")]
    return out


STAGE_OF_KERNEL = {"k_walk_max<true": "edge_walk", "k_vertex<0>": "vertex_a", "k_vertex<1>": "vertex_c",
                   "k_vertex<2>": "output", "k_walk_max_fixup": "walk_fixup", "k_reduce_q": "reduce_a",
                   "k_reduce_lse": "reduce_c"}


def _short(name):
    name = name.replace("void ", "").replace("ac::", "")
    return name.split("(")[0]


def summarize_profile(src, workload, out_dir):
    """Writes profiles/<workload>_kernels.md and profiles/<workload>_traffic.json from a directory holding
    rocprofv3 outputs: trace/ (--kernel-trace --stats), fetch/ and write/ (--pmc FETCH_SIZE / WRITE_SIZE of the
    bench), cfetch/ and cwrite/ (the same counters for --copy-calibration, a plain copy of 3 x 1 GiB)."""
    import collections
    import csv
    import glob

    def rows(sub):
        f = glob.glob(os.path.join(src, sub, "*_kernel_stats.csv")) or glob.glob(os.path.join(src, sub,
                                                                                                "*_counter_collection.csv"))
        return list(csv.DictReader(open(f[0]))) if f else []

    def per_kernel(sub, counter):
        agg = collections.defaultdict(list)
        for r in rows(sub):
            if r.get("Counter_Name") == counter:
                agg[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]) * 1024.0)  # KiB -> bytes
        return {k: sum(v) / len(v) for k, v in agg.items()}

    copy_bytes = 3 * (1 << 30) / 3  # per copy launch
    cf = per_kernel("cfetch", "FETCH_SIZE")
    cw = per_kernel("cwrite", "WRITE_SIZE")
    fk = next((k for k in cf if "copyBuffer" in k), None)
    wk = next((k for k in cw if "copyBuffer" in k), None)
    f_scale = copy_bytes / cf[fk] if fk else None
    w_scale = copy_bytes / cw[wk] if wk else None
    fetch = per_kernel("fetch", "FETCH_SIZE")
    write = per_kernel("write", "WRITE_SIZE")
    stats = rows("trace")
    lines = [f"# Kernel profile: {workload}", "",
             "Commands (run from the repository root on one MI355X):", "",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <dir>/trace -o s22 -- "
             "python -u bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-pruned --no-extras",
             "    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir>/fetch -o s22 -- "
             "python -u bench.py --steps 1 --warmup 0 --no-cpu-baseline --no-profile --no-pruned --no-extras",
             "    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir>/write -o s22 -- "
             "python -u bench.py --steps 1 --warmup 0 --no-cpu-baseline --no-profile --no-pruned --no-extras",
             "    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir>/cfetch -o copy -- python -u bench.py --copy-calibration",
             "    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir>/cwrite -o copy -- python -u bench.py --copy-calibration",
             f"    python bench.py --summarize-profile <dir>", "",
             f"Counter calibration on a plain 1 GiB device copy: FETCH_SIZE x {f_scale:.4f}, WRITE_SIZE x {w_scale:.4f} "
             "(bytes per counted byte)." if f_scale and w_scale else "Counter calibration: missing.", "",
             "| kernel | calls | avg us | total ms | % | HBM read GB/launch | HBM write GB/launch |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    traffic = {"workload": workload, "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"])):
        k = _short(r["Name"])
        fr = fetch.get(k)
        wr = write.get(k)
        fb = fr * f_scale if fr is not None and f_scale else None
        wb = wr * w_scale if wr is not None and w_scale else None
        lines.append(f"| {k} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | {float(r['TotalDurationNs']) / 1e6:.3f} "
                     f"| {float(r['Percentage']):.1f} | {fb / 1e9 if fb else float('nan'):.3f} | "
                     f"{wb / 1e9 if wb else float('nan'):.3f} |")
        traffic["kernels"][k] = {"avg_us": float(r["AverageNs"]) / 1e3, "calls": int(r["Calls"]),
                                 "read_bytes": fb, "write_bytes": wb}
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, f"{workload}_kernels.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(out_dir, f"{workload}_traffic.json"), "w") as f:
        json.dump(traffic, f, indent=1)
    print("\n".join(lines))


def summarize_counters(src, workload, out_dir):
    """Writes profiles/<workload>_counters.md from rocprofv3 counter passes under src: req/ and creq/ (L2 read requests
    to memory by size, of the bench and of --copy-calibration), hit/ (L2 hits and misses), sqa/ and sqb/ (wave
    counters).  Figures are averages per dispatch."""
    import collections
    import csv
    import glob

    def load(sub):
        agg = collections.defaultdict(lambda: collections.defaultdict(list))
        for f in glob.glob(os.path.join(src, sub, "**", "*_counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                agg[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in agg.items()}

    req, creq, hit, sqa, sqb = (load(d) for d in ("req", "creq", "hit", "sqa", "sqb"))
    cmd = "python -u bench.py --steps 1 --warmup 0 --no-cpu-baseline --no-profile --no-pruned --no-extras"
    passes = [("req", "TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum TCC_EA0_RDREQ_128B_sum", cmd),
              ("creq", "TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum TCC_EA0_RDREQ_128B_sum",
               "python -u bench.py --copy-calibration"),
              ("hit", "TCC_HIT_sum TCC_MISS_sum", cmd),
              ("sqa", "SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_INSTS_VALU "
                      "SQ_INSTS_SALU SQ_ACTIVE_INST_VALU", cmd),
              ("sqb", "SQ_INSTS_VMEM SQ_INST_LEVEL_VMEM SQ_ACTIVE_INST_ANY SQ_INSTS_BRANCH", cmd)]
    lines = [f"# Counters: {workload}", "", "Commands (run from the repository root on one MI355X, one pass each):", ""]
    lines += [f"    rocprofv3 --pmc {c} --output-format csv -d <dir>/{d} -o c -- {x}" for d, c, x in passes]
    lines += [f"    python bench.py --summarize-counters <dir>", "",
              "Per dispatch, averaged over the dispatches of each kernel. `lines` are L2 read requests to memory "
              "(HBM or the MALL) by size; `read GB` = 128 B x lines128 + 64 B x lines64 + 32 B x lines32. "
              "`L2 hit` = TCC_HIT / (TCC_HIT + TCC_MISS). `dep. wait` = SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES, the "
              "share of wave time spent waiting for a dependency (an outstanding load, mostly). `VALU/wave`, "
              "`VMEM/wave` = instructions per wave.", "",
              "| kernel | lines128 | lines64 | lines32 | read GB | L2 hit | waves | dep. wait | VALU/wave | VMEM/wave |",
              "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]

    def row(k, r, h, a, b):
        l128, l64, l32 = (r.get(f"TCC_EA0_RDREQ_{s}_sum", float("nan")) for s in ("128B", "64B", "32B"))
        gb = (128 * l128 + 64 * l64 + 32 * l32) / 1e9
        ht = h.get("TCC_HIT_sum"), h.get("TCC_MISS_sum")
        hr = f"{ht[0] / (ht[0] + ht[1]):.2f}" if None not in ht and ht[0] + ht[1] > 0 else "-"
        w = a.get("SQ_WAVES")
        wait = f"{a['SQ_WAIT_INST_ANY'] / a['SQ_WAVE_CYCLES']:.2f}" if a.get("SQ_WAVE_CYCLES") else "-"
        valu = f"{a['SQ_INSTS_VALU'] / w:.0f}" if w else "-"
        vmem = f"{b['SQ_INSTS_VMEM'] / w:.0f}" if w and "SQ_INSTS_VMEM" in b else "-"
        return (f"| {k} | {l128:.4g} | {l64:.4g} | {l32:.4g} | {gb:.3f} | {hr} | {w:.4g} | {wait} | {valu} | {vmem} |"
                if w else f"| {k} | {l128:.4g} | {l64:.4g} | {l32:.4g} | {gb:.3f} | {hr} | - | - | - | - |")

    order = sorted(req, key=lambda k: -req[k].get("TCC_EA0_RDREQ_sum", 0.0))
    for k in order[:12]:
        lines.append(row(k, req[k], hit.get(k, {}), sqa.get(k, {}), sqb.get(k, {})))
    for k in creq:
        if "copyBuffer" in k:
            lines.append(row(k + " (1 GiB copy)", creq[k], {}, {}, {}))
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, f"{workload}_counters.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def summarize_batched_profile(src, out_dir):
    """Per-pass kernel time and HBM traffic of the batched workload, sequential plain runs against batched runs, from
    rocprofv3 outputs under src: for each phase P in (seq, bat) and step count S in (1, 2), P<S>/trace (--kernel-trace
    --stats), P<S>/fetch and P<S>/write (--pmc FETCH_SIZE / WRITE_SIZE) and P<S>.json (the bench's line); cfetch/ and
    cwrite/ as for summarize_profile.  One pass is the difference of the two-step and one-step runs, so graph
    generation, uploads and setup cancel.  Writes profiles/batched_rmat20-d32-256-uneven_traffic.{md,json}."""
    import collections
    import csv
    import glob

    def table(sub, pattern):
        f = sorted(glob.glob(os.path.join(src, sub, "**", pattern), recursive=True))
        return list(csv.DictReader(open(f[0]))) if f else []

    def counter_total(sub, counter):
        tot = collections.defaultdict(float)
        for r in table(sub, "*_counter_collection.csv"):
            if r.get("Counter_Name") == counter:
                tot[_short(r["Kernel_Name"])] += float(r["Counter_Value"]) * 1024.0
        return tot

    def stats(sub):
        return {_short(r["Name"]): (int(r["Calls"]), float(r["TotalDurationNs"]) / 1e6)
                for r in table(sub, "*_kernel_stats.csv")}

    def copy_scale(sub, counter):
        # Per copy launch (--copy-calibration runs several 1 GiB copies; the raw table has a row per launch).
        vals = [float(r["Counter_Value"]) * 1024.0 for r in table(sub, "*_counter_collection.csv")
                if r.get("Counter_Name") == counter and "copyBuffer" in r["Kernel_Name"]]
        return (1 << 30) / (sum(vals) / len(vals)) if vals else None

    fs, ws = copy_scale("cfetch", "FETCH_SIZE"), copy_scale("cwrite", "WRITE_SIZE")
    out = {"workload": "rmat20-d32-256-uneven-patterns", "fetch_scale": fs, "write_scale": ws, "phases": {}}
    lines = ["# Batched workload: one pass, sequential plain runs against batched runs", "",
             "R-MAT scale 20 (32-dim features) with 256 uneven planted patterns (V in [4, 32], E in [V, 2V]); one pass "
             "scores all 256 patterns.  Every figure is the difference of a two-step and a one-step run, so it covers "
             "exactly one pass (generation, uploads and setup cancel).  Under the profiler batched runs launch their "
             "iterations directly instead of replaying graphs (same kernels, same order).", "",
             "Commands (run from the repository root on one MI355X; P = sequential or batched, S = 1 and 2, "
             "SLAB / WB = the autotuned slab width and walk batch of a plain `--config batched` run):", "",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <dir>/<p><S>/trace -o t -- python -u bench.py "
             "--config batched --phase P --steps S --warmup 0 --slab-edges SLAB --walk-batch WB",
             "    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir>/<p><S>/fetch -o f -- (the same bench command)",
             "    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir>/<p><S>/write -o w -- (the same bench command)",
             "    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir>/cfetch -o copy -- python -u bench.py "
             "--copy-calibration",
             "    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir>/cwrite -o copy -- python -u bench.py "
             "--copy-calibration",
             "    python bench.py --summarize-batched <dir>", "",
             f"Counter calibration on a plain 1 GiB device copy: FETCH_SIZE x {fs:.4f}, WRITE_SIZE x {ws:.4f}." if fs
             and ws else "Counter calibration: missing.", ""]
    for ph, label in (("seq", "sequential plain runs"), ("bat", "batched runs")):
        js = {}
        for st in (1, 2):
            path = os.path.join(src, f"{ph}{st}.json")
            if os.path.exists(path):
                txt = [ln for ln in open(path).read().splitlines() if ln.startswith("{")]
                js[st] = json.loads(txt[-1]) if txt else {}
        upd = (js.get(2) or js.get(1) or {}).get("updates")
        s1, s2 = stats(f"{ph}1/trace"), stats(f"{ph}2/trace")
        f1, f2 = counter_total(f"{ph}1/fetch", "FETCH_SIZE"), counter_total(f"{ph}2/fetch", "FETCH_SIZE")
        w1, w2 = counter_total(f"{ph}1/write", "WRITE_SIZE"), counter_total(f"{ph}2/write", "WRITE_SIZE")
        kern = {}
        for k in set(s1) | set(s2):
            calls = s2.get(k, (0, 0.0))[0] - s1.get(k, (0, 0.0))[0]
            ms = s2.get(k, (0, 0.0))[1] - s1.get(k, (0, 0.0))[1]
            if calls <= 0:
                continue
            rd = (f2.get(k, 0.0) - f1.get(k, 0.0)) * (fs or float("nan"))
            wr = (w2.get(k, 0.0) - w1.get(k, 0.0)) * (ws or float("nan"))
            kern[k] = {"calls": calls, "ms": ms, "read_bytes": rd, "write_bytes": wr}
        dev_ms = sum(v["ms"] for v in kern.values())
        rd = sum(v["read_bytes"] for v in kern.values())
        wr = sum(v["write_bytes"] for v in kern.values())
        ph_out = {"device_ms": dev_ms, "read_bytes": rd, "write_bytes": wr, "updates": upd,
                  "bytes_per_update": (rd + wr) / upd if upd else None,
                  "achieved_GBps": (rd + wr) / (dev_ms * 1e-3) / 1e9 if dev_ms else None, "kernels": kern}
        out["phases"][ph] = ph_out
        lines += [f"## {label}", "",
                  f"Kernel time {dev_ms:.1f} ms per pass; HBM {rd / 1e9:.2f} GB read + {wr / 1e9:.2f} GB written"
                  + (f" = {(rd + wr) / upd:.3f} bytes per edge-pair update ({upd:.3e} updates)" if upd else "")
                  + (f"; {ph_out['achieved_GBps']:.0f} GB/s over kernel time." if dev_ms else "."), "",
                  "| kernel | calls | total ms | % | HBM read GB | HBM write GB |", "|---|---:|---:|---:|---:|---:|"]
        for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"]):
            if v["ms"] < 0.005 * dev_ms:
                continue
            lines.append(f"| {k} | {v['calls']} | {v['ms']:.2f} | {100 * v['ms'] / dev_ms:.1f} | "
                         f"{v['read_bytes'] / 1e9:.3f} | {v['write_bytes'] / 1e9:.3f} |")
        lines.append("")
    plain = os.path.join(src, "plain.json")
    if os.path.exists(plain):  # the unprofiled `--config batched` run the configuration came from
        txt = [ln for ln in open(plain).read().splitlines() if ln.startswith("{")]
        pj = json.loads(txt[-1]) if txt else {}
        out["plain_run"] = {k: pj.get(k) for k in ("sequential_ms_per_pass", "ms_per_step", "sequential_device_ms",
                                                    "batched_device_ms", "vs_sequential_plain",
                                                    "device_vs_sequential_plain", "checks")}
        lines += ["## Unprofiled run (graph replay on)", "",
                  f"Sequential {pj.get('sequential_ms_per_pass', 0):.1f} ms wall / "
                  f"{pj.get('sequential_device_ms', 0):.1f} ms device per pass; batched "
                  f"{pj.get('ms_per_step', 0):.1f} ms wall / {pj.get('batched_device_ms', 0):.1f} ms device: "
                  f"x{pj.get('vs_sequential_plain', 0):.2f} wall clock, x{pj.get('device_vs_sequential_plain', 0):.2f} "
                  f"device time; sampled patterns against sequential: {pj.get('checks')}.", ""]
    sq, bt = out["phases"].get("seq"), out["phases"].get("bat")
    if sq and bt and sq["device_ms"] and bt["device_ms"]:
        lines.append(f"Sequential / batched: kernel time x{sq['device_ms'] / bt['device_ms']:.2f}, HBM bytes "
                     f"x{(sq['read_bytes'] + sq['write_bytes']) / (bt['read_bytes'] + bt['write_bytes']):.2f}.")
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "batched_rmat20-d32-256-uneven_traffic.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(out_dir, "batched_rmat20-d32-256-uneven_traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("\n".join(lines))


def summarize_diff_profile(src, out_dir):
    """Writes profiles/diff_rmat18-d32-64-patterns.{json,md} from one GPU call: src/diff.json (the line of
    `bench.py --config diff`) and src/trace/ (rocprofv3 --kernel-trace --stats of `bench.py --config diff --steps 1
    --warmup 1`); the .md's figures are the .json's."""
    import csv
    import glob

    txt = [ln for ln in open(os.path.join(src, "diff.json")).read().splitlines() if ln.startswith("{")]
    d = json.loads(txt[-1])
    f = sorted(glob.glob(os.path.join(src, "trace", "**", "*_kernel_stats.csv"), recursive=True))
    rows = list(csv.DictReader(open(f[0]))) if f else []
    tot = sum(float(r["TotalDurationNs"]) for r in rows) / 1e6
    c = d.get("config", {})
    lines = ["# Adjoint: 64 planted 8-16-vertex patterns on R-MAT scale 18 (32-dim), learnable cdist weights", "",
             "Commands (run from the repository root on one MI355X):", "",
             f"    python -u bench.py --config diff --steps {d['steps']} --warmup {d['warmup']} > <dir>/diff.json",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <dir>/trace -o diff -- "
             "python -u bench.py --config diff --steps 1 --warmup 1",
             "    python bench.py --summarize-diff <dir>", "",
             "| pass over 64 patterns | ms |", "|---|---|",
             f"| forward (ac_run, weighted cdist) | {d['forward_ms_per_step']:.1f} |",
             f"| forward with tape ({'a checkpoint at every iteration, no recompute' if c.get('checkpoint_every') == 1 else 'checkpoint every ' + str(c.get('checkpoint_every'))}) + ac_backward of the planted "
             f"cross-entropy | {d['ms_per_step']:.1f} |", "",
             f"forward+backward / forward = **{d['value']:.2f}** (best of {d['steps']} steps after {d['warmup']} "
             f"warm-up). Checks: {d.get('checks')}.", ""]
    if rows:
        lines += [f"Kernel time of the profiled command (one warm-up and one timed step of each pass): {tot:.1f} ms; "
                  "the sixteen largest:", "", "| kernel | calls | total ms | % |", "|---|---:|---:|---:|"]
        for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:16]:
            lines.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.1f} | "
                         f"{float(r['Percentage']):.1f} |")
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "diff_rmat18-d32-64-patterns.md"), "w") as fo:
        fo.write("\n".join(lines) + "\n")
    with open(os.path.join(out_dir, "diff_rmat18-d32-64-patterns.json"), "w") as fo:
        json.dump(d, fo, indent=1)
    print("\n".join(lines))


def measured_traffic(workload, stage):
    """HBM bytes (read + write) per launch of a stage's kernel from the committed profile summary, or None."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    prefixes = [p for p, st in STAGE_OF_KERNEL.items() if st == stage]
    for k, v in json.load(open(path)).get("kernels", {}).items():
        if any(k.startswith(p) for p in prefixes) and v.get("read_bytes") is not None and v.get("write_bytes") is not None:
            return v["read_bytes"] + v["write_bytes"]
    return None


def measure_diff(steps, warmup, K=1):
    """Differentiation: R-MAT scale 18 (32-dim) with 64 planted 8-16-vertex patterns; per pattern the forward
    (prep + iterations + output) alone, then forward with the tape (a checkpoint every K iterations; with K = 1 the run
    also records every walk's argmax rows, so ac_backward recomputes nothing) + ac_backward of
    the planted cross-entropy.  A step is one pass over the 64 patterns; the figure is the best step of each side."""
    from application_classification_amd import workloads as W
    from application_classification_amd.matcher import AC_OPT_TAPE, Matcher

    t = time.time()
    w = W.rmat(18, 32, 12, 16, seed=0)
    g = w.data
    pats = W.uneven_patterns(g, 64, seed=7, v_range=(8, 16))
    upd = sum(g.num_edges * p.num_edges * p.num_nodes for p in pats)
    log(f"diff workload: dn={g.num_nodes} de={g.num_edges}, {len(pats)} patterns, {upd:.3e} edge-pair updates, "
        f"generated in {time.time() - t:.1f} s")
    rng = np.random.default_rng(0)
    wv = rng.uniform(0.5, 1.5, 32).astype(np.float32)
    we = rng.uniform(0.5, 1.5, 32).astype(np.float32)
    grads = [np.zeros((len(pats), 32)), np.zeros((len(pats), 32))]

    def fwd_pass(m, tape):
        t0 = time.perf_counter()
        for i, p in enumerate(pats):
            m.set_pattern(p)
            mu = m.run()
            if tape:
                gm = np.zeros_like(mu)
                gm[p.vmap, np.arange(p.num_nodes)] = -1.0 / p.num_nodes
                gv, ge = m.backward(gm)
                grads[0][i], grads[1][i] = gv, ge
        return (time.perf_counter() - t0) * 1e3

    times = {}
    for tape in (False, True):
        m = Matcher(0)
        m.set_option(AC_OPT_TAPE, K if tape else 0)
        m.set_data(g)
        m.set_weights(wv, we)
        for _ in range(warmup):
            fwd_pass(m, tape)
        times[tape] = min(fwd_pass(m, tape) for _ in range(steps))
        m.close()
        log(f"{'forward+backward' if tape else 'forward'}: {times[tape]:.1f} ms per pass over {len(pats)} patterns")
    return {
        "ratio": times[True] / times[False],
        "forward_backward_ms": times[True],
        "forward_ms": times[False],
        "config": {"workload": "rmat18-d32-64-patterns-8to16", "data_vertices": g.num_nodes,
                   "data_edges": g.num_edges, "patterns": len(pats), "checkpoint_every": K,
                   "updates_per_pass": upd, "parallelism": "single-gpu"},
        "checks": {"grad_wv_norm": float(np.linalg.norm(grads[0].sum(0))),
                   "grad_we_norm": float(np.linalg.norm(grads[1].sum(0))),
                   "all_finite": bool(np.isfinite(grads[0]).all() and np.isfinite(grads[1]).all())},
    }


def bench_diff(args):
    r = measure_diff(args.steps, args.warmup, args.checkpoint_every)
    every = "every iteration" if args.checkpoint_every == 1 else f"every {args.checkpoint_every}"
    out = {
        "metric": f"forward+backward time / forward time (64 patterns, R-MAT 18, 32-dim, checkpoint {every})",
        "value": r["ratio"],
        "unit": "ratio",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["forward_backward_ms"],
        "forward_ms_per_step": r["forward_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE,
        "gradient_dtype": "f32",
        "data": "synthetic",
        "config": r["config"],
        "checks": r["checks"],
    }
    print(json.dumps(out), flush=True)


def measure_batched(steps, warmup, phase="both", tuned=None):
    """256 uneven planted patterns (V in [4, 32], E in [V, 2V]) against one resident R-MAT scale-20 graph (32-dim):
    the whole job (prep + iterations + output of every pattern) as 256 sequential plain runs, then as batched runs
    (autotuned slabs, graph-replayed iterations).  A step is one pass over all 256 patterns; figures are the best
    step of each side.  phase "sequential" or "batched" times one side only (for per-side hardware counters); tuned
    = (slab_edges, walk_batch) skips the autotuner."""
    from application_classification_amd import workloads as W
    from application_classification_amd.matcher import AC_OPT_GRAPH, Matcher

    def allclose_report(a, b, rtol, atol):
        err = np.abs(a.astype(np.float64) - b.astype(np.float64))
        tol = atol + rtol * np.abs(b.astype(np.float64))
        return int((err > tol).sum()), float((err / tol).max())

    t = time.time()
    g, pats = W.batched(0)
    upd = sum(g.num_edges * p.num_edges * p.num_nodes for p in pats)
    log(f"batched workload: dn={g.num_nodes} de={g.num_edges}, {len(pats)} patterns, "
        f"sum pe={sum(p.num_edges for p in pats)}, sum pn={sum(p.num_nodes for p in pats)}, "
        f"{upd:.3e} edge-pair updates, generated in {time.time() - t:.1f} s")
    m = Matcher(0)
    m.set_data(g)
    sample = list(range(0, len(pats), 32))
    # Sequential plain runs: one ac_set_graph + ac_run per pattern, as the reference's driver scores one pattern.
    # The sampled patterns' results are read back before the timed passes (neither side times its copies).
    ref = {}
    if phase == "both":
        for i in sample:
            m.set_pattern(pats[i])
            m.run()
            ref[i] = m.mu()
    seq_ms = []
    for step in range(steps if phase != "batched" else 0):
        dev = 0.0
        t0 = time.perf_counter()
        for i, p in enumerate(pats):
            m.set_pattern(p)
            m.run_async()
            m.sync()
            a, b = m.timing()
            dev += a + b
        seq_ms.append(((time.perf_counter() - t0) * 1e3, dev))
        log(f"sequential step {step + 1}/{steps}: wall {seq_ms[-1][0]:.1f} ms, device {dev:.1f} ms")
    if phase == "sequential":
        m.close()
        t_seq = min(w for w, _ in seq_ms)
        return {"value": upd / (t_seq * 1e-3), "sequential_ms_per_pass": t_seq,
                "sequential_device_ms": min(d for _, d in seq_ms), "updates": upd}
    m.batch_set_patterns(pats)
    if tuned:
        from application_classification_amd.matcher import AC_OPT_SLAB_EDGES, AC_OPT_WALK_BATCH

        cap, wb = tuned
        m.set_option(AC_OPT_SLAB_EDGES, cap)
        m.set_option(AC_OPT_WALK_BATCH, wb)
        log(f"slab edges {cap}, walk batch {wb} (given)")
    else:
        t = time.time()
        (cap, wb), est = m.batch_autotune(2)
        log(f"autotune: slab edges {cap}, walk batch {wb} (estimated {est:.1f} ms) in {time.time() - t:.1f} s")
    m.set_option(AC_OPT_GRAPH, 1)
    for i in range(warmup):
        m.batch_run()
        log(f"batched warmup {i + 1}/{warmup}: device {sum(m.batch_timing()):.1f} ms")
    bat_ms = []
    for step in range(steps):
        t0 = time.perf_counter()
        m.batch_run()
        wall = (time.perf_counter() - t0) * 1e3
        a, b = m.batch_timing()
        bat_ms.append((wall, a + b, a, b))
        log(f"batched step {step + 1}/{steps}: wall {wall:.1f} ms, device {a + b:.1f} ms (prep {a:.1f}, "
            f"loop {b:.1f})")
    info = m.batch_info()
    worst = 0.0
    bad = 0
    for i in (sample if phase != "batched" else []):
        nb, ratio = allclose_report(m.batch_mu(i), ref[i], 1e-5, 1e-8)
        bad += nb
        worst = max(worst, ratio)
    m.close()
    if phase == "batched":
        t_bat = min(w for w, *_ in bat_ms)
        return {"value": upd / (t_bat * 1e-3), "ms_per_pass": t_bat, "batched_device_ms": min(d for _, d, *_ in bat_ms),
                "updates": upd, "slabs": info["slabs"], "slab_edges": cap, "walk_batch": wb,
                "updates_computed": info["updates_computed"]}
    t_seq = min(w for w, _ in seq_ms)
    t_bat = min(w for w, *_ in bat_ms)
    dev_seq = min(d for _, d in seq_ms)
    dev_bat = min(d for _, d, *_ in bat_ms)
    return {
        "value": upd / (t_bat * 1e-3),
        "vs_sequential_plain": t_seq / t_bat,
        "device_vs_sequential_plain": dev_seq / dev_bat,
        "ms_per_pass": t_bat,
        "sequential_ms_per_pass": t_seq,
        "sequential_device_ms": dev_seq,
        "batched_device_ms": dev_bat,
        "batched_prep_ms": min(a for *_, a, _ in bat_ms),
        "batched_loop_ms": min(b for *_, b in bat_ms),
        "config": {"workload": "rmat20-d32-256-uneven-patterns", "data_vertices": g.num_nodes,
                   "data_edges": g.num_edges, "patterns": len(pats),
                   "sum_pattern_edges": int(sum(p.num_edges for p in pats)), "slab_edges": cap,
                   "walk_batch": wb, "slabs": info["slabs"], "updates": info["updates"],
                   "updates_computed": info["updates_computed"], "parallelism": "single-gpu"},
        "checks": {"sampled_patterns": len(sample), "violations_vs_sequential": int(bad),
                   "max_err_over_tol_vs_sequential": worst},
    }


def bench_batched(args):
    tuned = (args.slab_edges, args.walk_batch) if args.slab_edges else None
    r = measure_batched(args.steps, args.warmup, args.phase, tuned)
    if args.phase != "both":
        print(json.dumps(dict(r, phase=args.phase, metric="batched workload, one side only", steps=args.steps,
                              warmup=args.warmup)), flush=True)
        return
    out = {
        "metric": "batched edge-pair updates/s (256 patterns, whole job) vs 256 sequential plain runs",
        "value": r["value"],
        "unit": "edge-pair updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_pass"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE,
        "data": "synthetic",
    }
    out.update({k: v for k, v in r.items() if k not in ("value", "ms_per_pass")})
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="s22", choices=["s22", "s18", "small", "batched", "diff"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--phase", default="both", choices=["both", "sequential", "batched"],
                    help="--config batched: time both sides (default) or one (for per-side hardware counters)")
    ap.add_argument("--slab-edges", type=int, default=0,
                    help="--config batched: slab width in pattern edges (with --walk-batch; 0 = autotune)")
    ap.add_argument("--walk-batch", type=int, default=0, help="--config batched: walk batch (with --slab-edges)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched and forward+backward measurements that the default line carries")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-pruned", action="store_true",
                    help="skip the comparison runs, pruned and with FP32 costs (profiled commands then hold the "
                         "headline launches only)")
    ap.add_argument("--summarize-profile", metavar="DIR",
                    help="summarise rocprofv3 outputs under DIR into profiles/ and exit")
    ap.add_argument("--summarize-counters", metavar="DIR",
                    help="summarise the counter passes (req, creq, hit, sqa, sqb) under DIR into profiles/ and exit")
    ap.add_argument("--summarize-batched", metavar="DIR",
                    help="summarise the batched workload's per-pass rocprofv3 outputs under DIR into profiles/ and "
                         "exit")
    ap.add_argument("--summarize-diff", metavar="DIR",
                    help="write the adjoint profile (json and md) into profiles/ from DIR/diff.json and DIR/trace, "
                         "and exit")
    ap.add_argument("--out", metavar="DIR", help="--summarize-batched: write the summary here instead of profiles/")
    ap.add_argument("--checkpoint-every", type=int, default=1,
                    help="the adjoint measurement's checkpoint interval (AC_OPT_TAPE); 1 keeps every state and the "
                         "walks' argmax rows, so the backward recomputes nothing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (scales rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE to bytes), then time ten more (the copy's bandwidth, printed as one JSON line) "
                         "and exit")
    args = ap.parse_args()
    if args.summarize_profile:
        from application_classification_amd import workloads as W

        name = {"s22": "rmat22-d64-p16x24", "s18": "rmat18-d32-p12x16", "small": "small-er1000"}[args.config]
        summarize_profile(args.summarize_profile, name, os.path.join(ROOT, "profiles"))
        return
    if args.summarize_counters:
        name = {"s22": "rmat22-d64-p16x24", "s18": "rmat18-d32-p12x16", "small": "small-er1000"}[args.config]
        summarize_counters(args.summarize_counters, name, args.out or os.path.join(ROOT, "profiles"))
        return
    if args.summarize_diff:
        summarize_diff_profile(args.summarize_diff, args.out or os.path.join(ROOT, "profiles"))
        return
    if args.summarize_batched:
        summarize_batched_profile(args.summarize_batched, args.out or os.path.join(ROOT, "profiles"))
        return
    if args.copy_calibration:
        import torch

        n = 1 << 28
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied 3 x {n * 4} bytes", flush=True)
        # A plain device copy's bandwidth (read + write), HIP events around ten launches on torch's current stream
        # (where they run).
        def timed(fn, reps=10):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            fn()
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps / 1e3
        t_copy = timed(lambda: b.copy_(a))
        print(json.dumps({"copy_GBps": 2 * n * 4 / t_copy / 1e9, "bytes": n * 4}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the matcher is a single-GPU path here (multi-GPU sharding is a design "
              "note, not built); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.config == "batched":
        bench_batched(args)
        return
    if args.config == "diff":
        bench_diff(args)
        return

    from application_classification_amd import workloads as W
    from application_classification_amd.matcher import Matcher

    t = time.time()
    w = W.CONFIGS[args.config](0)
    d, p = w.data, w.pattern
    dn, de, pn, pe = d.num_nodes, d.num_edges, p.num_nodes, p.num_edges
    name, dim = w.name, int(d.vfeat.shape[1])
    log(f"workload {w.name}: dn={dn} de={de} pn={pn} pe={pe} generated in {time.time() - t:.1f} s")

    from application_classification_amd.matcher import AC_OPT_PACK_COSTS, AC_OPT_PRUNE

    m = Matcher(0)
    m.set_option(AC_OPT_PRUNE, 0)  # the headline walks every edge; the exact pruning is measured separately below
    t = time.time()
    m.set_data(d)
    m.set_pattern(p)
    log(f"graphs resident on the GPU in {time.time() - t:.1f} s")
    for i in range(args.warmup):
        m.run_async()
        m.sync()
        log(f"warmup {i + 1}/{args.warmup}: prep {m.timing()[0]:.3f} ms, loop {m.timing()[1]:.3f} ms")

    try:
        import torch

        sync_all = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing here; the matcher does not need it
        sync_all = lambda: None  # noqa: E731
    loop_ms = prep_ms = 0.0
    m.sync()
    sync_all()
    t0 = time.perf_counter()
    for i in range(args.steps):
        m.run_async()
        m.sync()
        a, b = m.timing()
        prep_ms += a
        loop_ms += b
    m.sync()
    sync_all()
    t1 = time.perf_counter()
    wall_ms = (t1 - t0) * 1e3 / args.steps
    loop_ms /= args.steps
    prep_ms /= args.steps
    log(f"timed {args.steps} steps: {wall_ms:.3f} ms/step (device: prep {prep_ms:.3f} ms, loop {loop_ms:.3f} ms)")
    nb, packed = m.cost_packing()
    packing = {"bundles": nb, "packed_bundles": packed, "frac": packed / max(nb, 1)}
    log(f"cost packing (AC_OPT_PACK_COSTS): {packed} of {nb} edge bundles walked as 24-bit codes")

    mu = m.mu()
    checks = {
        "planted_recovered": bool((mu.argmax(axis=0) == p.vmap).all()),
        "max_abs_column_logsumexp": float(np.abs(np.log(np.exp(mu.astype(np.float64)).sum(axis=0))).max()),
        "all_log_probs_nonpositive": bool((mu <= 0).all()),
    }
    log(f"checks: {checks}")

    updates = de * pe * pn
    B = algorithmic_bytes(dn, de, pn, pe)
    iter_frac = B["iteration"] * pn / (loop_ms * 1e-3) / HBM_PEAK
    roofline = None
    stages = None
    if not args.no_profile:
        stages = m.run_profiled()
        loop_stages = {k: v for k, v in stages.items() if k not in ("prep",)}
        dom = max(("edge_walk", "vertex_a", "vertex_c"), key=lambda k: loop_stages[k])
        avg_ms = stages[dom] / pn
        achieved = B[dom] / (avg_ms * 1e-3)
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": measured_traffic(name, dom),
                    "alg_bytes_per_launch": B[dom], "avg_launch_ms": avg_ms}
        log(f"stages (ms, summed over {pn} iterations): {json.dumps({k: round(v, 4) for k, v in stages.items()})}")

    # Exact pruning (AC_OPT_PRUNE, the library default): the same run skipping edge bundles that provably cannot
    # raise FMax above the fill; the output must be bit-identical.
    pruned = None
    if not args.no_pruned:
        m.set_option(AC_OPT_PRUNE, 1)
        m.run_async()
        m.sync()
        pr_loop = pr_prep = 0.0
        for i in range(args.steps):
            m.run_async()
            m.sync()
            a, b = m.timing()
            pr_prep += a
            pr_loop += b
        pr_loop /= args.steps
        pr_prep /= args.steps
        pruned = {"value": updates / (pr_loop * 1e-3), "loop_ms_per_step": pr_loop, "prep_ms_per_step": pr_prep,
                  "speedup_loop": loop_ms / pr_loop, "bit_identical": bool(np.array_equal(m.mu(), mu))}
        log(f"pruned: loop {pr_loop:.3f} ms ({pruned['speedup_loop']:.2f}x), identical: {pruned['bit_identical']}")
        m.set_option(AC_OPT_PRUNE, 0)

    # The same headline run with every edge cost read as FP32 (AC_OPT_PACK_COSTS = 0): the 24-bit codes are
    # lossless, so the output must be bit-identical.
    fp32_costs = None
    if not args.no_pruned:
        m.set_option(AC_OPT_PACK_COSTS, 0)
        m.run_async()
        m.sync()
        u_loop = u_prep = 0.0
        for i in range(args.steps):
            m.run_async()
            m.sync()
            a, b = m.timing()
            u_prep += a
            u_loop += b
        u_loop /= args.steps
        u_prep /= args.steps
        fp32_costs = {"value": updates / (u_loop * 1e-3), "loop_ms_per_step": u_loop, "prep_ms_per_step": u_prep,
                      "hbm_roofline_frac_iteration": B["iteration"] * pn / (u_loop * 1e-3) / HBM_PEAK,
                      "bit_identical": bool(np.array_equal(m.mu(), mu))}
        log(f"FP32 costs: loop {u_loop:.3f} ms, identical: {fp32_costs['bit_identical']}")
        m.set_option(AC_OPT_PACK_COSTS, 1)

    m.close()
    del m, mu, w, d, p  # the R-MAT 22 graphs leave the GPU and the host before the other measurements

    # The metric's other two parts: batched against sequential plain runs (one warmup pass and one timed pass each),
    # and forward+backward against forward (one warmup pass and the best of three, since a single pass of either side
    # can be slowed by what ran just before it); --config batched / diff time them with --steps / --warmup.
    batched = fwd_bwd = None
    if not args.no_extras:
        # Forward+backward first: its forward pass is host-heavy (64 small runs), and right after the batched
        # measurement the release of that workload's host arrays slowed it by about a third.
        import gc
        gc.collect()
        log("forward+backward: 64 patterns on R-MAT 18 ...")
        f = measure_diff(3, 1, args.checkpoint_every)
        fwd_bwd = {"ratio": f["ratio"], "forward_backward_ms": f["forward_backward_ms"], "forward_ms": f["forward_ms"],
                   "patterns": f["config"]["patterns"], "checkpoint_every": args.checkpoint_every,
                   "grads_finite": f["checks"]["all_finite"]}
        log(f"forward+backward: {json.dumps(fwd_bwd)}")
        gc.collect()
        log("batched: 256 uneven patterns on R-MAT 20 ...")
        b = measure_batched(1, 1)
        batched = {"vs_sequential_plain": b["vs_sequential_plain"],
                   "device_vs_sequential_plain": b["device_vs_sequential_plain"], "value": b["value"],
                   "ms_per_pass": b["ms_per_pass"], "sequential_ms_per_pass": b["sequential_ms_per_pass"],
                   "slabs": b["config"]["slabs"], "slab_edges": b["config"]["slab_edges"],
                   "max_err_over_tol_vs_sequential": b["checks"]["max_err_over_tol_vs_sequential"],
                   "sampled_patterns": b["checks"]["sampled_patterns"]}
        log(f"batched: {json.dumps(batched)}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"cpu baseline: C++ restatement, FP64 and FP32 on {threads} threads, FP64 on 1 thread ...")
        cpu = cpu_baseline(threads)
        log(f"cpu baseline: {json.dumps(cpu)}")

    out = {
        "metric": "edge-pair updates/s per iteration + % HBM roofline",
        "value": updates / (loop_ms * 1e-3),
        "unit": "edge-pair updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_ms,
        "loop_ms_per_step": loop_ms,
        "prep_ms_per_step": prep_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE,
        "data": "synthetic",
        "config": {"workload": name, "data_vertices": dn, "data_edges": de, "pattern_vertices": pn,
                   "pattern_edges": pe, "feature_dim": dim, "iterations": pn, "parallelism": "single-gpu"},
        "hbm_roofline_frac_iteration": iter_frac,
        "alg_bytes_per_iteration": B["iteration"],
        "roofline": roofline,
        "cost_packing": packing,
        "fp32_costs": fp32_costs,
        "pruned": pruned,
        "cpu_baseline": cpu,
        "batched": batched,
        "fwd_bwd": fwd_bwd,
        "stage_ms": stages,
        "checks": checks,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
