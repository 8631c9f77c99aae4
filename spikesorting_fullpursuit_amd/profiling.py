"""Summaries of rocprofv3 output for the pursuit kernels.

Collect and summarise in one command (runs rocprofv3 four times over the
program after --: kernel trace, FETCH_SIZE, WRITE_SIZE, matrix-core counters):

    python -m spikesorting_fullpursuit_amd.profiling collect --dir OUTDIR \
        --copy-bytes 768000000 --title TITLE --md OUT.md --json OUT.json \
        -- python3 bench.py --steps 1 ...

or summarise existing output:

    python -m spikesorting_fullpursuit_amd.profiling --stats DIR/stats_kernel_stats.csv \
        --fetch DIR/fetch_counter_collection.csv --write DIR/write_counter_collection.csv \
        --mfma DIR/mfma_counter_collection.csv --copy-bytes 768000000 --md OUT.md --json OUT.json

--stats is the `rocprofv3 --kernel-trace --stats --output-format csv` kernel
table; --fetch / --write are `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
counter collections of the same command.  FETCH_SIZE and WRITE_SIZE are in KiB
per dispatch; each is scaled by what it reads for the device-to-device copy of
the voltage (__amd_rocclr_copyBuffer moving --copy-bytes bytes), so the
figures are bytes moved to or from HBM per dispatch.  --mfma is a collection
of MfmaUtil (percent of SIMD cycles with the matrix cores busy) and
MfmaFlopsBF16 / MfmaFlopsF16 (matrix-core flop).
"""
import argparse
import collections
import csv
import json
import re


def kernel_name(raw):
    s = raw.replace("void ", "").replace("fbp::(anonymous namespace)::", "").replace("fbp::", "")
    m = re.match(r"([A-Za-z_0-9]+)(<[^(]*>)?", s)
    return (m.group(1) + (m.group(2) or "")) if m else s


def read_stats(path):
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append({"kernel": kernel_name(r["Name"]), "calls": int(r["Calls"]),
                     "total_ms": float(r["TotalDurationNs"]) / 1e6, "avg_us": float(r["AverageNs"]) / 1e3,
                     "pct": float(r["Percentage"])})
    return rows


def read_counter(path, counter):
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        key = (r.get("Dispatch_Id") or r.get("Correlation_Id"), kernel_name(r["Kernel_Name"]))
        per[key].append(float(r["Counter_Value"]))
    out = collections.defaultdict(list)
    for (_, k), vals in per.items():
        out[k].append(sum(vals))   # summed over the counter's instances (TCC channels)
    return out


def calibration(per_kernel, copy_bytes):
    """KiB counted for the largest copyBuffer dispatch, against its bytes."""
    vals = per_kernel.get("__amd_rocclr_copyBuffer", [])
    if not vals:
        return None
    kib = max(vals)
    return (copy_bytes / 1024.0) / kib if kib > 0 else None


MFMA_COUNTERS = ("MfmaUtil", "MfmaFlopsBF16", "MfmaFlopsF16")


def read_mfma(path):
    """Per kernel: mean MfmaUtil (percent) and matrix-core flop per dispatch."""
    out = {}
    util = read_counter(path, "MfmaUtil")
    flops = collections.defaultdict(list)
    for c in MFMA_COUNTERS[1:]:
        for k, v in read_counter(path, c).items():
            flops[k].extend(v)
    for k in set(util) | set(flops):
        u, fl = util.get(k, []), flops.get(k, [])
        out[k] = {"mfma_util_pct": sum(u) / len(u) if u else None,
                  "mfma_flop_per_dispatch": sum(fl) / max(1, len(u) or len(fl)) if fl else None}
    return out


def summarize(stats, fetch, write, copy_bytes, mfma=None):
    rows = read_stats(stats)
    f = read_counter(fetch, "FETCH_SIZE") if fetch else {}
    w = read_counter(write, "WRITE_SIZE") if write else {}
    cf = calibration(f, copy_bytes) if f else None
    cw = calibration(w, copy_bytes) if w else None
    traffic = {}
    for k in set(f) | set(w):
        fv, wv = f.get(k, []), w.get(k, [])
        traffic[k] = {
            "dispatches": max(len(fv), len(wv)),
            "fetch_bytes_per_dispatch": (sum(fv) / len(fv) * 1024.0 * cf) if fv and cf else None,
            "write_bytes_per_dispatch": (sum(wv) / len(wv) * 1024.0 * cw) if wv and cw else None,
        }
    m = read_mfma(mfma) if mfma else {}
    return {"kernels": rows, "traffic": traffic, "fetch_scale": cf, "write_scale": cw, "copy_bytes": copy_bytes,
            "mfma": {k: v for k, v in m.items() if v["mfma_flop_per_dispatch"]}}


def markdown(s, title, commands):
    out = ["# " + title, "", "Commands (repository root, one MI355X, TMPDIR=/tmp):", ""]
    out += ["    " + c for c in commands]
    out += ["", "## Kernel time (rocprofv3 --kernel-trace --stats)", "",
            "| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for r in s["kernels"]:
        out.append("| %s | %d | %.3f | %.1f | %.2f |" % (r["kernel"], r["calls"], r["total_ms"], r["avg_us"], r["pct"]))
    if s["traffic"]:
        out += ["", "## HBM traffic per dispatch (FETCH_SIZE, WRITE_SIZE)", "",
                "Calibration: the %d-byte device-to-device voltage copy (__amd_rocclr_copyBuffer); FETCH_SIZE "
                "scaled by %s, WRITE_SIZE by %s." % (s["copy_bytes"], "%.3f" % s["fetch_scale"] if s["fetch_scale"]
                                                   else "n/a", "%.3f" % s["write_scale"] if s["write_scale"] else "n/a"),
                "", "| kernel | dispatches | fetch MB/dispatch | write MB/dispatch |", "|---|---|---|---|"]
        order = {r["kernel"]: -r["total_ms"] for r in s["kernels"]}
        for k in sorted(s["traffic"], key=lambda k: order.get(k, 0)):
            t = s["traffic"][k]
            fb, wb = t["fetch_bytes_per_dispatch"], t["write_bytes_per_dispatch"]
            out.append("| %s | %d | %s | %s |" % (k, t["dispatches"], "%.2f" % (fb / 1e6) if fb is not None else "-",
                                                  "%.2f" % (wb / 1e6) if wb is not None else "-"))
    if s.get("mfma"):
        avg = {r["kernel"]: r["avg_us"] for r in s["kernels"]}
        out += ["", "## Matrix cores (MfmaUtil, MfmaFlopsBF16 + MfmaFlopsF16)", "",
                "MfmaUtil: percent of SIMD cycles with the matrix cores busy during the dispatch; TFLOP/s: "
                "matrix-core flop per dispatch over the traced average duration.", "",
                "| kernel | MfmaUtil % | matrix-core GFLOP/dispatch | TFLOP/s |", "|---|---|---|---|"]
        for k, v in sorted(s["mfma"].items(), key=lambda kv: -(kv[1]["mfma_flop_per_dispatch"] or 0)):
            fl = v["mfma_flop_per_dispatch"]
            t = avg.get(k)
            out.append("| %s | %s | %.3f | %s |" % (k, "%.1f" % v["mfma_util_pct"] if v["mfma_util_pct"] is not None
                                                     else "-", fl / 1e9,
                                                     "%.1f" % (fl / (t * 1e-6) / 1e12) if t else "-"))
    return "\n".join(out) + "\n"


def collect(out_dir, program, copy_bytes, title, md=None, js=None, timeout=600):
    """Run `program` (a list, python3 first) under rocprofv3 four times --
    kernel trace with statistics, FETCH_SIZE, WRITE_SIZE, matrix-core
    counters -- into out_dir and summarise; each pass is killed at `timeout`
    seconds.  Returns the summary."""
    import glob
    import os
    import shlex
    import subprocess
    os.makedirs(out_dir, exist_ok=True)
    passes = [("stats", ["--kernel-trace", "--stats"]), ("fetch", ["--pmc", "FETCH_SIZE"]),
              ("write", ["--pmc", "WRITE_SIZE"]), ("mfma", ["--pmc"] + list(MFMA_COUNTERS))]
    commands = []
    for name, flags in passes:
        cmd = ["rocprofv3"] + flags + ["-d", os.path.join(out_dir, name), "-o", name, "--output-format", "csv",
                                       "--"] + list(program)
        commands.append(" ".join(shlex.quote(c) for c in cmd))
        print("profiling pass %s: %s" % (name, commands[-1]), flush=True)
        with open(os.path.join(out_dir, name + ".log"), "w") as log:
            subprocess.run(["timeout", "-s", "KILL", str(timeout)] + cmd, stdout=log, stderr=subprocess.STDOUT,
                           check=True)

    def one(name, suffix):
        hits = glob.glob(os.path.join(out_dir, name, "**", "*" + suffix), recursive=True)
        return hits[0] if hits else None

    s = summarize(one("stats", "kernel_stats.csv"), one("fetch", "counter_collection.csv"),
                  one("write", "counter_collection.csv"), copy_bytes, one("mfma", "counter_collection.csv"))
    s["commands"] = commands
    if md:
        open(md, "w").write(markdown(s, title, commands))
    if js:
        json.dump(s, open(js, "w"), indent=1)
    return s


def main():
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "collect":
        argv = sys.argv[2:]
        if "--" not in argv:
            raise SystemExit("collect: the program to profile follows --")
        k = argv.index("--")
        cp = argparse.ArgumentParser()
        cp.add_argument("--dir", required=True)
        cp.add_argument("--copy-bytes", type=int, required=True)
        cp.add_argument("--title", default="Kernel time, HBM traffic and matrix cores")
        cp.add_argument("--md")
        cp.add_argument("--json")
        cp.add_argument("--timeout", type=int, default=600)
        a = cp.parse_args(argv[:k])
        collect(a.dir, argv[k + 1:], a.copy_bytes, a.title, a.md, a.json, a.timeout)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--mfma")
    ap.add_argument("--copy-bytes", type=int, required=True)
    ap.add_argument("--title", default="Kernel time and HBM traffic")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--md")
    ap.add_argument("--json")
    a = ap.parse_args()
    s = summarize(a.stats, a.fetch, a.write, a.copy_bytes, a.mfma)
    if a.md:
        open(a.md, "w").write(markdown(s, a.title, a.command))
    if a.json:
        json.dump(s, open(a.json, "w"), indent=1)
    if not a.md and not a.json:
        print(markdown(s, a.title, a.command))


if __name__ == "__main__":
    main()
