"""Summaries of rocprofv3 output for the pursuit kernels.

    python -m spikesorting_fullpursuit_amd.profiling --stats DIR/stats_kernel_stats.csv \
        --fetch DIR/fetch_counter_collection.csv --write DIR/write_counter_collection.csv \
        --copy-bytes 768000000 --md OUT.md --json OUT.json

--stats is the `rocprofv3 --kernel-trace --stats --output-format csv` kernel
table; --fetch / --write are `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
counter collections of the same command.  FETCH_SIZE and WRITE_SIZE are in KiB
per dispatch; each is scaled by what it reads for the device-to-device copy of
the voltage (__amd_rocclr_copyBuffer moving --copy-bytes bytes), so the
figures are bytes moved to or from HBM per dispatch.
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


def summarize(stats, fetch, write, copy_bytes):
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
    return {"kernels": rows, "traffic": traffic, "fetch_scale": cf, "write_scale": cw, "copy_bytes": copy_bytes}


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
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-bytes", type=int, required=True)
    ap.add_argument("--title", default="Kernel time and HBM traffic")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--md")
    ap.add_argument("--json")
    a = ap.parse_args()
    s = summarize(a.stats, a.fetch, a.write, a.copy_bytes)
    if a.md:
        open(a.md, "w").write(markdown(s, a.title, a.command))
    if a.json:
        json.dump(s, open(a.json, "w"), indent=1)
    if not a.md and not a.json:
        print(markdown(s, a.title, a.command))


if __name__ == "__main__":
    main()
