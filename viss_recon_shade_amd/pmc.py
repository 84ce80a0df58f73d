"""Turns rocprofv3 counter runs into the per-kernel figures bench.py reports.

    # a plain device copy of known size, the calibration for FETCH_SIZE / WRITE_SIZE
    rocprofv3 --pmc FETCH_SIZE -f csv -d OUT/copy_fetch -o copy -- python -m viss_recon_shade_amd.pmc copy
    rocprofv3 --pmc WRITE_SIZE -f csv -d OUT/copy_write -o copy -- python -m viss_recon_shade_amd.pmc copy
    # the workload, one counter per run
    rocprofv3 --pmc FETCH_SIZE -f csv -d OUT/fetch -o bench -- python bench.py ...
    rocprofv3 --pmc WRITE_SIZE -f csv -d OUT/write -o bench -- python bench.py ...
    python -m viss_recon_shade_amd.pmc traffic --fetch OUT/fetch --write OUT/write \\
        --copy-fetch OUT/copy_fetch --copy-write OUT/copy_write --out profiles/traffic_c2.json

    # instruction mix: any set of SQ_* / GRBM_* runs of the same command
    python -m viss_recon_shade_amd.pmc mix OUT/passA OUT/passB ... --out mix.json

    # markdown table of a --kernel-trace --stats run
    python -m viss_recon_shade_amd.pmc stats OUT/trace

    # per-sort breakdown of the uvw gridding in a --kernel-trace run (one table
    # per vrs_grid_uvw call, keyed by its step count)
    python -m viss_recon_shade_amd.pmc sorts OUT/trace

Counter values are summed over the dimensions rocprofv3 reports (shader
engines, XCDs) per dispatch and averaged over the dispatches of each kernel.
"""
import argparse
import collections
import csv
import glob
import json
import os
import re
import subprocess

COPY_BYTES = 256 * 1024 * 1024   # bytes read and bytes written by one calibration copy


def _csvs(path):
    if os.path.isfile(path):
        return [path]
    return sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True))


def _demangle(name):
    if not name.startswith("_Z"):
        return name
    try:
        out = subprocess.run(["c++filt", name], capture_output=True, text=True,
                             check=True).stdout.strip()
        if out and not out.startswith("_Z"):
            return out
    except (OSError, subprocess.CalledProcessError):
        pass
    # c++filt too old for the name (e.g. _Float16 arguments): for a plain nested
    # name _ZN <len><id>... E <args> keep the last identifier
    m = re.match(r"_ZN((?:\d+\w+?)+)E", name)
    if m:
        ids, rest = [], m.group(1)
        while rest and rest[0].isdigit():
            n = int(re.match(r"\d+", rest).group(0))
            rest = rest[len(str(n)):]
            ids.append(rest[:n])
            rest = rest[n:]
        if ids:
            return ids[-1]
    return name


def short_name(kernel):
    """Demangled kernel name without return type, namespaces and argument list."""
    k = _demangle(kernel.strip())
    if k.endswith(")"):                  # drop the argument list: the last balanced (...)
        depth = 0
        for i in range(len(k) - 1, -1, -1):
            depth += {")": 1, "(": -1}.get(k[i], 0)
            if depth == 0:
                k = k[:i]
                break
    k = re.sub(r"^void\s+", "", k)
    k = re.sub(r"^vrs::(\(anonymous namespace\)::)?", "", k)
    return k


def per_kernel(paths, skip=("__amd_rocclr",)):
    """{kernel: {counter: mean over dispatches, "_dispatches": n, "_ms": mean duration}}."""
    disp = collections.defaultdict(lambda: collections.defaultdict(float))
    dur = {}
    for path in paths:
        for f in _csvs(path):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    name = r["Kernel_Name"]
                    if any(s in name for s in skip):
                        continue
                    key = (short_name(name), f, r["Dispatch_Id"])
                    disp[key][r["Counter_Name"]] += float(r["Counter_Value"])
                    dur[key] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for key, cs in disp.items():
        for c, v in cs.items():
            acc[key[0]][c].append(v)
        acc[key[0]]["_ms"].append(dur[key])
    out = {}
    for k, cs in acc.items():
        out[k] = {c: sum(v) / len(v) for c, v in cs.items()}
        out[k]["_dispatches"] = len(cs["_ms"])
    return out


def copy_kib(path, counter):
    """Largest per-dispatch reading of `counter` among the calibration copy's kernels."""
    best = 0.0
    for f in _csvs(path):
        with open(f) as fh:
            per = collections.defaultdict(float)
            for r in csv.DictReader(fh):
                if r["Counter_Name"] == counter:
                    per[(f, r["Dispatch_Id"])] += float(r["Counter_Value"])
            if per:
                best = max(best, max(per.values()))
    return best


def traffic(args):
    fetch = per_kernel([args.fetch])
    write = per_kernel([args.write])
    cf, cw = copy_kib(args.copy_fetch, "FETCH_SIZE"), copy_kib(args.copy_write, "WRITE_SIZE")
    ff, wf = COPY_BYTES / (cf * 1024.0), COPY_BYTES / (cw * 1024.0)
    kern = {}
    for k in sorted(set(fetch) & set(write)):
        fb = fetch[k]["FETCH_SIZE"] * 1024.0 * ff
        wb = write[k]["WRITE_SIZE"] * 1024.0 * wf
        kern[k] = {"fetch_bytes": int(fb), "write_bytes": int(wb), "traffic_bytes": int(fb + wb),
                   "FETCH_SIZE_KiB": round(fetch[k]["FETCH_SIZE"], 1),
                   "WRITE_SIZE_KiB": round(write[k]["WRITE_SIZE"], 1),
                   "dispatches": fetch[k]["_dispatches"]}
    out = {"workload": args.workload,
           "calibration": {"copy_bytes_read": COPY_BYTES, "copy_FETCH_SIZE_KiB": round(cf, 1),
                           "fetch_factor": round(ff, 4), "copy_bytes_written": COPY_BYTES,
                           "copy_WRITE_SIZE_KiB": round(cw, 1), "write_factor": round(wf, 4)},
           "kernels": kern}
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def mix(args):
    res = per_kernel(args.paths)
    with open(args.out, "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    for k in sorted(res):
        print(k, {c: round(v, 4) for c, v in sorted(res[k].items())})


def stats(args):
    """Markdown table of a --kernel-trace --stats run (kernel_stats.csv)."""
    files = glob.glob(os.path.join(args.dir, "**", "*kernel_stats.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_stats.csv under {args.dir}")
    rows = []
    with open(files[0]) as fh:
        for r in csv.DictReader(fh):
            name = short_name(r["Name"])
            if any(s in name for s in ("__amd_rocclr",)) and not args.all:
                continue
            rows.append((name, int(r["Calls"]), float(r["AverageNs"]) * 1e-6,
                         float(r["MinNs"]) * 1e-6, float(r["MaxNs"]) * 1e-6,
                         float(r["Percentage"])))
    rows.sort(key=lambda x: -x[5])
    print("| kernel | calls | average ms | min ms | max ms | % of GPU time |")
    print("|---|---|---|---|---|---|")
    for n, c, a, mn, mx, pc in rows:
        print(f"| `{n}` | {c} | {a:.3f} | {mn:.3f} | {mx:.3f} | {pc:.2f} |")


def sorts(args):
    """One markdown table per gridding call in a kernel trace: the dispatches from
    one grid_keys_kernel to the next, summed per kernel."""
    files = glob.glob(os.path.join(args.dir, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_trace.csv under {args.dir}")
    with open(files[0]) as fh:
        rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Start_Timestamp"]))
    calls, cur = [], None
    for r in rows:
        name = short_name(r["Kernel_Name"])
        if name.startswith("grid_keys_kernel"):
            cur = {"steps": int(r["Grid_Size_X"]), "k": collections.OrderedDict()}
            calls.append(cur)
        elif cur is None or not any(name.startswith(p) for p in (
                "key_and_or", "tile_hist", "seg_sum", "scan64", "seg_offsets", "radix_scatter",
                "run_count", "run_starts", "run_counts", "decode")):
            continue
        ms = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
        c = cur["k"].setdefault(name, [0, 0.0])
        c[0] += 1
        c[1] += ms
    for c in calls:
        tot = sum(v[1] for v in c["k"].values())
        print(f"\nvrs_grid_uvw call, grid_keys_kernel launched over {c['steps']} threads "
              f"(steps): {tot:.3f} ms of kernels\n")
        print("| kernel | dispatches | total ms | average ms |")
        print("|---|---|---|---|")
        for n, (k, ms) in sorted(c["k"].items(), key=lambda kv: -kv[1][1]):
            print(f"| `{n}` | {k} | {ms:.3f} | {ms / k:.3f} |")


def copy(_args):
    import torch
    x = torch.ones(COPY_BYTES // 4, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied", COPY_BYTES, "bytes x 3")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic")
    for a in ("--fetch", "--write", "--copy-fetch", "--copy-write", "--out"):
        t.add_argument(a, required=True)
    t.add_argument("--workload", default="c2 (120 x 8640 visibilities x 196608 pixels)")
    m = sub.add_parser("mix")
    m.add_argument("paths", nargs="+")
    m.add_argument("--out", required=True)
    sub.add_parser("copy")
    st = sub.add_parser("stats")
    st.add_argument("dir")
    st.add_argument("--all", action="store_true")
    so = sub.add_parser("sorts")
    so.add_argument("dir")
    args = ap.parse_args()
    {"traffic": traffic, "mix": mix, "copy": copy, "stats": stats, "sorts": sorts}[args.cmd](args)


if __name__ == "__main__":
    main()
