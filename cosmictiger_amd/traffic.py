"""HBM traffic of the P2P kernels from two rocprofv3 counter runs.

    rocprofv3 --pmc FETCH_SIZE -f csv -d DIR_F -o run -- python3 bench.py ... --copy-calibration-mib M
    rocprofv3 --pmc WRITE_SIZE -f csv -d DIR_W -o run -- python3 bench.py ... --copy-calibration-mib M
    python3 -m cosmictiger_amd.traffic DIR_F DIR_W --copy-mib M --n N --workload clustered --bucket B --out FILE.json

FETCH_SIZE and WRITE_SIZE are the derived TCC counters (KiB per dispatch).
The calibration dispatch is the HIP runtime's copy kernel of a device-to-device
hipMemcpy of M MiB (the largest __amd_rocclr_copyBuffer dispatch of the run),
which reads M MiB and writes M MiB; each counter is scaled by
(M MiB) / (what it reads for that dispatch).  The P2P figure is the sum over
the k_pp dispatches (one per leaf-size class) of one kick, averaged over the
kicks of the run (a kick has one k_kick dispatch).
"""
import argparse
import csv
import glob
import json
import os
import sys


def _rows(path):
    files = [path] if os.path.isfile(path) else sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"),
                                                               recursive=True))
    if not files:
        raise FileNotFoundError("no counter_collection.csv under %s" % path)
    out = []
    for f in files:
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    return out


def per_dispatch(path, counter):
    """[(dispatch id, kernel name, value)] for one counter, in dispatch order."""
    vals = {}
    for r in _rows(path):
        if r.get("Counter_Name") != counter:
            continue
        d = int(r["Dispatch_Id"])
        name, v = vals.get(d, (r["Kernel_Name"], 0.0))
        vals[d] = (name, v + float(r["Counter_Value"]))
    return [(d, vals[d][0], vals[d][1]) for d in sorted(vals)]


def summarize(fetch_dir, write_dir, copy_mib):
    f = per_dispatch(fetch_dir, "FETCH_SIZE")
    w = per_dispatch(write_dir, "WRITE_SIZE")
    res = {}
    for key, rows in (("fetch", f), ("write", w)):
        cal = [v for _, n, v in rows if "copyBuffer" in n]
        if not cal:
            raise RuntimeError("no calibration (copyBuffer) dispatch in the %s run" % key)
        kib_expected = copy_mib * 1024.0
        scale = kib_expected / max(cal)
        pp = sum(v for _, n, v in rows if "k_pp<" in n)
        kicks = sum(1 for _, n, _ in rows if "k_kick<" in n)
        if kicks == 0:
            raise RuntimeError("no k_kick dispatch in the %s run" % key)
        res[key] = {"calibration_kib_read": max(cal), "calibration_kib_expected": kib_expected, "scale": scale,
                    "kicks": kicks, "pp_kib_raw_per_kick": pp / kicks, "pp_bytes_per_kick": pp / kicks * 1024.0 * scale}
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch_dir")
    ap.add_argument("write_dir")
    ap.add_argument("--copy-mib", type=float, required=True)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--workload", default="clustered")
    ap.add_argument("--bucket", type=int, required=True)
    ap.add_argument("--command", default="")
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    res = summarize(a.fetch_dir, a.write_dir, a.copy_mib)
    total = res["fetch"]["pp_bytes_per_kick"] + res["write"]["pp_bytes_per_kick"]
    out = {"n": a.n, "workload": a.workload, "bucket": a.bucket, "kernel": "k_pp (all leaf-size classes of one kick)",
           "bytes_per_launch": total, "fetch": res["fetch"], "write": res["write"], "command": a.command}
    if a.out:
        # the file holds one entry per (n, workload, bucket): this one replaces
        # an earlier measurement of the same case
        entries = []
        if os.path.exists(a.out):
            with open(a.out) as fh:
                old = json.load(fh)
            entries = old.get("entries", [old] if "bytes_per_launch" in old else [])
        entries = [e for e in entries if (e.get("n"), e.get("workload"), e.get("bucket")) != (a.n, a.workload, a.bucket)]
        entries.append(out)
        entries.sort(key=lambda e: (e.get("workload", ""), e.get("n", 0), e.get("bucket", 0)))
        with open(a.out, "w") as fh:
            fh.write(json.dumps({"entries": entries}, indent=1) + "\n")
    print(json.dumps(out, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
