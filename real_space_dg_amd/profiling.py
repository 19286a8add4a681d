"""Summaries of rocprofv3 output for the solver's kernels.

    python -m real_space_dg_amd.profiling stats  DIR  --command "CMD"
        kernel-trace --stats run -> markdown table (per kernel: calls, average,
        total, share), grouped by solver role and precision.
    python -m real_space_dg_amd.profiling traffic FETCH_DIR WRITE_DIR --command "CMD"
        two --pmc runs (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch) ->
        JSON of measured HBM bytes per launch per solver kernel, each scaled
        by what the same counter reads for the plain device copy of
        ``--copy-bytes`` bytes that the profiled program ran first.

Kernel names are the demangled template names of the library; ``role``
maps them to the names the library's own HIP-event profiler uses
(transport, theta_fft, theta_gemm1, theta_combine, theta_gemm2,
line_aggregate, line_carry, line_apply, arnoldi_dot, arnoldi_update) plus the precision of the
instantiation ('mixed' for float/bf16, 'fp64' for double).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import sys
from collections import defaultdict

_ROLES = [
    (r"transport_kernel<", "transport"),
    (r"combine_kernel<", "theta_combine"),
    (r"theta_fft_kernel<", "theta_fft"),
    (r"phi_fft_kernel", "phi_fft"),
    (r"gemm_(?:bf16x2|f64)_kernel<[^>]*?\bfalse\b", "theta_gemm1"),
    (r"gemm_(?:bf16x2|f64)_kernel<[^>]*?\btrue\b", "theta_gemm2"),
    (r"line_aggregate_kernel<", "line_aggregate"),
    (r"line_carry_kernel", "line_carry"),
    (r"line_apply_kernel<", "line_apply"),
    (r"arnoldi_dot_kernel<", "arnoldi_dot"),
    (r"arnoldi_dot_finalize_kernel", "arnoldi_dot_finalize"),
    (r"arnoldi_update_kernel<", "arnoldi_update"),
    (r"arnoldi_update_finalize_kernel", "arnoldi_update_finalize"),
    (r"arnoldi_normalize_kernel", "arnoldi_normalize"),
    (r"residual_kernel<", "residual"),
    (r"observables_kernel", "observables"),
    (r"poisson_newton_kernel", "poisson_newton"),
]
_COPY = re.compile(r"copyBuffer|copy_kernel|CopyKernel|elementwise_kernel|vectorized_elementwise", re.I)


def role(name: str):
    """(role, precision) of a library kernel, or None."""
    for pat, r in _ROLES:
        if re.search(pat, name):
            if "gemm_bf16x2" in name:
                prec = "mixed"
            elif "gemm_f64" in name:
                prec = "fp64"
            elif re.search(r"\bfloat\b|float2", name):
                prec = "mixed"
            elif re.search(r"\bdouble\b|double2", name):
                prec = "fp64"
            else:
                prec = "any"
            return r, prec
    return None


def _find(root: str, suffix: str):
    hits = sorted(glob.glob(os.path.join(root, "**", "*" + suffix), recursive=True))
    if not hits:
        raise SystemExit(f"no *{suffix} under {root}")
    return hits


def stats(root: str, command: str) -> str:
    rows = []
    for f in _find(root, "kernel_stats.csv"):
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    agg = defaultdict(lambda: [0, 0.0])
    other = [0, 0.0]
    for r in rows:
        rr = role(r["Name"])
        calls, tot = int(r["Calls"]), float(r["TotalDurationNs"])
        if rr is None:
            other[0] += calls
            other[1] += tot
            continue
        agg[rr][0] += calls
        agg[rr][1] += tot
    out = ["# rocprofv3 kernel-trace summary", "", "Command:", "", "```", command, "```", "",
           "| kernel (role) | precision | calls | avg us | total ms | share |",
           "|---|---|---:|---:|---:|---:|"]
    for (r, p), (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| {r} | {p} | {c} | {t / c / 1e3:.2f} | {t / 1e6:.3f} | {t / total:.3f} |")
    if other[0]:
        out.append(f"| (other) | - | {other[0]} | {other[1] / other[0] / 1e3:.2f} | {other[1] / 1e6:.3f} | "
                   f"{other[1] / total:.3f} |")
    out += ["", "Raw per-instantiation rows (top 25 by total time):", "",
            "| name | calls | avg us | total ms |", "|---|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:25]:
        nm = r["Name"]
        nm = nm if len(nm) < 110 else nm[:107] + "..."
        out.append(f"| `{nm}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.3f} |")
    return "\n".join(out) + "\n"


def _counter_rows(root: str, counter: str):
    """Per-dispatch (kernel name, value in KiB) of one counter."""
    vals = defaultdict(float)
    names = {}
    for f in _find(root, "counter_collection.csv"):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                key = (f, r["Dispatch_Id"])
                vals[key] += float(r["Counter_Value"])
                names[key] = r["Kernel_Name"]
    return [(names[k], v) for k, v in vals.items()]


def traffic(fetch_root: str, write_root: str, copy_bytes: float, command: str) -> dict:
    res = {"command": command, "copy_bytes": copy_bytes, "kernels": {}, "kernels_fp64": {}}
    per = {}
    for counter, root in (("FETCH_SIZE", fetch_root), ("WRITE_SIZE", write_root)):
        rows = _counter_rows(root, counter)
        copies = [v for n, v in rows if _COPY.search(n) and role(n) is None]
        if not copies:
            raise SystemExit(f"no copy-kernel dispatch found in {root} for calibration")
        # the calibration copy is the largest copy-kernel dispatch of the run
        ref_kib = max(copies)
        scale = (copy_bytes / 1024.0) / ref_kib
        res[f"{counter}_copy_kib"] = ref_kib
        res[f"{counter}_scale"] = scale
        acc = defaultdict(list)
        for n, v in rows:
            rr = role(n)
            if rr is not None:
                acc[rr].append(v * 1024.0 * scale)
        per[counter] = {k: sum(v) / len(v) for k, v in acc.items()}
    for (r, p) in sorted(set(per["FETCH_SIZE"]) | set(per["WRITE_SIZE"])):
        fb = per["FETCH_SIZE"].get((r, p))
        wb = per["WRITE_SIZE"].get((r, p))
        entry = {"fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
                 "bytes_per_launch": (fb or 0.0) + (wb or 0.0)}
        res["kernels_fp64" if p == "fp64" else "kernels"][r] = entry
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--command", required=True)
    a = ap.parse_args(argv)
    if a.what == "stats":
        sys.stdout.write(stats(a.dir, a.command))
    else:
        json.dump(traffic(a.fetch_dir, a.write_dir, a.copy_bytes, a.command), sys.stdout, indent=1)
        sys.stdout.write("\n")


if __name__ == "__main__":
    main()
