"""Summaries of rocprofv3 databases for the transport engine.

Reads the rocpd SQLite files that ``rocprofv3 -d DIR -o run`` writes (one
kernel-trace run and one run per PMC pass) and prints the figures the
engine's roofline uses:

* the kernel's average dispatch duration (``kernels`` view);
* ``FETCH_SIZE`` and ``WRITE_SIZE`` (KiB per dispatch, ``counters_collection``
  view), each scaled by what the same counter reads for a plain device copy of
  known size run in the same process (``bench.py --copy-calibration-mb``):
  factor = copy bytes / counter bytes;
* optional SQ instruction counters, per primary photon.

Usage::

    python -m ggems_amd.profiling --kt D1 --fetch D2 --write D3 [--sq D4] \
        --copy-mb 256 --photons 16777216 --out summary.txt --json traffic.json

Without ``--fetch``/``--write`` it summarises the kernel trace alone.
"""

import argparse
import glob
import hashlib
import json
import os
import sqlite3

KERNEL = "transport_kernel"
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _db(path):
    if os.path.isdir(path):
        found = sorted(glob.glob(os.path.join(path, "**", "*.db"), recursive=True))
        if not found:
            raise FileNotFoundError("no rocpd database under %s" % path)
        path = found[0]
    return sqlite3.connect(path)


def kernel_durations(path, kernel=KERNEL):
    """(calls, average ns, min ns, max ns) of every dispatch whose name holds ``kernel``."""
    con = _db(path)
    rows = con.execute("select duration from kernels where name like ?", ("%" + kernel + "%",)).fetchall()
    d = [r[0] for r in rows]
    if not d:
        raise ValueError("kernel %s not in %s" % (kernel, path))
    return len(d), sum(d) / len(d), min(d), max(d)


def counter_values(path, counter, kernel):
    """Per-dispatch values (KiB for the byte counters) of ``counter`` for ``kernel``."""
    con = _db(path)
    rows = con.execute("select dispatch_id, grid_size, value from counters_collection "
                       "where counter_name = ? and kernel_name like ? order by dispatch_id",
                       (counter, "%" + kernel + "%")).fetchall()
    return [(r[0], r[1], r[2]) for r in rows]


def top_kernels(path, n=6):
    """(name, calls, total us, average us, percent) rows of the kernel-trace statistics (the top_kernels view
    holds microseconds)."""
    con = _db(path)
    return con.execute("select name, total_calls, total_duration, average, percentage from top_kernels "
                       "order by total_duration desc limit ?", (n,)).fetchall()


def copy_factor(path, counter, copy_bytes):
    """copy bytes / counter bytes for the largest plain copy dispatch of the run."""
    vals = counter_values(path, counter, COPY_KERNEL)
    if not vals:
        raise ValueError("no %s dispatch in %s" % (COPY_KERNEL, path))
    dispatch, grid, kib = max(vals, key=lambda v: v[1])
    return copy_bytes / (kib * 1024.0), dispatch, kib


def source_digest(repo=None):
    """SHA-256 over the engine sources, to tie a traffic measurement to the code it measured."""
    repo = repo or os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    h = hashlib.sha256()
    for rel in ("ggems_amd/csrc/engine/engine.hip", "ggems_amd/csrc/engine/transport_device.hpp"):
        with open(os.path.join(repo, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def summarize(kt, fetch=None, write=None, sq=None, copy_mb=256, photons=None, kernel=KERNEL, extra=(), steps=None,
              top=6):
    """Kernel-trace figures of ``kernel``; HBM traffic too when both PMC runs are given."""
    copy_bytes = copy_mb * (1 << 20)
    calls, avg, lo, hi = kernel_durations(kt, kernel)
    out = {
        "kernel": kernel,
        "calls": calls,
        "avg_ns": avg,
        "min_ns": lo,
        "max_ns": hi,
        "photons": photons,
        "top": top_kernels(kt, top),
    }
    if fetch and write:
        ff, fd, fkib = copy_factor(fetch, "FETCH_SIZE", copy_bytes)
        wf, wd, wkib = copy_factor(write, "WRITE_SIZE", copy_bytes)
        fetch_k = [v[2] for v in counter_values(fetch, "FETCH_SIZE", kernel)]
        write_k = [v[2] for v in counter_values(write, "WRITE_SIZE", kernel)]
        fetch_b = sum(fetch_k) / len(fetch_k) * 1024.0 * ff
        write_b = sum(write_k) / len(write_k) * 1024.0 * wf
        out.update({
            "copy_bytes": copy_bytes,
            "fetch_factor": ff,
            "write_factor": wf,
            "copy_fetch_kib": fkib,
            "copy_write_kib": wkib,
            "fetch_kib_raw": sum(fetch_k) / len(fetch_k),
            "write_kib_raw": sum(write_k) / len(write_k),
            "fetch_bytes": fetch_b,
            "write_bytes": write_b,
            "traffic_bytes": fetch_b + write_b,
        })
    if photons and "traffic_bytes" in out:
        out["traffic_bytes_per_primary"] = out["traffic_bytes"] / photons
    counters = {}
    for path in ([sq] if sq else []) + list(extra):
        con = _db(path)
        rows = con.execute("select counter_name, avg(value) from counters_collection where kernel_name like ? "
                           "group by counter_name", ("%" + kernel + "%",)).fetchall()
        counters.update({r[0]: r[1] for r in rows})
    if counters:
        out["sq"] = counters
        hit, miss = counters.get("TCC_HIT_sum"), counters.get("TCC_MISS_sum")
        if hit is not None and miss is not None and hit + miss > 0:
            out["l2_hit_rate"] = hit / (hit + miss)
    out["steps"] = steps
    return out


def format_summary(s, commands=()):
    lines = []
    lines.append("kernel: %s" % s["kernel"])
    lines.append("dispatches traced: %d   avg %.3f ms   min %.3f ms   max %.3f ms"
                 % (s["calls"], s["avg_ns"] * 1e-6, s["min_ns"] * 1e-6, s["max_ns"] * 1e-6))
    lines.append("kernel-trace statistics (all kernels of the traced run):")
    lines.append("  %-72s %6s %12s %12s %7s" % ("name", "calls", "total ms", "avg ms", "%"))
    for name, calls, total, avg, pct in s["top"]:
        lines.append("  %-72s %6d %12.3f %12.3f %7.3f" % (name[:72], calls, total * 1e-3, avg * 1e-3, pct))
    if "traffic_bytes" in s:
        lines.append("copy calibration (%d MiB device copy): FETCH_SIZE read %.1f KiB -> factor %.4f; "
                     "WRITE_SIZE read %.1f KiB -> factor %.4f"
                     % (s["copy_bytes"] >> 20, s["copy_fetch_kib"], s["fetch_factor"], s["copy_write_kib"],
                        s["write_factor"]))
        lines.append("FETCH_SIZE per dispatch: %.1f KiB raw -> %.3f GB calibrated"
                     % (s["fetch_kib_raw"], s["fetch_bytes"] * 1e-9))
        lines.append("WRITE_SIZE per dispatch: %.1f KiB raw -> %.3f GB calibrated"
                     % (s["write_kib_raw"], s["write_bytes"] * 1e-9))
        lines.append("traffic per dispatch: %.3f GB" % (s["traffic_bytes"] * 1e-9))
        if s.get("photons"):
            lines.append("primaries per dispatch: %d   traffic per primary: %.1f B"
                         % (s["photons"], s["traffic_bytes_per_primary"]))
    elif s.get("photons"):
        lines.append("primaries per dispatch: %d" % s["photons"])
    if s.get("sq"):
        lines.append("counters per dispatch:")
        for k in sorted(s["sq"]):
            v = s["sq"][k]
            extra = ""
            if s.get("photons") and k.startswith("SQ_INSTS"):
                extra = "   (%.1f per primary" % (v / s["photons"])
                if s.get("steps"):
                    extra += ", %.3f per voxel step" % (v / s["steps"])
                extra += ")"
            lines.append("  %-20s %.6g%s" % (k, v, extra))
        if "l2_hit_rate" in s:
            lines.append("L2 hit rate (TCC_HIT / (TCC_HIT + TCC_MISS)): %.3f" % s["l2_hit_rate"])
    if commands:
        lines.append("")
        lines.append("commands:")
        lines.extend("  " + c for c in commands)
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kt", required=True)
    ap.add_argument("--fetch", help="FETCH_SIZE run (with --write: HBM traffic per dispatch)")
    ap.add_argument("--write", help="WRITE_SIZE run")
    ap.add_argument("--sq")
    ap.add_argument("--extra", action="append", default=[], help="further PMC run directories")
    ap.add_argument("--steps", type=float, help="voxel steps per dispatch (per-step instruction counts)")
    ap.add_argument("--copy-mb", type=int, default=256)
    ap.add_argument("--photons", type=int)
    ap.add_argument("--kernel", default=KERNEL)
    ap.add_argument("--top", type=int, default=6, help="rows of the kernel-trace statistics to print")
    ap.add_argument("--out")
    ap.add_argument("--json")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--sampling", default="reference", help="the transport's sampling mode in the measured run")
    args = ap.parse_args()
    s = summarize(args.kt, args.fetch, args.write, args.sq, args.copy_mb, args.photons, args.kernel, args.extra,
                  args.steps, args.top)
    text = format_summary(s, args.command)
    print(text, end="")
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    if args.json:
        if "traffic_bytes" not in s:
            ap.error("--json needs --fetch and --write")
        rec = {"kernel": s["kernel"], "photons": s["photons"], "traffic_bytes": s["traffic_bytes"],
               "traffic_bytes_per_primary": s.get("traffic_bytes_per_primary"),
               "fetch_factor": s["fetch_factor"], "write_factor": s["write_factor"],
               "avg_ms": s["avg_ns"] * 1e-6, "source_sha256": source_digest(), "sampling": args.sampling}
        c = s.get("sq", {})
        if args.steps and "SQ_INSTS_VALU" in c:
            rec["valu_instr_per_step"] = c["SQ_INSTS_VALU"] / args.steps
            rec["salu_instr_per_step"] = c.get("SQ_INSTS_SALU", 0.0) / args.steps
        if "l2_hit_rate" in s:
            rec["l2_hit_rate"] = s["l2_hit_rate"]
        with open(args.json, "w") as f:
            json.dump(rec, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
