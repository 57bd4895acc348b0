"""Summaries of rocprofv3 runs of the scene-flow path (kernel times, HBM bytes).

Reads the rocpd SQLite databases rocprofv3 writes (`<dir>/run_results.db`)
and produces the short text summaries kept with the project:

  python rgbd-flow_amd/profiling.py stats <db> <out.md> --title T --command C [--detail K ...]
      per-kernel dispatch counts and durations of a --kernel-trace run;
  python rgbd-flow_amd/profiling.py copy --mib M
      one device-to-device copy of M MiB (torch), to be profiled with
      --pmc FETCH_SIZE and --pmc WRITE_SIZE: the counters' reading for a copy
      of known size gives their bytes-per-unit factor;
  python rgbd-flow_amd/profiling.py traffic --fetch <db> --write <db> --copy-fetch <db>
      --copy-write <db> --copy-mib M --kernel K --batch B --variant V --out <json> --command C
      HBM bytes per launch of the kernels whose name contains K, each counter
      scaled by the copy calibration (bench.py reads the JSON as roofline.traffic);
  python rgbd-flow_amd/profiling.py counters <db> [<db> ...] --kernel K --out <json> --command C [--largest N]
      per-dispatch averages of the counters of --pmc runs and the derived
      utilisations (VALU busy, LDS bank conflicts, waves per SIMD); with
      --largest, over each counter's N largest dispatches.
"""
from __future__ import annotations

import argparse
import json
import sqlite3
import sys


def _short(name: str) -> str:
    name = name.replace("rgbdflow::(anonymous namespace)::", "").replace("(anonymous namespace)::", "")
    return name.split("(")[0]


def kernel_stats(db_path):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, count(*), avg(end-start), min(end-start), max(end-start), sum(end-start) "
                      "from kernels group by name order by sum(end-start) desc").fetchall()
    return [(_short(r[0]), r[1], r[2] / 1e3, r[3] / 1e3, r[4] / 1e3, r[5] / 1e6) for r in rows]


def write_stats(db_path, out_md, title, command, detail=()):
    rows = kernel_stats(db_path)
    total = sum(r[5] for r in rows)
    L = [f"# {title}", "", "Command (repo root on the GPU host):", "", f"`{command}`", "",
         "Durations from the rocprofv3 kernel trace, all dispatches of the profiled process.", "",
         "| kernel | calls | avg us | min us | max us | total ms | share |", "|---|---|---|---|---|---|---|"]
    for r in rows:
        L.append(f"| {r[0]} | {r[1]} | {r[2]:.2f} | {r[3]:.2f} | {r[4]:.2f} | {r[5]:.2f} | {100 * r[5] / total:.1f}% |")
    for k in detail:
        sel = [r for r in rows if k in r[0]]
        n = sum(r[1] for r in sel)
        if n:
            tot = sum(r[5] for r in sel)
            L += ["", f"All kernels matching `{k}`: {n} dispatches, {tot:.2f} ms, average {tot * 1e3 / n:.2f} us."]
    with open(out_md, "w") as f:
        f.write("\n".join(L) + "\n")
    print("\n".join(L))


def counter_per_dispatch(db_path, counter, kernel_like):
    """Per-dispatch value of a counter (summed over the hardware blocks that report it)."""
    db = sqlite3.connect(db_path)
    rows = db.execute("select dispatch_id, sum(value) from counters_collection where counter_name = ? and "
                      "kernel_name like ? group by dispatch_id", (counter, f"%{kernel_like}%")).fetchall()
    return [r[1] for r in rows]


def counters(a):
    """Average per-dispatch value of every counter in the given --pmc runs for
    the kernels whose name contains a.kernel (with a.largest: over the
    a.largest largest values of each counter, the dispatches of the largest
    level), and the derived utilisations (CU count from a.cus; each SIMD
    issues one VALU instruction of a wave in 4 cycles)."""
    vals = {}
    n = {}
    for db_path in a.dbs:
        db = sqlite3.connect(db_path)
        names = [r[0] for r in db.execute("select distinct counter_name from counters_collection")]
        for c in names:
            v = counter_per_dispatch(db_path, c, a.kernel)
            if a.largest:
                v = sorted(v, reverse=True)[:a.largest]
            if v:
                vals[c] = sum(v) / len(v)
                n[c] = len(v)
    cus = a.cus
    d = {}
    # GRBM_GUI_ACTIVE is summed over the XCDs; the utilisations use one XCD's cycles
    g = vals.get("GRBM_GUI_ACTIVE") / a.xcds if vals.get("GRBM_GUI_ACTIVE") else None
    if g:
        if "SQ_ACTIVE_INST_VALU" in vals:
            d["valu_busy_pct"] = 100 * vals["SQ_ACTIVE_INST_VALU"] * 4 / (4 * cus) / g
        if "SQ_LDS_BANK_CONFLICT" in vals:
            d["lds_bank_conflict_pct_of_cycles"] = 100 * vals["SQ_LDS_BANK_CONFLICT"] / g / cus
        if "SQ_ACTIVE_INST_LDS" in vals:
            d["lds_inst_busy_pct"] = 100 * vals["SQ_ACTIVE_INST_LDS"] * 4 / (4 * cus) / g
        if "SQ_ACTIVE_INST_VMEM" in vals:
            d["vmem_inst_busy_pct"] = 100 * vals["SQ_ACTIVE_INST_VMEM"] * 4 / (4 * cus) / g
        if "TA_ADDR_STALLED_BY_TC_CYCLES" in vals and "TA_TA_BUSY" in vals and vals["TA_TA_BUSY"]:
            d["ta_addr_stalled_by_tc_pct_of_busy"] = 100 * vals["TA_ADDR_STALLED_BY_TC_CYCLES"] / vals["TA_TA_BUSY"]
    if "SQ_WAVE_CYCLES" in vals and "SQ_BUSY_CYCLES" in vals and vals["SQ_BUSY_CYCLES"]:
        d["mean_waves_per_sq_while_busy"] = vals["SQ_WAVE_CYCLES"] / vals["SQ_BUSY_CYCLES"]
    if "SQ_WAIT_INST_LDS" in vals and "SQ_WAVE_CYCLES" in vals and vals["SQ_WAVE_CYCLES"]:
        d["wave_cycles_waiting_lds_pct"] = 100 * vals["SQ_WAIT_INST_LDS"] / vals["SQ_WAVE_CYCLES"]
    if "SQ_WAIT_ANY" in vals and "SQ_WAVE_CYCLES" in vals and vals["SQ_WAVE_CYCLES"]:
        d["wave_cycles_waiting_any_pct"] = 100 * vals["SQ_WAIT_ANY"] / vals["SQ_WAVE_CYCLES"]
    out = {"kernel": a.kernel, "dispatches": max(n.values()) if n else 0, "largest": a.largest, "cus": cus,
           "per_dispatch": {k: round(v, 1) for k, v in sorted(vals.items())},
           "derived": {k: round(v, 2) for k, v in d.items()}, "command": a.command}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def traffic(a):
    fetch = counter_per_dispatch(a.fetch, "FETCH_SIZE", a.kernel)
    write = counter_per_dispatch(a.write, "WRITE_SIZE", a.kernel)
    copy_bytes = a.copy_mib * 2 ** 20
    cf = counter_per_dispatch(a.copy_fetch, "FETCH_SIZE", "")
    cw = counter_per_dispatch(a.copy_write, "WRITE_SIZE", "")
    # the copy is the largest dispatch of its run
    f_fetch = copy_bytes / max(cf)   # bytes per counter unit
    f_write = copy_bytes / max(cw)
    fb = sum(fetch) / len(fetch) * f_fetch
    wb = sum(write) / len(write) * f_write
    out = {"batch": a.batch, "variant": a.variant, "kernel": a.kernel, "launches": len(fetch),
           "fetch_bytes_per_launch": round(fb), "write_bytes_per_launch": round(wb),
           "bytes_per_launch": round(fb + wb), "copy_mib": a.copy_mib,
           "copy_factor_fetch": f_fetch, "copy_factor_write": f_write, "command": a.command}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def copy(mib):
    import torch
    n = mib * 2 ** 20 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("db")
    s.add_argument("out")
    s.add_argument("--title", required=True)
    s.add_argument("--command", required=True)
    s.add_argument("--detail", nargs="*", default=[])
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    t = sub.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write", "kernel", "variant", "out", "command"):
        t.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    t.add_argument("--copy-mib", dest="copy_mib", type=int, required=True)
    t.add_argument("--batch", type=int, required=True)
    k = sub.add_parser("counters")
    k.add_argument("dbs", nargs="+")
    k.add_argument("--kernel", required=True)
    k.add_argument("--cus", type=int, default=256)
    k.add_argument("--xcds", type=int, default=8)
    k.add_argument("--largest", type=int, default=0, help="average the largest N values of each counter")
    k.add_argument("--out", required=True)
    k.add_argument("--command", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        write_stats(a.db, a.out, a.title, a.command, a.detail)
    elif a.cmd == "counters":
        counters(a)
    elif a.cmd == "copy":
        copy(a.mib)
    else:
        traffic(a)


if __name__ == "__main__":
    main(sys.argv[1:])
