"""Roofline summaries from rocprofv3 CSV output (measurement tool, not part of the library).

    python3 speedball_amd/calib/roofline_summary.py ceilings <dir> > profiles/roofline_ceilings.md
    python3 speedball_amd/calib/roofline_summary.py kernels <dir> profiles/roofline_ceilings.md > profiles/walk_instanced_pmc.md

`ceilings <dir>` reads <dir>/plain.txt (the output of speedball_amd/calib/ceilings) and the
counter passes <dir>/cls (SQ_INSTS_VALU and its classes) and <dir>/mem (TCP/TCC requests)
over the same program, and derives:
  * per VALU instruction class, the issue rate (wave-instructions per second) from the
    stream kernel of that class; a class measured beside another one (TRANS beside ADD,
    INT32 and CVT beside xor/shift) has the companion's time subtracted at the
    companion's own rate;
  * for the instructions outside every counted class ("other": compares, selects, min/max,
    moves, bit logic) two rates: the fast one of xor/shift (kernel 9) and the slow one of
    min/max (kernel 1); the counters cannot split them, so a kernel's VALU issue time is
    bracketed by the two;
  * the VALU issue peak: the fastest stream's rate;
  * the L2->L1 read ceiling in bytes of 128-byte TCP->TCC requests per second (and the
    ratio of loaded bytes to request bytes of the L2 stream), and the L1 ceiling in tag
    accesses per second.

`kernels <dir> <ceilings summary>` reads the passes of one profiled program: kt (kernel trace),
fetch (FETCH_SIZE), write (WRITE_SIZE), cls (VALU classes), mem (TCP/TCC), mfma
(SQ_VALU_MFMA_BUSY_CYCLES, GRBM_GUI_ACTIVE, ...), and prints per kernel the average launch
time, HBM bytes (calibrated against the run's 256 MiB copy kernel), VALU issue bracket,
L2->L1 bytes against the L2 ceiling, and MFMA busy, plus the machine-readable lines that
bench.py reads (traffic, valu_mix, l2_req, mfma_busy).
"""
from __future__ import annotations

import collections
import csv
import os
import re
import sys

CLASSES = ("ADD_F32", "MUL_F32", "FMA_F32", "TRANS_F32", "INT32", "CVT", "INT64")
REQ_BYTES = 128  # one TCP->TCC read request moves a 128-byte line
SIMDS = 1024     # 256 CUs x 4 SIMDs
XCDS = 8


def kname(raw: str) -> str:
    s = raw.replace("(anonymous namespace)::", "")
    s = re.sub(r"^void ", "", s)
    depth, out = 0, []
    for ch in s:  # drop the argument list, keep template arguments
        if ch == "(" and depth == 0:
            break
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        out.append(ch)
    return "".join(out).replace(" ", "")


def counters(path: str):
    """{dispatch id: (kernel, {counter: value summed over dimensions})}"""
    d = collections.defaultdict(lambda: collections.defaultdict(float))
    names = {}
    with open(os.path.join(path, "run_counter_collection.csv")) as fh:
        for r in csv.DictReader(fh):
            di = int(r["Dispatch_Id"])
            d[di][r["Counter_Name"]] += float(r["Counter_Value"])
            names[di] = kname(r["Kernel_Name"])
    return {di: (names[di], dict(v)) for di, v in d.items()}


def per_kernel(path: str):
    """{kernel: {counter: mean per dispatch}}"""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for _, (k, v) in counters(path).items():
        for c, x in v.items():
            acc[k][c].append(x)
    return {k: {c: sum(x) / len(x) for c, x in v.items()} for k, v in acc.items()}


def trace_ms(path: str):
    """{kernel: (calls, mean ms)} from a kernel trace."""
    acc = collections.defaultdict(list)
    with open(os.path.join(path, "run_kernel_trace.csv")) as fh:
        for r in csv.DictReader(fh):
            acc[kname(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


def ceilings(d: str):
    plain = {}
    cache = {}
    with open(os.path.join(d, "plain.txt")) as fh:
        for line in fh:
            m = re.match(r"valu (\d+) (\S+) ([\d.]+) ms", line)
            if m:
                plain[int(m.group(1))] = (m.group(2), float(m.group(3)) * 1e-3)
            m = re.match(r"cache (l[12]) (\d+) bytes ([\d.]+) ms", line)
            if m:
                cache[m.group(1)] = (float(m.group(2)), float(m.group(3)) * 1e-3)
    cnt = {}
    for k, v in per_kernel(os.path.join(d, "cls")).items():
        m = re.match(r"k_stream<(\d+)>", k)
        if m:
            row = {c: v.get("SQ_INSTS_VALU_" + c, 0.0) for c in CLASSES}
            row["OTHER"] = v["SQ_INSTS_VALU"] - sum(row.values())
            row["TOTAL"] = v["SQ_INSTS_VALU"]
            cnt[int(m.group(1))] = row
    t = {k: plain[k][1] for k in plain}
    rate = {}
    rate["ADD_F32"] = cnt[6]["TOTAL"] / t[6]
    rate["MUL_F32"] = cnt[5]["TOTAL"] / t[5]
    rate["FMA_F32"] = cnt[0]["TOTAL"] / t[0]
    rate["OTHER_FAST"] = cnt[9]["TOTAL"] / t[9]
    rate["OTHER_SLOW"] = cnt[1]["TOTAL"] / t[1]
    rest = lambda k, main: (cnt[k][main], t[k] - (cnt[k]["TOTAL"] - cnt[k][main]) / (
        rate["ADD_F32"] if main == "TRANS_F32" else rate["OTHER_FAST"]))
    n, s = rest(4, "TRANS_F32")
    rate["TRANS_F32"] = n / s
    n, s = rest(2, "INT32")
    rate["INT32"] = n / s
    n, s = rest(7, "CVT")
    rate["CVT"] = n / s
    rate["INT64"] = cnt[8]["TOTAL"] / t[8]  # its 64-bit shifts counted at the same rate
    rate["PEAK"] = max(cnt[k]["TOTAL"] / t[k] for k in t)  # the fastest stream's issue rate
    mem = per_kernel(os.path.join(d, "mem"))
    l2 = next(v for k, v in mem.items() if k.startswith("k_cache_stream<false>"))
    l1 = next(v for k, v in mem.items() if k.startswith("k_cache_stream<true>"))
    l2_bytes, l2_s = cache["l2"]
    req_bytes = l2["TCP_TCC_READ_REQ_sum"] * REQ_BYTES
    return {"rate": rate, "cnt": cnt, "t": t, "names": {k: plain[k][0] for k in plain},
            "l2_ceiling": req_bytes / l2_s, "l2_load_ceiling": l2_bytes / l2_s,
            "l2_req_per_load_byte": req_bytes / l2_bytes,
            "l1_ceiling": l1["TCP_TOTAL_CACHE_ACCESSES_sum"] / cache["l1"][1], "l1_load_ceiling": cache["l1"][0] / cache["l1"][1]}


def issue_bracket(rate, row):
    base = sum(row[c] / rate[c] for c in CLASSES)
    return base + row["OTHER"] / rate["OTHER_FAST"], base + row["OTHER"] / rate["OTHER_SLOW"]


def read_ceilings(path: str):
    """The machine-readable lines of a ceilings summary: {rate name: per second} and
    {l2_ceiling: bytes/s, l1_ceiling: accesses/s}."""
    rate, other = {}, {}
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if len(p) == 3 and p[0] == "rate":
                rate[p[1].upper()] = float(p[2]) * 1e9
            if len(p) == 2 and p[0] in ("l2_ceiling", "l1_ceiling"):
                other[p[0]] = float(p[1]) * 1e9
    return rate, other


def cmd_ceilings(d: str):
    c = ceilings(d)
    r = c["rate"]
    clk = 256 * 2.4e9
    print("# VALU issue rates and the L2 read ceiling of this GPU (speedball_amd/calib/ceilings)\n")
    print("Commands (from the repository root, TMPDIR=/tmp):\n")
    print("    speedball_amd/calib/ceilings > gpurun_out/cal/plain.txt")
    print("    rocprofv3 --kernel-trace --pmc SQ_INSTS_VALU SQ_INSTS_VALU_ADD_F32 SQ_INSTS_VALU_MUL_F32 "
          "SQ_INSTS_VALU_FMA_F32 SQ_INSTS_VALU_TRANS_F32 SQ_INSTS_VALU_INT32 SQ_INSTS_VALU_CVT "
          "SQ_INSTS_VALU_INT64 -d gpurun_out/cal/cls -o run --output-format csv -- speedball_amd/calib/ceilings")
    print("    rocprofv3 --kernel-trace --pmc TCP_TCC_READ_REQ_sum TCP_TOTAL_CACHE_ACCESSES_sum TCC_HIT_sum "
          "TCC_MISS_sum GRBM_GUI_ACTIVE -d gpurun_out/cal/mem -o run --output-format csv -- "
          "speedball_amd/calib/ceilings")
    print("    python3 speedball_amd/calib/roofline_summary.py ceilings gpurun_out/cal\n")
    print("Stream kernels (4,096 blocks x 256 threads, best of 3), wave-instructions by class from the counters:\n")
    print("| kernel | best ms | wave-instr (M) | by class (M) | instr/s (G) | per CU-clock at 2.4 GHz |")
    print("|---|---:|---:|---|---:|---:|")
    for k in sorted(c["t"]):
        row = c["cnt"][k]
        cls = ", ".join(f"{x.lower()} {row[x] / 1e6:.0f}" for x in CLASSES + ("OTHER",) if row[x] > 0.01 * row["TOTAL"])
        ips = row["TOTAL"] / c["t"][k]
        print(f"| {k} {c['names'][k]} | {c['t'][k] * 1e3:.4f} | {row['TOTAL'] / 1e6:.1f} | {cls} | "
              f"{ips / 1e9:.1f} | {ips / clk:.2f} |")
    print("\nClass rates (TRANS, INT32 and CVT with the companion add or xor/shift time removed at "
          "its own rate; INT64 with its 64-bit shifts at the same rate). Instructions outside the "
          "counted classes issue between the xor/shift rate (fast) and the min/max rate (slow): a "
          "kernel's VALU issue time is bracketed by counting them at one or the other.\n")
    print(f"The VALU issue peak (`rate peak`) is the fastest stream's rate, {r['PEAK'] / 1e9:.0f} G wave-instructions/s "
          f"({r['PEAK'] / clk:.2f} per CU per 2.4 GHz clock): no instruction issues faster, so instructions / "
          "(peak x time) is a lower bound on a kernel's share of the issue capacity, whatever its mix.\n")
    print("```")
    for name in CLASSES + ("OTHER_FAST", "OTHER_SLOW", "PEAK"):
        print(f"rate {name.lower()} {r[name] / 1e9:.1f}")
    print(f"l2_ceiling {c['l2_ceiling'] / 1e9:.1f}")
    print(f"l1_ceiling {c['l1_ceiling'] / 1e9:.1f}")
    print("```\n")
    print(f"L2 read stream: {c['l2_load_ceiling'] / 1e9:.0f} GB/s of loaded bytes; TCP->TCC requests x "
          f"{REQ_BYTES} B = {c['l2_req_per_load_byte']:.3f} of the loaded bytes (the rest hit L1), so the "
          f"ceiling in request bytes is {c['l2_ceiling'] / 1e9:.0f} GB/s (`l2_ceiling`, GB/s). L1 read stream: "
          f"{c['l1_load_ceiling'] / 1e9:.0f} GB/s of loaded bytes, {c['l1_ceiling'] / 1e9:.0f} G tag accesses/s "
          "(TCP_TOTAL_CACHE_ACCESSES; `l1_ceiling`, G accesses/s).")
    # the model on its own calibration kernels
    print("\nThe bracket on the calibration kernels themselves (issue time / measured time):\n")
    print("| kernel | low | high |")
    print("|---|---:|---:|")
    for k in sorted(c["t"]):
        lo, hi = issue_bracket(r, c["cnt"][k])
        print(f"| {k} {c['names'][k]} | {lo / c['t'][k]:.2f} | {hi / c['t'][k]:.2f} |")


def cmd_kernels(d: str, cal: str):
    rate, ceil = read_ceilings(cal)
    l2_ceiling, l1_ceiling = ceil["l2_ceiling"], ceil["l1_ceiling"]
    kt = trace_ms(os.path.join(d, "kt"))
    fetch = per_kernel(os.path.join(d, "fetch"))
    write = per_kernel(os.path.join(d, "write"))
    cls = per_kernel(os.path.join(d, "cls"))
    mem = per_kernel(os.path.join(d, "mem"))
    mf = per_kernel(os.path.join(d, "mfma"))
    # the run's 256 MiB device-to-device copy: the largest copy-kernel dispatch
    copy_bytes = 256 << 20
    big = lambda path, ctr: max(v[ctr] for k, v in counters(path).values() if k.startswith("__amd_rocclr_copyBuffer"))
    f_scale = copy_bytes / (big(os.path.join(d, "fetch"), "FETCH_SIZE") * 1024)
    w_scale = copy_bytes / (big(os.path.join(d, "write"), "WRITE_SIZE") * 1024)
    rows = []
    for k, (calls, ms) in sorted(kt.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        if k not in cls or k.startswith("__amd"):
            continue
        s = ms * 1e-3
        c = cls[k]
        row = {x: c.get("SQ_INSTS_VALU_" + x, 0.0) for x in CLASSES}
        row["OTHER"] = c["SQ_INSTS_VALU"] - sum(row.values())
        lo, hi = issue_bracket(rate, row)
        rd = fetch.get(k, {}).get("FETCH_SIZE", 0.0) * 1024 * f_scale
        wr = write.get(k, {}).get("WRITE_SIZE", 0.0) * 1024 * w_scale
        req = mem.get(k, {}).get("TCP_TCC_READ_REQ_sum", 0.0) * REQ_BYTES
        l1a = mem.get(k, {}).get("TCP_TOTAL_CACHE_ACCESSES_sum", 0.0)
        m = mf.get(k, {})
        busy = None
        if m.get("GRBM_GUI_ACTIVE"):
            busy = m.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / (m["GRBM_GUI_ACTIVE"] / XCDS * SIMDS)
        rows.append((k, calls, ms, rd, wr, c["SQ_INSTS_VALU"], c["SQ_INSTS_VALU"] / s / rate["PEAK"], lo / s, hi / s,
                     req, req / s / l2_ceiling, l1a / s / l1_ceiling, busy, m.get("SQ_INSTS_SALU", 0.0), row))
    print("# Probe update, instanced workload: counters per kernel (rocprofv3 --pmc)\n")
    print("Workload: python3 bench.py --steps 3 --warmup 1 --no-cpu --no-check --no-adjoint --no-dynamic "
          "--no-batch --no-lights --no-reflect --no-room --copy-calibration (the headline field: 12,288 probes "
          "x 256 rays, instanced scene, BF16x3 blend, ordered traversal). One rocprofv3 run per counter "
          "pass, TMPDIR=/tmp, from the repository root:\n")
    for tag, ctr in (("kt", "--kernel-trace --stats"), ("fetch", "--pmc FETCH_SIZE"), ("write", "--pmc WRITE_SIZE"),
                     ("cls", "--pmc SQ_INSTS_VALU SQ_INSTS_VALU_ADD_F32 SQ_INSTS_VALU_MUL_F32 SQ_INSTS_VALU_FMA_F32 "
                             "SQ_INSTS_VALU_TRANS_F32 SQ_INSTS_VALU_INT32 SQ_INSTS_VALU_CVT SQ_INSTS_VALU_INT64"),
                     ("mem", "--pmc TCP_TCC_READ_REQ_sum TCP_TOTAL_CACHE_ACCESSES_sum TCC_HIT_sum TCC_MISS_sum "
                             "GRBM_GUI_ACTIVE"),
                     ("mfma", "--pmc SQ_VALU_MFMA_BUSY_CYCLES GRBM_GUI_ACTIVE SQ_INSTS_MFMA SQ_INSTS_LDS "
                              "SQ_INSTS_VMEM_RD SQ_INSTS_SALU SQ_WAVES")):
        print(f"    rocprofv3 {ctr} -d gpurun_out/prof/{tag} -o run --output-format csv -- python3 bench.py "
              f"<workload flags above>")
    print(f"    python3 speedball_amd/calib/roofline_summary.py kernels gpurun_out/prof {cal}\n")
    print(f"HBM bytes: FETCH_SIZE x 1024 x {f_scale:.4f} and WRITE_SIZE x 1024 x {w_scale:.4f} (the factors "
          f"that make the run's 256 MiB device-to-device copy read and write 256 MiB).\n")
    print(f"Ceilings from {cal}. VALU peak: wave-instructions / (the fastest stream's issue rate x launch "
          "time), a lower bound on the share of the issue capacity used. VALU mix low / high: the same with "
          "each class at its own measured rate and the unclassified instructions at the fast or the slow rate "
          "(an additive model that ignores overlap between classes, so not a bound: it exceeds 1 on "
          "k_resolve). L2: TCP->TCC read requests x "
          f"{REQ_BYTES} B per second over the L2 read ceiling. L1: TCP tag accesses per second over the L1 "
          "ceiling. MFMA busy: SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE per XCD x 1,024 SIMDs).\n")
    print("| kernel | calls | avg ms | HBM rd MB | HBM wr MB | VALU (M) | SALU (M) | VALU peak | mix low | mix high | "
          "L2 req MB | L2 | L1 | MFMA busy |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    for k, calls, ms, rd, wr, v, pk, lo, hi, req, l2f, l1f, busy, salu, _ in rows:
        b = f"{busy:.3f}" if busy is not None else "-"
        print(f"| `{k}` | {calls} | {ms:.4f} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | {v / 1e6:.1f} | {salu / 1e6:.1f} | "
              f"{pk:.2f} | {lo:.2f} | {hi:.2f} | {req / 1e6:.1f} | {l2f:.2f} | {l1f:.2f} | {b} |")
    print("\nMachine-readable (bench.py reads these lines): traffic = HBM read and write bytes per dispatch; "
          "valu_mix = wave-instructions per dispatch, total then " + ", ".join(c.lower() for c in CLASSES)
          + ", other; l2_req = TCP->TCC request bytes per dispatch; mfma_busy = MFMA busy fraction.\n")
    print("```")
    for k, calls, ms, rd, wr, v, pk, lo, hi, req, l2f, l1f, busy, salu, row in rows:
        print(f"traffic {k} {rd:.0f} {wr:.0f}")
        print(f"valu_mix {k} {v:.0f} " + " ".join(f"{row[c]:.0f}" for c in CLASSES + ("OTHER",)))
        print(f"l2_req {k} {req:.0f}")
        if busy is not None:
            print(f"mfma_busy {k} {busy:.4f}")
    print("```")


def cmd_stats(d: str, title: str, command: str):
    """Per-kernel time table from a --kernel-trace --stats run."""
    rows = []
    with open(os.path.join(d, "run_kernel_stats.csv")) as fh:
        for r in csv.DictReader(fh):
            rows.append((kname(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]) / 1e6,
                         float(r["AverageNs"]) / 1e6, float(r["Percentage"])))
    print(f"# {title}\n")
    print("Command (from the repository root, TMPDIR=/tmp):\n")
    print(f"    {command}")
    print(f"    python3 speedball_amd/calib/roofline_summary.py stats {d} \"<title>\" \"<the command above>\"\n")
    print("| kernel | calls | total ms | avg ms | % |")
    print("|---|---:|---:|---:|---:|")
    for k, n, tot, avg, pct in sorted(rows, key=lambda x: -x[2]):
        print(f"| `{k}` | {n} | {tot:.3f} | {avg:.4f} | {pct:.1f} |")


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "ceilings":
        cmd_ceilings(sys.argv[2])
    elif len(sys.argv) >= 5 and sys.argv[1] == "stats":
        cmd_stats(sys.argv[2], sys.argv[3], sys.argv[4])
    elif len(sys.argv) >= 4 and sys.argv[1] == "kernels":
        cmd_kernels(sys.argv[2], sys.argv[3])
    else:
        sys.exit(__doc__)
