"""Text summaries of rocprofv3 databases (rocpd SQLite, ``*_results.db``).

    python trixicuda.jl_amd/profiling.py kernels DB [--match S] [--command CMD]
    python trixicuda.jl_amd/profiling.py traffic --match S --fetch DB --write DB \
        --copy-fetch DB --copy-write DB --copy-mib MIB --label L --algorithmic-bytes B [--command CMD]
    python trixicuda.jl_amd/profiling.py flops --match S --f64 DB [--f32 DB] --dofs N --ms T [--label L]
    python trixicuda.jl_amd/profiling.py valu --match S --db DB --dofs N --ms T [--label L]

``kernels`` lists per-kernel dispatch counts and durations (the numbers of
``rocprofv3 --kernel-trace --stats``) with the launch resources.  ``traffic``
turns FETCH_SIZE / WRITE_SIZE (KiB per dispatch, one counter per run) into
HBM bytes per launch, each scaled by what the same counter reads for a plain
device copy of known size (the largest ``__amd_rocclr_copyBuffer`` dispatch of
a ``bench.py --copy-calibration-mib MIB`` run), and prints a
``traffic_bytes <label> <bytes>`` line that bench.py reads.  ``flops`` turns
SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 (and _F32) wave-instruction counts into
executed flop per dispatch: 64 lanes x (ADD + MUL + TRANS + 2 FMA).  ``valu``
prints instruction counts (SQ_INSTS_VALU, _CVT, _INT32, SALU, LDS, LDS bank
conflicts, VMEM) per dispatch and per DOF, and the fraction of the SIMDs'
VALU issue cycles the kernel's VALU instructions take; with ``--waits``, the
wave-cycle breakdown (waiting at s_waitcnt or a barrier, waiting to issue,
issuing VALU / LDS) of a second run's SQ_WAVE_CYCLES group.
"""
import argparse
import collections
import sqlite3
import statistics
import sys


def kernel_rows(db):
    c = sqlite3.connect(db)
    rows = c.execute("select name, duration, lds_size, vgpr_count, scratch_size, workgroup_x, grid_x "
                     "from kernels").fetchall()
    by = collections.defaultdict(list)
    meta = {}
    for name, dur, lds, vgpr, scratch, wg, grid in rows:
        by[name].append(dur)
        meta[name] = (lds, vgpr, scratch, wg, grid)
    out = []
    for name, ds in by.items():
        lds, vgpr, scratch, wg, grid = meta[name]
        out.append({"name": name, "calls": len(ds), "avg_us": statistics.fmean(ds) / 1e3, "min_us": min(ds) / 1e3,
                    "max_us": max(ds) / 1e3, "total_us": sum(ds) / 1e3, "lds": lds, "vgpr": vgpr,
                    "scratch": scratch, "workgroup": wg, "grid": grid})
    out.sort(key=lambda r: -r["total_us"])
    return out


def counter_per_dispatch(db, counter):
    """{kernel name: [value summed over the counter's instances, per dispatch]}"""
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                     "where counter_name = ? group by dispatch_id, kernel_name", (counter,)).fetchall()
    by = collections.defaultdict(list)
    for _, name, v in rows:
        by[name].append(v)
    return by


def _pick(by, match):
    names = [n for n in by if match in n]
    if len(names) != 1:
        raise SystemExit(f"'{match}' matches {len(names)} kernels: {names}")
    return names[0], by[names[0]]


def _copy_factor(db, counter, nbytes):
    by = counter_per_dispatch(db, counter)
    vals = by.get("__amd_rocclr_copyBuffer")
    if not vals:
        raise SystemExit(f"no copy dispatches in {db}")
    return nbytes / (max(vals) * 1024.0), max(vals)


def cmd_kernels(a):
    lines = []
    if a.command:
        lines.append(f"command: {a.command}")
    lines.append(f"{'calls':>6} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'total_us':>11}  vgpr  lds  scratch  wg  grid  kernel")
    for r in kernel_rows(a.db):
        if a.match and a.match not in r["name"]:
            continue
        lines.append(f"{r['calls']:6d} {r['avg_us']:10.2f} {r['min_us']:10.2f} {r['max_us']:10.2f} {r['total_us']:11.1f}"
                     f"  {r['vgpr']:4d} {r['lds']:5d} {r['scratch']:4d} {r['workgroup']:4d} {r['grid']:8d}  {r['name']}")
    print("\n".join(lines))


def cmd_traffic(a):
    nbytes = a.copy_mib * (1 << 20)
    ff, fcopy = _copy_factor(a.copy_fetch, "FETCH_SIZE", nbytes)
    wf, wcopy = _copy_factor(a.copy_write, "WRITE_SIZE", nbytes)
    name, fetch = _pick(counter_per_dispatch(a.fetch, "FETCH_SIZE"), a.match)
    name_w, write = _pick(counter_per_dispatch(a.write, "WRITE_SIZE"), a.match)
    assert name == name_w, (name, name_w)
    f_kib, w_kib = statistics.fmean(fetch), statistics.fmean(write)
    fb, wb = f_kib * 1024 * ff, w_kib * 1024 * wf
    if a.command:
        print(f"command: {a.command}")
    print(f"kernel: {name}")
    print(f"dispatches: FETCH_SIZE run {len(fetch)}, WRITE_SIZE run {len(write)}")
    print(f"copy calibration ({a.copy_mib} MiB read + written): FETCH_SIZE {fcopy:.0f} KiB -> factor {ff:.4f}; "
          f"WRITE_SIZE {wcopy:.0f} KiB -> factor {wf:.4f}")
    print(f"FETCH_SIZE {f_kib:.1f} KiB/dispatch -> {fb / 1e6:.2f} MB; WRITE_SIZE {w_kib:.1f} KiB/dispatch -> "
          f"{wb / 1e6:.2f} MB")
    if a.algorithmic_bytes:
        print(f"algorithmic bytes {a.algorithmic_bytes / 1e6:.2f} MB; measured / algorithmic "
              f"{(fb + wb) / a.algorithmic_bytes:.3f}")
    print(f"traffic_bytes {a.label} {fb + wb:.0f}")


def cmd_flops(a):
    def one(db, prec):
        r = {}
        name = None
        for ctr in ("ADD", "MUL", "FMA", "TRANS"):
            name, vals = _pick(counter_per_dispatch(db, f"SQ_INSTS_VALU_{ctr}_{prec}"), a.match)
            r[f"{ctr}_{prec}"] = statistics.fmean(vals)
        flop = 64 * (r[f"ADD_{prec}"] + r[f"MUL_{prec}"] + r[f"TRANS_{prec}"] + 2 * r[f"FMA_{prec}"])
        return name, r, flop
    name, r64, f64 = one(a.f64, "F64")
    r32, f32 = {}, 0.0
    if a.f32:
        _, r32, f32 = one(a.f32, "F32")
    print(f"## {a.label}: {name}")
    print("wave instructions per dispatch:", ", ".join(f"{k} {v:.0f}" for k, v in {**r64, **r32}.items()))
    print(f"executed FP64 flop {f64:.4e} ({f64 / a.dofs:.0f}/DOF), FP32 flop {f32:.4e} ({f32 / a.dofs:.0f}/DOF) "
          f"per dispatch")
    print(f"at {a.ms * 1e3:.1f} us per launch: FP64 {f64 / a.ms / 1e9:.1f} TFLOP/s of 78.6, "
          f"FP32 {f32 / a.ms / 1e9:.1f} TFLOP/s of 157.3")


def read_traffic(path):
    """{label: bytes} from the ``traffic_bytes`` lines of a summary file."""
    out = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 3 and p[0] == "traffic_bytes":
                out[p[1]] = float(p[2])
    return out


# MI355X: 256 CUs x 4 SIMDs, 2.4 GHz; a VALU instruction of a 64-lane wave
# takes 4 cycles on a 16-lane SIMD (FP64, and FP32 without packing)
SIMDS, CLOCK_HZ, VALU_CYCLES = 1024, 2.4e9, 4
ISSUE_COUNTERS = ("SQ_INSTS_VALU", "SQ_INSTS_VALU_CVT", "SQ_INSTS_VALU_INT32", "SQ_INSTS_SALU", "SQ_INSTS_LDS",
                  "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR")


WAIT_COUNTERS = ("SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_WAIT_INST_LDS", "SQ_ACTIVE_INST_VALU",
                 "SQ_ACTIVE_INST_LDS", "SQ_ACTIVE_INST_VMEM")


def cmd_valu(a):
    """Instruction counts per dispatch (wave instructions, summed over the
    counter instances) and per DOF (lane instructions), and the VALU issue
    fraction: SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x 2.4 GHz x kernel time)."""
    name = None
    print(f"## {a.label}")
    for ctr in ISSUE_COUNTERS:
        by = counter_per_dispatch(a.db, ctr)
        if not any(a.match in n for n in by):
            continue
        name, vals = _pick(by, a.match)
        v = statistics.fmean(vals)
        print(f"{ctr:24s} {v:14.6g} per dispatch  {v * 64 / a.dofs:8.1f} per DOF")
        if ctr == "SQ_INSTS_VALU":
            frac = v * VALU_CYCLES / (SIMDS * CLOCK_HZ * a.ms * 1e-3)
            print(f"VALU issue at {a.ms * 1e3:.1f} us per launch: {frac:.3f} of the 1024 SIMDs' VALU cycles")
    print(f"kernel: {name}")
    if a.waits:
        vals = {}
        for ctr in WAIT_COUNTERS:
            by = counter_per_dispatch(a.waits, ctr)
            wname, vs = _pick(by, a.match)
            vals[ctr] = statistics.fmean(vs)
        wc = vals["SQ_WAVE_CYCLES"]
        print("wave-cycle breakdown (fractions of SQ_WAVE_CYCLES):")
        for ctr in WAIT_COUNTERS[1:]:
            print(f"  {ctr:22s} {vals[ctr] / wc:6.3f}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--match", default="")
    k.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    for n in ("--fetch", "--write", "--copy-fetch", "--copy-write", "--match", "--label"):
        t.add_argument(n, required=True)
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--algorithmic-bytes", type=float, default=0.0)
    t.add_argument("--command", default="")
    f = sub.add_parser("flops")
    f.add_argument("--match", required=True)
    f.add_argument("--f64", required=True)
    f.add_argument("--f32", default="")
    f.add_argument("--dofs", type=float, required=True)
    f.add_argument("--ms", type=float, required=True)
    f.add_argument("--label", default="")
    v = sub.add_parser("valu")
    v.add_argument("--match", required=True)
    v.add_argument("--db", required=True)
    v.add_argument("--dofs", type=float, required=True)
    v.add_argument("--ms", type=float, required=True)
    v.add_argument("--waits", default="")
    v.add_argument("--label", default="")
    a = ap.parse_args(argv)
    {"kernels": cmd_kernels, "traffic": cmd_traffic, "flops": cmd_flops, "valu": cmd_valu}[a.cmd](a)


if __name__ == "__main__":
    main(sys.argv[1:])
