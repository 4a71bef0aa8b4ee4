"""Summaries of rocprofv3 runs (kernel trace + HBM byte counters) for profiles/.

    python -m simstack_amd.profiling --name NAME --cmd "COMMAND" \\
        --trace KT.db --fetch F.db --write W.db --copy-fetch CF.db --copy-write CW.db \\
        [--l2 L2.db] [--sq SQ.db] --kernel simk_gen --copy-bytes 1073741824 \\
        --out profiles/NAME.md [--traffic-json profiles/traffic_NAME.json] \\
        [--dispatch I] [--attempts A]

FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are scaled by what the same
counters read for the module's plain copy kernel of known size
(simstack_amd.calibrate), so traffic is in bytes. --dispatch picks one
dispatch of the kernel (in dispatch order) instead of the average of all;
--attempts is that dispatch's step attempts (from the bench line), giving
bytes per attempt.
"""
import argparse
import json
import sqlite3


def _q(db, sql):
    c = sqlite3.connect(db)
    try:
        return c.execute(sql).fetchall()
    finally:
        c.close()


def kernel_durations(db, prefix):
    rows = _q(db, "select name, duration from kernels")
    return [d for n, d in rows if n.startswith(prefix)]


def counter(db, prefix, name):
    rows = _q(db, "select kernel_name, counter_name, value, dispatch_id from counters_collection")
    per = {}
    for k, c, v, d in rows:
        if k.startswith(prefix) and c == name:
            per[d] = per.get(d, 0.0) + v
    return [per[d] for d in sorted(per)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--name", required=True)
    ap.add_argument("--cmd", required=True)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--copy-kernel", default="simk_copy_kernel",
                    help="simk_copy_kernel (16-byte lanes) or simk_copy8_kernel (8-byte lanes)")
    ap.add_argument("--l2")
    ap.add_argument("--sq")
    ap.add_argument("--kernel", default="simk_gen")
    ap.add_argument("--algo-flops-per-launch", type=float, default=0.0)
    ap.add_argument("--algo-bytes-per-launch", type=float, default=0.0)
    ap.add_argument("--out", required=True)
    ap.add_argument("--traffic-json")
    ap.add_argument("--dispatch", type=int, default=-1)
    ap.add_argument("--trace-dispatch", type=int, default=None,
                    help="the kernel trace's dispatch index when it differs from the counter runs' (warm-up launches)")
    ap.add_argument("--attempts", type=float, default=0.0)
    a = ap.parse_args()

    pick = (lambda v: [v[a.dispatch]]) if a.dispatch >= 0 else (lambda v: v)
    td = a.dispatch if a.trace_dispatch is None else a.trace_dispatch
    dur = kernel_durations(a.trace, a.kernel)
    dur = [dur[td]] if td >= 0 else dur
    avg_ns = sum(dur) / len(dur)
    if not a.fetch:   # kernel trace only: every kernel of the run
        rows = _q(a.trace, "select name, count(*), avg(duration), sum(duration) from kernels group by name "
                           "order by sum(duration) desc")
        lines = [f"# {a.name}", "", "Command:", "", "```",
                 f"rocprofv3 --kernel-trace --stats -d <dir> -o kt -- {a.cmd}", "```", "",
                 "| kernel | dispatches | average ms | total ms |", "|---|---|---|---|"]
        lines += [f"| `{n[:60]}` | {c} | {av / 1e6:.3f} | {tot / 1e6:.1f} |" for n, c, av, tot in rows[:12]]
        open(a.out, "w").write("\n".join(lines) + "\n")
        print("\n".join(lines))
        return
    cf = counter(a.copy_fetch, a.copy_kernel, "FETCH_SIZE")
    cw = counter(a.copy_write, a.copy_kernel, "WRITE_SIZE")
    f_scale = a.copy_bytes / (min(cf) * 1024.0)
    w_scale = a.copy_bytes / (min(cw) * 1024.0)
    fetch = pick(counter(a.fetch, a.kernel, "FETCH_SIZE"))
    write = pick(counter(a.write, a.kernel, "WRITE_SIZE"))
    rd = sum(fetch) / len(fetch) * 1024.0 * f_scale
    wr = sum(write) / len(write) * 1024.0 * w_scale
    lines = [f"# {a.name}", "", "Command (kernel trace; the counter passes run the same command under",
             "`rocprofv3 --pmc FETCH_SIZE`, `--pmc WRITE_SIZE`, ...):", "", "```",
             f"rocprofv3 --kernel-trace --stats -d <dir> -o kt -- {a.cmd}", "```", "",
             "Calibration: `rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE -- python3 -m simstack_amd.calibrate "
             f"--mib {a.copy_bytes >> 20} --reps 2{' --width 8' if a.copy_kernel == 'simk_copy8_kernel' else ''}` "
             f"(`{a.copy_kernel}`: a copy of {a.copy_bytes} bytes reads and writes that many).", "",
             "| quantity | value |", "|---|---|",
             f"| kernel `{a.kernel}*` dispatches | {len(dur)}"
             f"{f' (dispatch {td} of the traced run, {a.dispatch} of the counter runs)' if a.dispatch >= 0 else ''} |",
             f"| average duration per dispatch | {avg_ns / 1e6:.3f} ms |",
             f"| FETCH_SIZE scale (copy kernel) | {f_scale:.3f} |",
             f"| WRITE_SIZE scale (copy kernel) | {w_scale:.3f} |",
             f"| HBM bytes read per dispatch | {rd:.4e} |",
             f"| HBM bytes written per dispatch | {wr:.4e} |",
             f"| HBM traffic per dispatch | {rd + wr:.4e} |",
             f"| HBM bandwidth (traffic / avg duration) | {(rd + wr) / avg_ns:.1f} GB/s |", "",
             "FETCH_SIZE and WRITE_SIZE count the L2's traffic to the data fabric, so the "
             "\"HBM\" rows include what the 256 MB Infinity Cache serves.", ""]
    extra = []
    if a.algo_flops_per_launch:
        extra.append(f"| algorithmic FP64 flops per dispatch | {a.algo_flops_per_launch:.4e} |")
        extra.append(f"| achieved FP64 | {a.algo_flops_per_launch / avg_ns / 1e3:.3f} TFLOP/s |")
    if a.algo_bytes_per_launch:
        extra.append(f"| algorithmic bytes per dispatch | {a.algo_bytes_per_launch:.4e} |")
        extra.append(f"| achieved (algorithmic bytes / avg duration) | {a.algo_bytes_per_launch / avg_ns:.1f} GB/s |")
        extra.append(f"| counted traffic / algorithmic bytes | {(rd + wr) / a.algo_bytes_per_launch:.3f} |")
    if a.attempts:
        extra.append(f"| step attempts in the dispatch | {a.attempts:.4e} |")
        extra.append(f"| HBM traffic per attempt | {(rd + wr) / a.attempts:.1f} B |")
    if a.l2:
        hit = sum(counter(a.l2, a.kernel, "TCC_HIT_sum"))
        miss = sum(counter(a.l2, a.kernel, "TCC_MISS_sum"))
        extra.append(f"| L2 hit rate | {hit / max(hit + miss, 1):.3f} |")
    if a.sq:
        for c in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR", "SQ_WAVES",
                  "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_WAIT_INST_ANY"):
            v = counter(a.sq, a.kernel, c)
            if v:
                extra.append(f"| {c} (sum per dispatch) | {sum(v) / len(v):.4e} |")
    if extra:
        lines += ["| quantity | value |", "|---|---|"] + extra
    open(a.out, "w").write("\n".join(lines) + "\n")
    if a.traffic_json:
        json.dump({"kernel": a.kernel, "bytes_per_launch": rd + wr, "bytes_read": rd, "bytes_written": wr,
                   "avg_duration_ms": avg_ns / 1e6, "fetch_scale": f_scale, "write_scale": w_scale,
                   "attempts": a.attempts or None,
                   "bytes_per_attempt": (rd + wr) / a.attempts if a.attempts else None,
                   "command": a.cmd}, open(a.traffic_json, "w"), indent=1)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
