"""Summaries of rocprofv3 output for the batched step.

    python3 -m constraintphysics_amd.profiling DIR [--first-step 241] [--steps 20]
        [--fetch-factor 2.0] [--write-factor 1.0] [--contacts N] [--prefix c3]
    python3 -m constraintphysics_amd.profiling --calibrate

--calibrate runs the plain copy kernel the counter factors come from: an
elementwise float32 product over 64 Mi elements (256 MiB read, 256 MiB
written), five times; profile it with the same counters as the bench.

DIR holds `<prefix>_kernel_trace.csv` (rocprofv3 --kernel-trace -o <prefix>) and, optionally,
`fetch_counter_collection.csv` / `write_counter_collection.csv` (rocprofv3
--pmc FETCH_SIZE / WRITE_SIZE) of the same bench command.  Steps are counted
by k_gravity dispatches (one per step); the window is steps
[first_step, first_step + steps).  The PGS solve of a substep is every PGS
dispatch between that substep's k_build_rows and k_rows_writeback.  Counter
values (KiB per dispatch) are scaled by the factors a plain copy kernel of
known size gives for the same counters.
"""
import argparse
import collections
import csv
import os

PGS_KERNELS = ("k_pgs_colored_lds", "k_pgs_colored(", "k_pgs_colored32", "k_pgs_ordered", "k_pgs_islands")


def _rows(path):
    with open(path) as f:
        return sorted(csv.DictReader(f), key=lambda r: int(r["Dispatch_Id"]))


def _is_pgs(name):
    return any(p in name for p in PGS_KERNELS)


def pgs_groups(rows, value=None):
    """{step: [per substep: the sum of value(row) over the PGS dispatches, or
    with value=None the span from the first PGS start to the last PGS end (the
    size classes run concurrently)]}."""
    step, cur = 0, None
    out = collections.defaultdict(list)
    for r in rows:
        n = r["Kernel_Name"]
        if "k_gravity" in n:
            step += 1
        if "k_build_rows" in n:
            cur = []
        elif _is_pgs(n) and cur is not None:
            cur.append(r)
        elif "k_rows_writeback" in n and cur is not None:
            if value is None:
                t0 = min(int(x["Start_Timestamp"]) for x in cur) if cur else 0
                t1 = max(int(x["End_Timestamp"]) for x in cur) if cur else 0
                out[step].append((t1 - t0) * 1e-6)
            else:
                out[step].append(sum(value(x) for x in cur))
            cur = None
    return out


def _dur_ms(r):
    return (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6


def summarize(d, first_step=241, steps=20, fetch_factor=2.0, write_factor=1.0, contacts=None, top=14, prefix="c3"):
    lines = []
    win = range(first_step, first_step + steps)
    tr = _rows(os.path.join(d, f"{prefix}_kernel_trace.csv"))
    g = pgs_groups(tr)
    vals = [v for s in win for v in g.get(s, [])]
    pgs_ms = sum(vals) / max(1, len(vals))
    lines.append(f"window steps {first_step}..{first_step + steps - 1}: {len(vals)} PGS solves, avg {pgs_ms:.4f} ms per solve "
                 "(first PGS dispatch start to last end)")
    per = collections.defaultdict(list)
    step = 0
    for r in tr:
        if "k_gravity" in r["Kernel_Name"]:
            step += 1
        if step in win:
            per[r["Kernel_Name"].split("(")[0]].append(_dur_ms(r))
    tot = sum(sum(v) for v in per.values())
    for k, v in sorted(per.items(), key=lambda x: -sum(x[1]))[:top]:
        lines.append("%-48s calls %5d  avg %9.4f ms  per step %8.3f ms  %5.1f%%" % (k[-48:], len(v), sum(v) / len(v), sum(v) / steps,
                                                                                    100 * sum(v) / tot))
    lines.append("sum of kernel time per step: %.3f ms" % (tot / steps))
    traffic = 0.0
    for c, nm, f in (("FETCH_SIZE", "fetch", fetch_factor), ("WRITE_SIZE", "write", write_factor)):
        p = os.path.join(d, f"{nm}_counter_collection.csv")
        if not os.path.exists(p):
            return lines, pgs_ms, None
        g = pgs_groups(_rows(p), lambda r: float(r["Counter_Value"]))
        v = [x for s in win for x in g.get(s, [])]
        kib = sum(v) / max(1, len(v))
        b = f * kib * 1024
        traffic += b
        lines.append(f"{c}: {len(v)} PGS solves, avg {kib:.1f} KiB x {f} = {b / 1e9:.4f} GB per solve")
    lines.append(f"traffic_bytes_per_launch: {traffic:.6e}")
    if contacts:
        lines.append(f"per contact point: {traffic / contacts:.1f} B")
    return lines, pgs_ms, traffic


def kernel_counters(d, name, first_step=241, steps=20, factor=1.0):
    """{kernel: (dispatches, mean counter value x factor)} over the window of a
    counter run (`name`_counter_collection.csv), steps counted by k_gravity."""
    p = os.path.join(d, f"{name}_counter_collection.csv")
    if not os.path.exists(p):
        return {}
    win = range(first_step, first_step + steps)
    acc = collections.defaultdict(list)
    step = 0
    seen = set()
    for r in _rows(p):
        k = r["Kernel_Name"].split("(")[0]
        did = r["Dispatch_Id"]
        if "k_gravity" in k and did not in seen:
            step += 1
        seen.add(did)
        if step in win:
            acc[(k, r["Counter_Name"])].append(float(r["Counter_Value"]) * factor)
    return {key: (len(v), sum(v) / len(v)) for key, v in acc.items()}


def calibrate(reps=5, n=64 * 1024 * 1024):
    """The copy kernel of known size: y = 2 x over n float32 values (4n bytes
    read, 4n bytes written per dispatch)."""
    import torch
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    print(f"copy kernel: {reps} dispatches of {4 * n / 2**20:.0f} MiB read + {4 * n / 2**20:.0f} MiB written")


def main():
    import sys
    if "--calibrate" in sys.argv:
        calibrate()
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--first-step", type=int, default=241)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--fetch-factor", type=float, default=2.0)
    ap.add_argument("--write-factor", type=float, default=1.0)
    ap.add_argument("--contacts", type=float, default=None)
    ap.add_argument("--prefix", default="c3", help="rocprofv3 -o name of the kernel trace")
    ap.add_argument("--counters", nargs="*", default=[],
                    help="name[:factor] of further counter runs (name_counter_collection.csv) to summarise per kernel")
    a = ap.parse_args()
    lines, _, _ = summarize(a.dir, a.first_step, a.steps, a.fetch_factor, a.write_factor, a.contacts, prefix=a.prefix)
    for spec in a.counters:
        name, _, f = spec.partition(":")
        for (k, c), (n, v) in sorted(kernel_counters(a.dir, name, a.first_step, a.steps, float(f or 1.0)).items()):
            lines.append("%-8s %-40s %-16s dispatches %5d  mean %14.3f" % (name, k[-40:], c, n, v))
    print("\n".join(lines))


if __name__ == "__main__":
    main()
