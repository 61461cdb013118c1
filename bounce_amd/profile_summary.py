"""Summaries of rocprofv3 output for the batched step (measurement support).

    python3 -m bounce_amd.profile_summary --trace-dir D1 --bench-json B --fetch-dir D2 --write-dir D3 \
        --warmup W --steps K --out-dir profiles

Reads the CSV output of
  * `rocprofv3 --kernel-trace --stats --output-format csv -d D1 -o run -- python3 bench.py ...`
    (the bench's own process: its JSON line is in B),
  * `rocprofv3 --pmc FETCH_SIZE ... -d D2 -o run -- python3 -m bounce_amd.profile_step ...` and the
    same with WRITE_SIZE into D3,
and writes
  * kernels_pyramids<worlds>.md: rocprof's per-kernel statistics over all calls, the average over
    the calls of the bench's timed region (steps W+1 .. W+K) from the kernel trace, and the HIP-event
    average the bench printed for the same steps;
  * traffic_pyramids<worlds>.md and traffic.json: HBM bytes per launch of each step kernel over the
    same steps, from FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by the factor that maps
    the same counters on k_copy (a copy of a known size) to its true bytes.
"""
import argparse
import csv
import json
import os
import re
import sys

STEP_KERNELS = ["k_update_contacts", "k_islands", "k_integrate_vel", "k_init_contacts", "k_solve_velocity",
                "k_mid", "k_solve_position", "k_finish", "k_sync", "k_find_pairs"]


def short(name):
    m = re.search(r"bx::(k_\w+)(?:<[^>]*>)?\(", name)
    if m:
        # the eight-lane sweep kernels fill the velocity- and position-sweep slots of the step
        return {"k_solve_velocity_quad": "k_solve_velocity",
                "k_solve_position_quad": "k_solve_position"}.get(m.group(1), m.group(1))
    m = re.search(r"(k_copy|k_fill)\(", name)
    return m.group(1) if m else name


def instance(name):
    """The kernel's short name with its template arguments (k_solve_velocity<16u, 256u>)."""
    m = re.search(r"bx::(k_\w+(?:<[^>]*>)?)\(", name)
    return m.group(1) if m else short(name)


def per_kernel_calls(rows, key):
    out = {}
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        out.setdefault(short(r["Kernel_Name"]), []).append(key(r))
    return out


def timed_slice(calls, name, warmup, steps):
    # k_find_pairs runs once more than the other kernels: the creation-time pair search of step 1
    first = warmup + (1 if name == "k_find_pairs" else 0)
    return calls[first:first + steps]


def read_json_line(path):
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("{"):
                return json.loads(line)
    raise ValueError(f"no JSON line in {path}")


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--trace-dir", required=True)
    p.add_argument("--bench-json", required=True)
    p.add_argument("--fetch-dir", required=True)
    p.add_argument("--write-dir", required=True)
    p.add_argument("--profile-json", required=True, help="JSON line printed by profile_step in the FETCH run")
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--steps", type=int, default=60)
    p.add_argument("--out-dir", default="profiles")
    p.add_argument("--commands", default="", help="text of the commands that produced the inputs")
    a = p.parse_args()

    bench = read_json_line(a.bench_json)
    prof = read_json_line(a.profile_json)
    worlds = bench["config"]["worlds"]
    tag = f"pyramids{worlds}"
    os.makedirs(a.out_dir, exist_ok=True)

    # --- kernel durations -------------------------------------------------------------------------
    with open(os.path.join(a.trace_dir, "run_kernel_stats.csv")) as f:
        stats = list(csv.DictReader(f))
    with open(os.path.join(a.trace_dir, "run_kernel_trace.csv")) as f:
        trace = list(csv.DictReader(f))
    dur = per_kernel_calls(trace, lambda r: (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6)
    lines = [f"# Kernel times: batched stacking, {worlds} worlds x {bench['config']['bodies'] // worlds} bodies",
             "", "Commands (from the repository root):", "", "```", a.commands.strip(), "```", "",
             "## rocprofv3 --stats (all calls of the bench process)", "",
             "| kernel | calls | average ms | share % |", "|---|---|---|---|"]
    for r in stats:
        lines.append(f"| {instance(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) * 1e-6:.4f} | "
                     f"{float(r['Percentage']):.2f} |")
    lines += ["", f"## Timed region (steps {a.warmup + 1}..{a.warmup + a.steps}): kernel trace vs the bench's "
              "HIP events", "", "| kernel | trace average ms | HIP-event average ms | ratio |", "|---|---|---|---|"]
    agree = {}
    for k in STEP_KERNELS:
        calls = timed_slice(dur.get(k, []), k, a.warmup, a.steps)
        if not calls:
            continue
        t = sum(calls) / len(calls)
        e = bench["kernel_ms"].get(k)
        agree[k] = (t, e)
        lines.append(f"| {k} | {t:.4f} | {e:.4f} | {e / t:.3f} |")
    lines += ["", f"bench: value {bench['value']:.4g} body-steps/s, {bench['ms_per_step']:.3f} ms/step, "
              f"roofline kernel {bench['roofline']['kernel']}, achieved {bench['roofline']['achieved']:.1f} GB/s "
              f"({bench['roofline']['frac'] * 100:.2f}% of {bench['roofline']['peak']:.0f} GB/s)", ""]
    with open(os.path.join(a.out_dir, f"kernels_{tag}.md"), "w") as f:
        f.write("\n".join(lines))

    # --- HBM traffic -----------------------------------------------------------------------------
    def counter(d, name):
        with open(os.path.join(d, "run_counter_collection.csv")) as f:
            rows = [r for r in csv.DictReader(f) if r["Counter_Name"] == name]
        return per_kernel_calls(rows, lambda r: float(r["Counter_Value"]))

    fetch = counter(a.fetch_dir, "FETCH_SIZE")
    write = counter(a.write_dir, "WRITE_SIZE")
    copy_bytes = prof["copy_bytes_each_way"]
    f_fetch = copy_bytes / (sum(fetch["k_copy"]) / len(fetch["k_copy"]) * 1024.0)
    f_write = copy_bytes / (sum(write["k_copy"]) / len(write["k_copy"]) * 1024.0)
    units = prof["units"]
    alg = {
        "k_solve_velocity": 428 * units["island_contacts"] + 48 * units["island_bodies"],
        "k_solve_position": 272 * units["island_contacts"] + 128 * units["island_bodies"],
        "k_update_contacts": (2 * 332 + 2 * (72 + 24 + 88)) * units["contacts"],
    }
    traffic = {"worlds": worlds, "levels": prof["levels"], "steps": f"{a.warmup + 1}..{a.warmup + a.steps}",
               "fetch_factor": f_fetch, "write_factor": f_write, "bytes_per_launch": {}}
    lines = [f"# HBM traffic per launch: batched stacking, {worlds} worlds", "",
             "Commands (from the repository root):", "", "```", a.commands.strip(), "```", "",
             f"Calibration: k_copy of {copy_bytes} B each way reads FETCH_SIZE "
             f"{copy_bytes / f_fetch / 1024:.0f} KiB and WRITE_SIZE {copy_bytes / f_write / 1024:.0f} KiB per "
             f"dispatch: factors {f_fetch:.4f} (fetch) and {f_write:.4f} (write) bytes per counted byte.", "",
             f"Steps {a.warmup + 1}..{a.warmup + a.steps}; units of work at the last step: "
             f"{units['island_contacts']:.0f} island contacts, {units['island_bodies']:.0f} island bodies, "
             f"{units['contacts']:.0f} contacts.", "",
             "| kernel | FETCH bytes | WRITE bytes | traffic bytes | algorithmic bytes | traffic / algorithmic |",
             "|---|---|---|---|---|---|"]
    for k in STEP_KERNELS:
        fc = timed_slice(fetch.get(k, []), k, a.warmup, a.steps)
        wc = timed_slice(write.get(k, []), k, a.warmup, a.steps)
        if not fc or not wc:
            continue
        fb = sum(fc) / len(fc) * 1024.0 * f_fetch
        wb = sum(wc) / len(wc) * 1024.0 * f_write
        traffic["bytes_per_launch"][k] = fb + wb
        ab = alg.get(k)
        lines.append(f"| {k} | {fb:.4g} | {wb:.4g} | {fb + wb:.4g} | "
                     f"{'' if ab is None else f'{ab:.4g}'} | {'' if ab is None else f'{(fb + wb) / ab:.2f}'} |")
    lines.append("")
    with open(os.path.join(a.out_dir, f"traffic_{tag}.md"), "w") as f:
        f.write("\n".join(lines))
    with open(os.path.join(a.out_dir, "traffic.json"), "w") as f:
        json.dump(traffic, f, indent=1)
    print(json.dumps({"agreement": agree, "traffic": traffic["bytes_per_launch"]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
