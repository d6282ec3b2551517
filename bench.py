"""Benchmark: full D3Q19 IBM shear-cell steps on one MI355X.

Workload (default): the 128^3 power-law configuration -- box 128x128x128 between walls moving at
+-0.02 (LatticeSize 130x130x131), power-law fluid n = 0.7 with K = nu*g0^(1-n), nu from tau 0.9 and
g0 = 2U/H, local tau from the strain rate clamped to [0.51, 5], BGK collision, 32 spheres of radius
6 with 192 markers each at seeded uniform-random centres (surface gap >= 2), Peskin 4-point IBM,
fp32 populations.  A step is the whole time step (dynamics, collide+stream, IBM, contacts).

Prints one JSON line: MLUPS of the fp32 step (value), the fp16-shifted-storage step, the
collide+stream kernel's HBM roofline (per-launch event timing on the simulation's stream), and a
CPU baseline (the oracle, OpenMP, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dispersion-shear_d3q19-opencl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_collide_c2_powerlaw.md")


def measured_traffic(key):
    """HBM bytes per collide launch from the committed counter-pass summary (FETCH_SIZE and
    WRITE_SIZE, calibrated with a copy kernel of known size), or None when absent."""
    try:
        for line in open(TRAFFIC_SUMMARY):
            if line.startswith(key + ":"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


METRIC = "MLUPS & %HBM roofline, FP32 vs FP16-store; shear-cell solves/s + p99 latency"


def c2_deck(storage=0, steps=None):
    from cases import deck
    n, tau, U, H = 0.7, 0.9, 0.02, 128.0
    nu = (tau - 0.5) / 3.0
    K = nu * (2 * U / H) ** (1 - n)
    return deck(size=128, walls=(0, 0, 1), tau=tau, uw=U, np_=32, diam=12.0, markers=192, dist=4, seed=2024, gap=2.0,
                coll=1, ibm=1, visc=2, vparams=(K, n, 0, 0), local=1, clamp=(0.51, 5.0), storage=storage)


def cpu_baseline(sample_steps):
    """The oracle (float32, OpenMP) on `sample_steps` steps of the same workload."""
    import d3q19
    import oracle_py
    c = d3q19.Case(lines=c2_deck())
    o = oracle_py.Oracle(c, "f32").load_case()
    o.step(1)  # warm-up (page-in)
    t0 = time.perf_counter()
    o.step(sample_steps)
    dt = time.perf_counter() - t0
    nodes = (c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(nodes * sample_steps / dt / 1e6, 3), "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": "%d steps of the same 128^3 power-law, 32-sphere workload (oracle/d3q19_oracle.c, float32, "
                      "OpenMP), %.1f s" % (sample_steps, dt)}


def run(args):
    import d3q19
    c = d3q19.Case(lines=c2_deck(storage=0))
    nodes = (c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2)
    sim = d3q19.Sim(c)
    sim.upload_case()
    sim.run(1, args.warmup)
    sim.sync()
    total_ms, collide_ms = sim.time_steps(1 + args.warmup, args.steps)
    sim.sync()
    assert sim.check_finite() == 0, "non-finite state"
    mlups = nodes * args.steps / (total_ms * 1e-3) / 1e6
    # collide+stream roofline: algorithmic bytes 19 loads + 19 stores of fp32 per node
    bytes_per_launch = 152.0 * nodes
    achieved = bytes_per_launch / (collide_ms * 1e-3) / 1e9
    sim.close()

    c16 = d3q19.Case(lines=c2_deck(storage=1))
    s16 = d3q19.Sim(c16)
    s16.upload_case()
    s16.run(1, args.warmup)
    s16.sync()
    t16, col16 = s16.time_steps(1 + args.warmup, args.steps)
    s16.close()
    mlups16 = nodes * args.steps / (t16 * 1e-3) / 1e6
    ach16 = 76.0 * nodes / (col16 * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sphere placement, fluid at rest)",
        "config": {"workload": "128^3 power-law shear cell (LatticeSize 130x130x131, walls +-0.02, n=0.7, "
                               "tau clamp [0.51,5]), 32 spheres r=6 x 192 markers, BGK, Peskin-4 IBM, fp32 populations",
                   "interior_nodes": nodes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("traffic_bytes_per_launch_fp32"),
                     "traffic_source": "profiles/hbm_traffic_collide_c2_powerlaw.md (rocprofv3 FETCH_SIZE x2.0, WRITE_SIZE x1.0, "
                                       "scaled by a 1 GiB copy)",
                     "algorithmic_bytes": int(bytes_per_launch),
                     "kernel": "collide_kernel (fused collide+stream), 152 B/node algorithmic",
                     "kernel_ms": round(collide_ms, 5)},
        "fp16_store": {"mlups": round(mlups16, 1), "ms_per_step": round(t16 / args.steps, 5),
                       "collide_ms": round(col16, 5), "collide_achieved_GBs": round(ach16, 1),
                       "collide_traffic": measured_traffic("traffic_bytes_per_launch_fp16"),
                       "speedup_vs_fp32": round(mlups16 / mlups, 3)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


COPY_ELEMS = 1 << 28  # float32 elements of the calibration copy (1 GiB read + 1 GiB written)


def traffic_probe():
    """Launches for the HBM-traffic counter passes (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE):
    5 direct collide+stream launches of the workload in fp32 and in fp16-shifted storage, then a
    plain elementwise copy of known size (torch.neg over 1 GiB of float32) whose counter readings
    calibrate the counters (expected 1 GiB fetched and 1 GiB written per launch)."""
    import d3q19
    import torch
    for storage in (0, 1):
        c = d3q19.Case(lines=c2_deck(storage=storage))
        s = d3q19.Sim(c)
        s.upload_case()
        s.run(1, 4)
        s.sync()
        s.time_collide(5)
        s.sync()
        s.close()
        print("probe: collide storage=%d done" % storage, flush=True)
    src = torch.ones(COPY_ELEMS, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print("probe: copy of %d bytes read + %d bytes written, 3 launches" % (4 * COPY_ELEMS, 4 * COPY_ELEMS), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only the launches profiled by the HBM-traffic counter passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (a cell is not split across GPUs); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe()
        return
    run(args)


if __name__ == "__main__":
    main()
