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


from workloads import c2_deck, c3_deck, c4_decks, c5_requests, poisson_arrivals, precision_study  # noqa: E402


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
                       "speedup_vs_fp32": round(mlups16 / mlups, 3),
                       "rounding": "stochastic (drag within 2e-4 of the fp64 build after 5000 steps)",
                       "note": "the two 80 MB fp16 lattices fit the 256 MB Infinity Cache, so this collide reads partly "
                               "from cache; see roofline_256x256x128 for a cache-independent lattice"},
    }
    if not args.no_large:
        out["roofline_256x256x128"] = large_lattice_collide()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


def large_lattice_collide():
    """The collide on the 256x256x128 Casson lattice with 500 spheres (8.45 M nodes, 1.28 GB of
    fp32 populations moved per launch, five times the 256 MB Infinity Cache), so that the achieved
    bandwidth cannot come from cache residency; 16 launches timed with HIP events on the
    simulation's stream."""
    import d3q19
    c = d3q19.Case(lines=c3_deck(storage=0))
    nodes = (c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2)
    s = d3q19.Sim(c)
    s.write_particles(c.parKin)
    s.run(1, 10)
    s.sync()
    ms = s.time_collide(16)
    assert s.check_finite() == 0, "non-finite state"
    s.close()
    ach = 152.0 * nodes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "kernel_ms": round(ms, 5), "interior_nodes": nodes,
            "workload": "256x256x128 Casson shear cell, 500 spheres r=5 x 128 markers, fp32 populations, 152 B/node"}


def pct(a, q):
    import numpy as np
    return float(np.percentile(np.asarray(a, dtype=np.float64), q)) if len(a) else None


def ensemble(args):
    """64 independent shear cells (uneven edges, sphere sets and rheologies) in fp16-shifted storage,
    stepped together by the batched persistent kernels with the autotuned launch shape; compared with
    the same cells stepped one after another as single simulations (graph-replayed too)."""
    import d3q19
    decks = c4_decks(storage=1)
    cases = [d3q19.Case(lines=d) for d in decks]
    nodes = sum((c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2) for c in cases)
    b = d3q19.Batch(storage=1, collision=d3q19.COLL_BGK, ibm=d3q19.IBM_PESKIN4, arena_bytes=16 << 30)
    ids = [b.add(c) for c in cases]
    print("ensemble: %d cells, %d nodes, %d spheres" % (len(ids), nodes, sum(c.np for c in cases)), flush=True)
    b.run(args.ens_warmup)
    b.sync()
    t = b.autotune()
    print("ensemble: autotuned %s" % t.as_dict(), flush=True)
    total_ms, col_ms = b.time_steps(args.ens_steps)
    mlups = nodes * args.ens_steps / (total_ms * 1e-3) / 1e6
    mem = b.memory()
    b.close()
    seq_steps = args.ens_seq_steps
    seq_ms = 0.0
    for c in cases:
        sm = d3q19.Sim(c)
        if c.np:
            sm.write_particles(c.parKin)
        sm.run(1, 4)
        sm.sync()
        ms, _ = sm.time_steps(5, seq_steps)
        seq_ms += ms
        sm.close()
    mlups_seq = nodes * seq_steps / (seq_ms * 1e-3) / 1e6
    return {"cells": len(ids), "nodes": nodes, "steps": args.ens_steps, "storage": "fp16_shift",
            "mlups": round(mlups, 1), "ms_per_step": round(total_ms / args.ens_steps, 4),
            "collide_ms": round(col_ms, 4), "collide_achieved_GBs": round(76.0 * nodes / (col_ms * 1e-3) / 1e9, 1),
            "tuning": t.as_dict(), "arena_used_bytes": mem[1],
            "sequential_single_mlups": round(mlups_seq, 1), "sequential_steps_per_cell": seq_steps,
            "speedup_vs_sequential": round(mlups / mlups_seq, 3)}


def serving(args):
    """Shear-cell solve requests as sequential plain fp32 solves (single simulations, one after another),
    then the same requests arriving as a Poisson stream into the continuous-batching server.  The
    arrival rate is `load` times the sequential capacity measured first (offered load above what the
    sequential solver sustains), unless --rate is given.  Latency = completion - scheduled arrival;
    the sequential latencies are those of a FIFO single server with the measured service times under
    the same arrivals."""
    import d3q19
    decks = c5_requests(count=args.requests, edge_range=(32, args.max_edge), storage=0)
    cases = [d3q19.Case(lines=d) for d in decks]
    sp = d3q19.solve_params(tolerance=1e-4, check_interval=args.check_interval, min_steps=10 * args.check_interval,
                            max_steps=args.max_steps)
    service, seq = [], {}
    for i, c in enumerate(cases):
        a = time.perf_counter()
        seq[i] = d3q19.solve(c, sp)
        service.append(time.perf_counter() - a)
        print("serving: sequential %d/%d %s %d steps %.3f s" % (i + 1, len(cases), c.N, seq[i].steps, service[-1]), flush=True)
    cap_seq = len(cases) / sum(service)
    rate = args.rate if args.rate > 0 else args.load * cap_seq
    arr = poisson_arrivals(rate, len(cases))
    srv = d3q19.Server(storage=args.serve_storage, collision=d3q19.COLL_BGK, ibm=d3q19.IBM_PESKIN4,
                       arena_bytes=args.arena_gb << 30)
    t0 = d3q19.clock()
    nxt = 0
    rid_of = {}
    done = {}
    while len(done) < len(cases):
        now = d3q19.clock() - t0
        while nxt < len(cases) and arr[nxt] <= now:
            rid_of[srv.submit(cases[nxt], sp)] = nxt
            nxt += 1
        for r in srv.progress():
            done[rid_of[r.id]] = r
            print("serving: batched %d/%d done (request %d, %d steps)" % (len(done), len(cases), rid_of[r.id], r.steps),
                  flush=True)
        q, res, _ = srv.status()
        if q == 0 and res == 0 and nxt < len(cases):
            time.sleep(max(0.0, arr[nxt] - (d3q19.clock() - t0)))
    srv.close()
    lat = [done[i].t_finish - (t0 + arr[i]) for i in range(len(cases))]
    span = max(done[i].t_finish for i in done) - (t0 + arr[0])
    fin, lat_seq = 0.0, []
    for i in range(len(cases)):
        fin = max(fin, arr[i]) + service[i]
        lat_seq.append(fin - arr[i])
    agree = max(abs(done[i].obs.wall_stress - seq[i].obs.wall_stress) / max(abs(seq[i].obs.wall_stress), 1e-300)
                for i in range(len(cases)))
    steps_equal = sum(1 for i in range(len(cases)) if done[i].steps == seq[i].steps)
    names = {0: "fp32", 1: "fp16_shift", 2: "bf16_shift", 3: "fp64"}
    return {"requests": len(cases), "arrival_rate_per_s": round(rate, 4), "offered_load_vs_sequential": round(rate / cap_seq, 3),
            "edges": [c.N[0] - 2 for c in cases], "spheres": [c.np for c in cases], "check_interval": args.check_interval,
            "tolerance": 1e-4, "max_steps": args.max_steps, "server_storage": names[args.serve_storage],
            "converged": sum(1 for r in done.values() if r.status == 0),
            "batched": {"solves_per_s": round(len(cases) / span, 4), "p50_latency_s": round(pct(lat, 50), 4),
                        "p99_latency_s": round(pct(lat, 99), 4), "makespan_s": round(span, 3)},
            "sequential": {"solves_per_s": round(cap_seq, 4), "p50_latency_s": round(pct(lat_seq, 50), 4),
                           "p99_latency_s": round(pct(lat_seq, 99), 4), "busy_s": round(sum(service), 3)},
            "speedup_solves_per_s": round((len(cases) / span) / cap_seq, 3),
            "max_rel_diff_wall_stress_vs_sequential": agree, "same_step_count": steps_equal}


def large(args):
    """The 256x256x128 Casson workload (500 spheres r=5 x 128 markers, 8.45 M interior nodes, five
    times the Infinity Cache in fp32), fp32 and fp16-shifted populations: the collide alone (16
    event-timed launches) and the whole step (fused step kernel + IBM pass, graph-replayed, timed
    with HIP events on the simulation's stream)."""
    import d3q19
    out = {}
    for st, name in ((0, "fp32"), (1, "fp16_shift")):
        c = d3q19.Case(lines=c3_deck(storage=st))
        nodes = (c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2)
        s = d3q19.Sim(c)
        s.write_particles(c.parKin)
        s.run(1, args.warmup)
        s.sync()
        col = s.time_collide(16)
        tot, _ = s.time_steps(args.warmup + 1, args.steps)
        assert s.check_finite() == 0, "non-finite state"
        s.close()
        ms = tot / args.steps
        out[name] = {"step_ms": round(ms, 5), "mlups": round(nodes / (ms * 1e-3) / 1e6, 1), "collide_ms": round(col, 5),
                     "collide_achieved_GBs": round((152.0 if st == 0 else 76.0) * nodes / (col * 1e-3) / 1e9, 1)}
        print("large: %s done" % name, flush=True)
    out["fp16_speedup_step"] = round(out["fp32"]["step_ms"] / out["fp16_shift"]["step_ms"], 3)
    out["interior_nodes"] = nodes
    return out


def precision(args):
    """The 256x256x128 Casson precision workload: fp32 / fp16-shifted / bf16-shifted populations
    against the fp64 build, errors of u, drag coefficient and wall stress."""
    return precision_study(c3_deck, args.prec_steps)


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
    ap.add_argument("--no-large", action="store_true", help="skip the 256x256x128 collide roofline")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only the launches profiled by the HBM-traffic counter passes")
    ap.add_argument("--mode", default="step", choices=["step", "ensemble", "serving", "precision", "large"],
                    help="step: the per-step benchmark line (default); ensemble / serving / precision / large: the "
                         "batched, serving, storage-precision and 256x256x128 workloads, one JSON line each")
    ap.add_argument("--ens-steps", type=int, default=1000)
    ap.add_argument("--ens-warmup", type=int, default=10)
    ap.add_argument("--ens-seq-steps", type=int, default=100)
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=0.0, help="arrivals per second (0: load x sequential capacity)")
    ap.add_argument("--load", type=float, default=4.0)
    ap.add_argument("--serve-storage", type=int, default=0)
    ap.add_argument("--max-edge", type=int, default=160)
    ap.add_argument("--check-interval", type=int, default=100)
    ap.add_argument("--max-steps", type=int, default=40000)
    ap.add_argument("--arena-gb", type=int, default=64)
    ap.add_argument("--prec-steps", type=int, default=3000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (a cell is not split across GPUs); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe()
        return
    if args.mode == "ensemble":
        print(json.dumps({"workload": "batched ensemble", **ensemble(args)}), flush=True)
    elif args.mode == "serving":
        print(json.dumps({"workload": "serving", **serving(args)}), flush=True)
    elif args.mode == "precision":
        print(json.dumps({"workload": "precision study", **precision(args)}), flush=True)
    elif args.mode == "large":
        print(json.dumps({"workload": "256x256x128 Casson shear cell, 500 spheres", **large(args)}), flush=True)
    else:
        run(args)


if __name__ == "__main__":
    main()
