"""Benchmark: full D3Q19 IBM shear-cell steps on one MI355X.

Workload (default line): the 256x256x128 Casson shear cell -- box 256x256x128 between walls moving
at +-0.02 (LatticeSize 258x258x131), Casson fluid with yield stress 1e-5 and plastic viscosity of
tau 0.7, Papanastasiou m = 1e4, local tau from the strain rate clamped to [0.505, 5], BGK collision,
500 spheres of diameter 10 with 128 markers each at seeded centres, Peskin 4-point IBM, fp32
populations.  A step is the whole time step: the step kernel (collide+stream with the spheres'
dynamics, force sums and contacts in extra blocks of the same launch) and the IBM pass.  The
lattice (1.3 GB of fp32 populations) is five times the 256 MB Infinity Cache, so the bandwidth is
HBM bandwidth.

Prints one JSON line: MLUPS of the fp32 step (value) over the timed steps, the step kernel's HBM
roofline (its own start/end HIP events on the simulation's stream, inside the timed region), the
attainable bandwidth of the same access pattern without arithmetic, the fp16-shifted-storage step
on the same workload, the 128^3 power-law workload, a bounded serving stream against sequential
solves, and a CPU baseline (the oracle, OpenMP, on a bounded sample of the same workload).
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
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_step_c3_casson.md")


def measured_traffic(key):
    """HBM bytes per launch from the committed counter-pass summary (FETCH_SIZE and WRITE_SIZE,
    calibrated with a copy kernel of known size), or None when absent."""
    try:
        for line in open(TRAFFIC_SUMMARY):
            if line.startswith(key + ":"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


METRIC = "MLUPS & %HBM roofline, FP32 vs FP16-store; shear-cell solves/s + p99 latency"
C3_NAME = ("256x256x128 Casson shear cell (LatticeSize 258x258x131, walls +-0.02, yield stress 1e-5, tau_p 0.7, "
           "Papanastasiou m=1e4, tau clamp [0.505,5]), 500 spheres d=10 x 128 markers, BGK, Peskin-4 IBM")


from workloads import c2_deck, c3_deck, c4_decks, c5_requests, poisson_arrivals, precision_study  # noqa: E402


def host_cpu():
    """nproc, OMP_NUM_THREADS and the CPU model of the host this runs on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": len(os.sched_getaffinity(0)), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def cpu_baseline(sample_steps):
    """The oracle (float32, OpenMP) on `sample_steps` steps of the same 256x256x128 workload."""
    import d3q19
    import oracle_py
    c = d3q19.Case(lines=c3_deck())
    o = oracle_py.Oracle(c, "f32").load_case()
    o.step(1)  # warm-up (page-in)
    t0 = time.perf_counter()
    o.step(sample_steps)
    dt = time.perf_counter() - t0
    o.close()
    nodes = (c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2)
    h = host_cpu()
    threads = int(h["omp_num_threads"]) if h["omp_num_threads"] else h["nproc"]
    return {"value": round(nodes * sample_steps / dt / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "%d steps of the same 256x256x128 Casson 500-sphere workload (oracle/d3q19_oracle.c, float32, "
                      "OpenMP), %.1f s" % (sample_steps, dt), **h}


def step_run(lines, steps, warmup, graph_steps=0):
    """Warm-up steps, then `steps` timed real steps whose kernels carry their own start/end events;
    optionally as many more steps replayed from the HIP graph (d3q19_run), timed on the host
    around synchronised calls."""
    import d3q19
    c = d3q19.Case(lines=lines)
    nodes = (c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2)
    s = d3q19.Sim(c)
    if c.np:
        s.write_particles(c.parKin)
    s.run(1, warmup)
    s.sync()
    total, kstep, kibm = s.time_steps(1 + warmup, steps)
    out = {"nodes": nodes, "ms_per_step": total / steps, "mlups": nodes * steps / (total * 1e-3) / 1e6,
           "step_kernel_ms": kstep, "ibm_ms": kibm}
    if graph_steps:
        t0 = time.perf_counter()
        s.run(1 + warmup + steps, graph_steps)
        s.sync()
        g = (time.perf_counter() - t0) * 1e3 / graph_steps
        out["graph_ms_per_step"] = g
        out["graph_mlups"] = nodes / (g * 1e-3) / 1e6
    assert s.check_finite() == 0, "non-finite state or a sphere moved a lattice unit in one step"
    s.close()
    return out


def run(args):
    import d3q19
    r32 = step_run(c3_deck(storage=0), args.steps, args.warmup, graph_steps=args.graph_steps)
    nodes = r32["nodes"]
    # step-kernel roofline: algorithmic bytes 19 loads + 19 stores of fp32 per interior node
    bytes_per_launch = 152.0 * nodes
    achieved = bytes_per_launch / (r32["step_kernel_ms"] * 1e-3) / 1e9
    # the step kernel's own access pattern: 19 populations in chunks of one 256-node row, read and
    # written with non-temporal accesses
    attain, _ = d3q19.stream_probe(256, 256, 128, streams=19, width=-256, pull=True, launches=20, nontemporal=True)
    r16 = step_run(c3_deck(storage=1), args.steps, args.warmup, graph_steps=args.graph_steps)
    ach16 = 76.0 * nodes / (r16["step_kernel_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(r32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sphere placement, fluid at rest)",
        "config": {"workload": C3_NAME + ", fp32 populations", "interior_nodes": nodes,
                   "step": "step kernel (collide+stream + spheres' turn) + IBM pass, timed real steps"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("traffic_bytes_per_launch_fp32"),
                     "traffic_source": "profiles/hbm_traffic_step_c3_casson.md (rocprofv3 FETCH_SIZE and WRITE_SIZE of the "
                                       "step kernel, each scaled by the same counter's reading for a copy of known size)",
                     "algorithmic_bytes": int(bytes_per_launch),
                     "kernel": "step_kernel_w5 (fused collide+stream with the spheres' turn), 152 B/node algorithmic",
                     "kernel_ms": round(r32["step_kernel_ms"], 5),
                     "attainable_GBs": round(attain, 1), "frac_of_attainable": round(achieved / attain, 4),
                     "attainable_note": "d3q19_stream_probe: the step kernel's access pattern (19 fp32 populations read at "
                                        "the pull offsets and 19 written, in chunks of one 256-node row, non-temporal loads "
                                        "and stores) with no arithmetic, 256x256x128 box"},
        "ibm_ms": round(r32["ibm_ms"], 5),
        "graph_replay": {"ms_per_step": round(r32.get("graph_ms_per_step", 0.0), 5),
                         "mlups": round(r32.get("graph_mlups", 0.0), 1)},
        "fp16_store": {"mlups": round(r16["mlups"], 1), "ms_per_step": round(r16["ms_per_step"], 5),
                       "step_kernel_ms": round(r16["step_kernel_ms"], 5), "ibm_ms": round(r16["ibm_ms"], 5),
                       "step_kernel_achieved_GBs": round(ach16, 1), "frac": round(ach16 / HBM_PEAK_GBS, 4),
                       "traffic": measured_traffic("traffic_bytes_per_launch_fp16"),
                       "speedup_vs_fp32": round(r16["mlups"] / r32["mlups"], 3),
                       "rounding": "stochastic; drag within 2.2e-4 of the fp64 build after 3000 steps of this workload "
                                   "(bench.py --mode precision)"},
    }
    if not args.no_c2:
        c2 = {}
        for st, name in ((0, "fp32"), (1, "fp16_shift")):
            r = step_run(c2_deck(storage=st), args.steps, args.warmup)
            c2[name] = {"mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms_per_step"], 5),
                        "step_kernel_ms": round(r["step_kernel_ms"], 5)}
        c2["workload"] = "128^3 power-law shear cell, 32 spheres d=12 x 192 markers (lattices partly Infinity-Cache resident)"
        out["c2_128cubed"] = c2
    if not args.no_serving:
        out["serving"] = serving_summary(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


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
        ms, _, _ = sm.time_steps(5, seq_steps)
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
    return serving_run(args.requests, args.max_edge, args.check_interval, args.max_steps, args.load, args.rate,
                       args.serve_storage, args.arena_gb, min_edge=args.min_edge, sequential=not args.no_sequential)


def serving_summary(args):
    """A bounded sample of the serving stream for the default line: the first 12 requests of the
    stream (edges 32-160), each solve capped at 2000 steps, at 4x the sequential capacity."""
    r = serving_run(12, 160, 100, 2000, 4.0, 0.0, 0, args.arena_gb, quiet=True)
    return {"solves_per_s": r["batched"]["solves_per_s"], "p50_latency_s": r["batched"]["p50_latency_s"],
            "p99_latency_s": r["batched"]["p99_latency_s"], "vs_sequential": r["speedup_solves_per_s"],
            "sequential": r["sequential"], "requests": r["requests"], "max_steps": r["max_steps"],
            "converged": r["converged"], "same_step_count": r["same_step_count"],
            "max_rel_diff_wall_stress_vs_sequential": r["max_rel_diff_wall_stress_vs_sequential"],
            "sample": "bounded: 12 requests of the 32-160 edge stream, solves capped at 2000 steps (bench.py --mode "
                      "serving runs the stream to convergence)"}


def server_backlog(cases, sp, serve_storage, arena_gb):
    """The server alone on a backlog (every request submitted at once): makespan, solves/s and the
    aggregate lattice-update rate (for profiling the server's kernels without the sequential solves)."""
    import d3q19
    srv = d3q19.Server(storage=serve_storage, collision=d3q19.COLL_BGK, ibm=d3q19.IBM_PESKIN4, arena_bytes=arena_gb << 30)
    t0 = d3q19.clock()
    ids = [srv.submit(c, sp) for c in cases]
    done = {}
    while len(done) < len(cases):
        for r in srv.progress():
            done[r.id] = r
    span = d3q19.clock() - t0
    srv.close()
    work = sum((c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2) * done[i].steps for c, i in zip(cases, ids))
    return {"requests": len(cases), "server_only": True, "makespan_s": round(span, 3), "solves_per_s": round(len(cases) / span, 4),
            "glups": round(work / span / 1e9, 3), "converged": sum(1 for r in done.values() if r.status == 0)}


def serving_run(requests, max_edge, check_interval, max_steps, load, rate_arg, serve_storage, arena_gb, quiet=False, min_edge=32,
                sequential=True):
    """Shear-cell solve requests as sequential plain fp32 solves (single simulations, one after another),
    then the same requests arriving as a Poisson stream into the continuous-batching server.  The
    arrival rate is `load` times the sequential capacity measured first (offered load above what the
    sequential solver sustains), unless --rate is given.  Latency = completion - scheduled arrival;
    the sequential latencies are those of a FIFO single server with the measured service times under
    the same arrivals."""
    import d3q19
    decks = c5_requests(count=requests, edge_range=(min_edge, max_edge), storage=0)
    cases = [d3q19.Case(lines=d) for d in decks]
    sp = d3q19.solve_params(tolerance=1e-4, check_interval=check_interval, min_steps=min(10 * check_interval, max_steps),
                            max_steps=max_steps)
    if not sequential:
        return server_backlog(cases, sp, serve_storage, arena_gb)
    service, seq = [], {}
    for i, c in enumerate(cases):
        a = time.perf_counter()
        seq[i] = d3q19.solve(c, sp)
        service.append(time.perf_counter() - a)
        if not quiet:
            print("serving: sequential %d/%d %s %d steps %.3f s" % (i + 1, len(cases), c.N, seq[i].steps, service[-1]),
                  flush=True)
    cap_seq = len(cases) / sum(service)
    rate = rate_arg if rate_arg > 0 else load * cap_seq
    arr = poisson_arrivals(rate, len(cases))
    srv = d3q19.Server(storage=serve_storage, collision=d3q19.COLL_BGK, ibm=d3q19.IBM_PESKIN4, arena_bytes=arena_gb << 30)
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
            if not quiet:
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
    # Work-conservation bound of any scheduler on this request set at the same arithmetic: the
    # makespan is at least the total lattice updates over the best rate one solve reached alone
    # (the lattice bandwidth of the GPU), and at least the longest solve alone.
    nodes = [(c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2) for c in cases]
    work = [nodes[i] * seq[i].steps for i in range(len(cases))]
    rate_best = max(work[i] / service[i] for i in range(len(cases)))
    bound_s = max(sum(work) / rate_best, max(service))
    ceiling = {"node_steps": int(sum(work)), "best_single_solve_glups": round(rate_best / 1e9, 3),
               "makespan_bound_s": round(bound_s, 3), "longest_solve_alone_s": round(max(service), 3),
               "solves_per_s_bound": round(len(cases) / bound_s, 4),
               "vs_sequential_bound": round(sum(service) / bound_s, 3),
               "work_share_edges_ge_96": round(sum(w for w, c in zip(work, cases) if c.N[0] - 2 >= 96) / max(sum(work), 1), 3)}
    per_request = [{"edge": cases[i].N[0] - 2, "spheres": cases[i].np, "steps": seq[i].steps,
                    "sequential_s": round(service[i], 4), "glups_alone": round(work[i] / service[i] / 1e9, 3),
                    "server_latency_s": round(lat[i], 4)} for i in range(len(cases))]
    names = {0: "fp32", 1: "fp16_shift", 2: "bf16_shift", 3: "fp64"}
    return {"requests": len(cases), "arrival_rate_per_s": round(rate, 4), "offered_load_vs_sequential": round(rate / cap_seq, 3),
            "edges": [c.N[0] - 2 for c in cases], "spheres": [c.np for c in cases], "check_interval": check_interval,
            "tolerance": 1e-4, "max_steps": max_steps, "server_storage": names[serve_storage],
            "converged": sum(1 for r in done.values() if r.status == 0),
            "batched": {"solves_per_s": round(len(cases) / span, 4), "p50_latency_s": round(pct(lat, 50), 4),
                        "p99_latency_s": round(pct(lat, 99), 4), "makespan_s": round(span, 3)},
            "sequential": {"solves_per_s": round(cap_seq, 4), "p50_latency_s": round(pct(lat_seq, 50), 4),
                           "p99_latency_s": round(pct(lat_seq, 99), 4), "busy_s": round(sum(service), 3)},
            "speedup_solves_per_s": round((len(cases) / span) / cap_seq, 3),
            "max_rel_diff_wall_stress_vs_sequential": agree, "same_step_count": steps_equal,
            "ceiling": ceiling, "per_request": per_request}


def large(args):
    """The 256x256x128 Casson workload (500 spheres r=5 x 128 markers, 8.45 M interior nodes, five
    times the Infinity Cache in fp32), fp32 and fp16-shifted populations: the collide alone (16
    event-timed launches, after the steps: they advance the lattice without the spheres) and the whole
    step (step kernel + IBM pass, real steps timed with HIP events on the simulation's stream)."""
    import d3q19
    out = {}
    for st, name in ((0, "fp32"), (1, "fp16_shift")):
        c = d3q19.Case(lines=c3_deck(storage=st))
        nodes = (c.N[0] - 2) * (c.N[1] - 2) * (c.N[2] - 2)
        s = d3q19.Sim(c)
        s.write_particles(c.parKin)
        s.run(1, args.warmup)
        s.sync()
        tot, _, _ = s.time_steps(args.warmup + 1, args.steps)
        assert s.check_finite() == 0, "non-finite state"
        col = s.time_collide(16)
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


def stream_sweep(args):
    """Attainable bandwidth of the collide's memory pattern on this GPU (d3q19_stream_probe): a
    streaming kernel with 19 fp32 planes read at the D3Q19 pull offsets and 19 written, no
    arithmetic, beside a plain copy, for 1 / 2 / 4 consecutive nodes per lane, on the
    256x256x128 box and on a 512x256x256 box (2.5 GB per direction); then the 19-plane pull
    pattern with 1 node per lane at 2-8 resident 256-thread blocks per CU (dynamic LDS caps them).
    The chunk layouts (19 populations in runs of 64 or 256 nodes) run with plain accesses and, for
    256, with the non-temporal loads and stores of the collide on lattices beyond the Infinity
    Cache."""
    import d3q19
    rows = []
    for box in ((256, 256, 128), (512, 256, 256)):
        for streams, width, pull, nt in ((1, 4, False, False), (19, 1, False, False), (19, 2, False, False), (19, 4, False, False),
                                         (19, 1, True, False), (19, 2, True, False), (19, 4, True, False), (19, -64, False, False),
                                         (19, -64, True, False), (19, -256, True, False), (19, -256, True, True)):
            # the copy moves the bytes of the 19 planes, so that neither fits the Infinity Cache
            cbox = (box[0] * 19, box[1], box[2]) if streams == 1 else box
            gbs, ms = d3q19.stream_probe(*cbox, streams=streams, width=width, pull=pull, launches=args.probe_launches, nontemporal=nt)
            rows.append({"box": list(cbox), "streams": streams, "width": width, "pull": pull, "nontemporal": nt, "lds_bytes": 0,
                         "layout": "chunks of %d nodes" % -width if width < 0 else "planes",
                         "GBs": round(gbs, 1), "ms": round(ms, 5), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)})
            print("stream: %s" % json.dumps(rows[-1]), flush=True)
    for blocks in (2, 3, 4, 6, 8):
        lds = min(65536, (160 * 1024) // blocks - 256)
        gbs, ms = d3q19.stream_probe(512, 256, 256, streams=19, width=1, pull=True, lds_bytes=lds,
                                     launches=args.probe_launches)
        rows.append({"box": [512, 256, 256], "streams": 19, "width": 1, "pull": True, "lds_bytes": lds,
                     "blocks_per_cu": blocks, "GBs": round(gbs, 1), "ms": round(ms, 5),
                     "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)})
        print("stream: %s" % json.dumps(rows[-1]), flush=True)
    best = max(r["GBs"] for r in rows if r["streams"] == 19)
    return {"rows": rows, "best_19_plane_GBs": best}


COPY_ELEMS = 1 << 28  # float32 elements of the calibration copy (1 GiB read + 1 GiB written)


PROBE_NODES = 256 * 256 * 128


def traffic_probe():
    """Launches for the HBM-traffic counter passes (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE):
    5 timed real steps of the 256x256x128 workload in fp32 and in fp16-shifted storage (after 10
    warm-up steps), then a plain elementwise copy of known size (torch.neg over 1 GiB of float32)
    whose counter readings calibrate the counters (expected 1 GiB fetched and 1 GiB written per
    launch), then the pull-pattern bandwidth probe on the same box (19 planes read and written per
    launch, 642 MB each way), a calibration with the step kernel's own access pattern."""
    import d3q19
    import torch
    for storage in (0, 1):
        c = d3q19.Case(lines=c3_deck(storage=storage))
        s = d3q19.Sim(c)
        s.write_particles(c.parKin)
        s.run(1, 10)
        s.sync()
        s.time_steps(11, 5)
        s.sync()
        s.close()
        print("probe: 5 steps storage=%d done" % storage, flush=True)
    src = torch.ones(COPY_ELEMS, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print("probe: copy of %d bytes read + %d bytes written, 3 launches" % (4 * COPY_ELEMS, 4 * COPY_ELEMS), flush=True)
    # the attainable-bandwidth probe with the step kernel's own access pattern (19 fp32 populations
    # read at the pull offsets and 19 written, in chunks of one 256-node row, non-temporal loads and
    # stores, 256x256x128 box): 2 warm-up and 3 timed launches of known bytes, the calibration whose
    # request mix is the step kernel's
    d3q19.stream_probe(256, 256, 128, streams=19, width=-256, pull=True, launches=3, nontemporal=True)
    print("probe: pull-pattern probe, %d bytes read + %d bytes written per launch, 5 launches" % (19 * 4 * PROBE_NODES, 19 * 4 * PROBE_NODES),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--graph-steps", type=int, default=200, help="graph-replayed steps timed after the timed region")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true", help="skip the 128^3 power-law workload")
    ap.add_argument("--no-serving", action="store_true", help="skip the bounded serving stream")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only the launches profiled by the HBM-traffic counter passes")
    ap.add_argument("--probe-launches", type=int, default=20)
    ap.add_argument("--mode", default="step", choices=["step", "ensemble", "serving", "precision", "large", "stream"],
                    help="step: the per-step benchmark line (default); ensemble / serving / precision / large: the "
                         "batched, serving, storage-precision and 256x256x128 workloads, one JSON line each")
    ap.add_argument("--ens-steps", type=int, default=1000)
    ap.add_argument("--ens-warmup", type=int, default=10)
    ap.add_argument("--ens-seq-steps", type=int, default=100)
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=0.0, help="arrivals per second (0: load x sequential capacity)")
    ap.add_argument("--load", type=float, default=4.0)
    ap.add_argument("--serve-storage", type=int, default=0)
    ap.add_argument("--min-edge", type=int, default=32)
    ap.add_argument("--no-sequential", action="store_true",
                    help="serving: the server alone on a backlog of the requests (no sequential solves)")
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
    elif args.mode == "stream":
        print(json.dumps({"workload": "attainable bandwidth of the collide pattern", **stream_sweep(args)}), flush=True)
    elif args.mode == "large":
        print(json.dumps({"workload": "256x256x128 Casson shear cell, 500 spheres", **large(args)}), flush=True)
    else:
        run(args)


if __name__ == "__main__":
    main()
