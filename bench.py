"""Benchmark of the PN-CSD forward+adjoint hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload serve|raster2d_grad] [--requests R]

serve (default): the mixed serving stream -- R forward+gradient requests (PN
order 3-9, 1D 100-400 cells or 2D 64^2-192^2, 10-25 keV, 2-4 elements) with
Poisson arrivals, served by the continuous-batching server under an HBM
budget.  A "step" is one pass of the whole stream; the timed region covers K
passes with all request inputs prepared in host memory beforehand.  Reported
beside it: p50/p99 latency (arrival to result), the plain one-request-at-a-time
GPU loop on a sample of the same requests, and the C oracle on one core.
The timed passes offer 1.5x the capacity measured on the head of the stream,
so the server stays saturated and the throughput is its own; p50/p99 come
from an untimed pass at 0.8x the measured throughput.

raster2d_grad: 2D P7 128x96, 32 beams, forward+adjoint of the batch per step.

Every mode also measures the roofline figure of the dominant kernel (the fused
even-parity half step of the 2D P7 128x96 batch) with HIP events on the
library's stream.  One JSON line is printed.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The server gives each group of PN orders its own stream; the runtime maps
# streams onto GPU_MAX_HW_QUEUES hardware queues per process (4 by default) and
# serialises streams that share one.  Six queues leave one each for the default
# stream, the context stream (1D lane) and four 2D/3D lanes.  Read by the HIP
# runtime when it initialises, so it is set before anything touches the GPU
# (raised to 6 when the environment asks for fewer, as the default 4 does).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 6:
    os.environ["GPU_MAX_HW_QUEUES"] = "6"

HBM_PEAK_GBS = 8000.0
METRIC = "served fwd+adj solves/s, p50/p99 latency vs plain GPU loop @ k-ratio err<=1e-10"


def _prep_one(spec):
    from pnepma_amd import serving
    return serving.prepare(spec)


def prepare_stream(n, seed=3):
    """Draw and prepare the request stream on the host (process pool, before the GPU is touched)."""
    import multiprocessing as mp
    from pnepma_amd import serving, physics
    for e in physics.ELEMENTS.values():
        physics.transport_coefficients(e, 1000.0, 9)
    specs = serving.serving_requests(n, seed=seed)
    workers = max(1, min(16, (os.cpu_count() or 1)))
    with mp.get_context("fork").Pool(workers) as pool:
        preps = pool.map(_prep_one, specs, chunksize=8)
    return specs, preps


PROBE_BEAMS = 160  # 160 beams x 3.57 MB of state: the working set exceeds the 256 MB Infinity Cache
PROFILE = os.path.join(ROOT, "profiles", "half_step_p7_2d_128x96_160beams.md")


def measured_traffic():
    """HBM bytes per launch of the probe kernel from the committed rocprofv3 --pmc summary."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def roofline_probe(reps=50):
    """Fused even-parity half step (forward) of the 2D P7 128x96 raster problem with
    PROBE_BEAMS beams, timed with HIP events on the library stream."""
    from pnepma_amd import workloads
    from pnepma_amd.problem import solve_inputs
    from pnepma_amd.solver import DeviceProblem
    import numpy as np
    p, mat, stds, det = workloads.raster2d(n_steps=400, n_beams=PROBE_BEAMS)
    si = solve_inputs(p, mat, det)
    dp = DeviceProblem(p, si, adjoint=False)
    ms, by = dp.half_step_bench(0, reps)
    achieved = by / (ms * 1e-3) / 1e9
    dp.close()
    # the fused reverse pass (the largest kernel of the served stream) on the same batch
    dp = DeviceProblem(p, si, ckpt_every=20)
    ms_r, by_r = dp.rev_bench(reps)
    dp.close()
    achieved_r = by_r / (ms_r * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
            "kernel": f"half_step_kernel<7,2,0,fwd> (even parity, 2D P7 128x96, {PROBE_BEAMS} beams)",
            "kernel_ms": ms, "algorithmic_bytes": by,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --workload kernel, "
                              "calibrated on a 1 GiB copy (profiles/half_step_p7_2d_128x96_160beams.md)",
            "reverse": {"kernel": f"rev_kernel_cw<7> (fused reverse pass, 2D P7 128x96, {PROBE_BEAMS} beams)",
                        "achieved": achieved_r, "frac": achieved_r / HBM_PEAK_GBS, "kernel_ms": ms_r,
                        "algorithmic_bytes": by_r, "bound": "latency (occupancy 3)"}}


def serve_run(preps, rate, budget, max_active, tape_fraction=None):
    """Submit every prepared request with Poisson arrivals of the given rate; returns
    (wall seconds from first arrival to last completion, latencies in seconds, stats)."""
    import numpy as np
    from pnepma_amd import serving
    rng = np.random.default_rng(7)
    gaps = rng.exponential(1.0 / rate, len(preps)) if rate > 0 else np.zeros(len(preps))
    arr = np.cumsum(gaps)
    arr -= arr[0]
    srv = serving.Server(budget, max_active=max_active)
    if tape_fraction is not None:
        srv.set_tape_limit(int(tape_fraction * budget))
    t0 = srv.clock() + 0.05
    tickets = [srv.submit(pr, arrival_s=t0 + a) for pr, a in zip(preps, arr)]
    progress(f"submitted {len(tickets)} requests in {srv.clock() - t0 + 0.05:.3f} s")
    srv.drain()
    tm = [srv.wait(t) for t in tickets]
    stats = srv.stats()
    stats["taped"] = sum(x.taped for x in tm)
    srv.close()
    lat = np.array([x.done_s - x.arrival_s for x in tm])
    wall = max(x.done_s for x in tm) - t0
    bad = sum(1 for x in tm if x.status != 0)
    if bad:
        raise RuntimeError(f"{bad} served requests failed")
    return wall, lat, stats


def latency_split(specs, lat, passes):
    """p50/p99 latency (ms) of the 1D and of the 2D requests of the stream."""
    import numpy as np
    is2d = np.tile([s.problem.gspec.axes[1].M > 1 for s in specs], passes)
    out = {}
    for name, sel in (("1d", ~is2d), ("2d", is2d)):
        if sel.any():
            out[name] = {"n": int(sel.sum()), "p50": float(np.percentile(lat[sel], 50) * 1e3),
                         "p99": float(np.percentile(lat[sel], 99) * 1e3)}
    return out


def sequential_run(specs, preps, max_seconds=20.0, max_requests=64):
    """The plain loop: one request at a time with the same kernels, no batching, no graph."""
    import numpy as np
    from pnepma_amd.solver import DeviceProblem
    times, outs = [], []
    t_all = time.perf_counter()
    for spec, pr in zip(specs, preps):
        K = max(1, math.ceil(math.sqrt(spec.problem.n_steps)))
        dp = DeviceProblem(spec.problem, pr.si, ckpt_every=K)
        t = time.perf_counter()
        I, _, _ = dp.forward_adjoint(spec.I_bar[None, :])
        times.append(time.perf_counter() - t)
        outs.append(I[0])
        dp.close()
        if len(times) >= max_requests or time.perf_counter() - t_all > max_seconds:
            break
    return np.array(times), outs


def request_cost(problem):
    """Algorithmic bytes of one forward+adjoint request (one beam), used only to
    weigh requests of different size against each other: per energy step the
    two forward sweeps (even half step, fused odd pair) read both parities and
    write the updated one, and the reverse reads u and lambda once more."""
    from pnepma_amd.grid import class_moments, odd_in
    cm = class_moments(problem.N)
    M = [a.M for a in problem.gspec.axes]
    dims = sum(m > 1 for m in M)
    vol = {0: 0.0, 1: 0.0}
    for c in problem.gspec.active_classes():
        pts = 1
        for a in range(dims):
            pts *= M[a] - 1 if odd_in(c, a) else M[a]
        vol[bin(c).count("1") % 2] += pts * len(cm[c]) * 8.0
    fwd = 2 * vol[0] + vol[1] + 2 * vol[1] + vol[0]
    rev = 2 * (vol[0] + vol[1]) + vol[1] + 2 * vol[0] + vol[1]
    return problem.n_steps * (fwd + rev)


def cpu_sample(specs, preps, budget_s=24.0, candidates=256):
    """The C oracle (single core) on a bounded sample of the stream, per stratum
    (1D requests get a fifth of the budget, 2D/3D the rest, as the latter carry
    nearly all the bytes).  Within a stratum requests are taken in stream order
    from the head, skipping any whose predicted time (its cost times the
    seconds per cost unit measured so far) would overrun the stratum's budget;
    the first one taken is the first of at most the stratum's median cost.
    The whole stream's oracle time is then predicted from each request's cost
    and its stratum's measured seconds per cost unit.  Returns the CPU
    throughput in solves/s, the sample size, its seconds and the sampled
    request indices."""
    import numpy as np
    from oracle import oracle
    cost = np.array([request_cost(s.problem) for s in specs])
    multi = np.array([s.problem.gspec.axes[1].M > 1 for s in specs])
    head = np.arange(min(candidates, len(specs)))
    sec_per_cost, n_all, el_all, taken = {}, 0, 0.0, []
    for stratum, share in ((False, 0.2), (True, 0.8)):
        idx = head[multi[head] == stratum]
        if idx.size == 0:
            continue
        b, med = share * budget_s, float(np.median(cost[idx]))
        el, done, n = 0.0, 0.0, 0
        for i in idx:
            if n == 0 and cost[i] > med:
                continue
            if n > 0 and el + cost[i] * el / done > b:
                continue
            t = time.perf_counter()
            oracle.forward_adjoint(specs[i].problem, preps[i].si, specs[i].I_bar)
            el += time.perf_counter() - t
            done += cost[i]
            n += 1
            taken.append(int(i))
            if el > 0.75 * b:
                break
        sec_per_cost[stratum] = el / done
        n_all += n
        el_all += el
    if not sec_per_cost:
        return 0.0, 0, 0.0, taken
    spc = np.array([sec_per_cost.get(bool(m), next(iter(sec_per_cost.values()))) for m in multi])
    return len(specs) / float(np.sum(cost * spc)), n_all, el_all, taken


_CPU_SAMPLE = None


def _run_sample(_):
    specs, preps, idx = _CPU_SAMPLE
    from oracle import oracle
    t = time.perf_counter()
    for i in idx:
        oracle.forward_adjoint(specs[i].problem, preps[i].si, specs[i].I_bar)
    return time.perf_counter() - t


def cpu_all_cores(specs, preps, idx, single_s, workers):
    """The same oracle sample run as `workers` concurrent copies, one process
    each (before the GPU is touched).  Every copy takes longer than the one-core
    run when the cores contend for memory; the all-core throughput is the
    one-core throughput times workers over that slowdown.  Returns (slowdown,
    wall seconds)."""
    import multiprocessing as mp
    global _CPU_SAMPLE
    _CPU_SAMPLE = (specs, preps, idx)
    t = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        times = pool.map(_run_sample, range(workers), chunksize=1)
    wall = time.perf_counter() - t
    return sum(times) / len(times) / single_s, wall


def progress(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main_serve(args):
    import numpy as np
    specs, preps = prepare_stream(args.requests)
    progress(f"prepared {len(preps)} requests")
    # CPU baselines first, while no process has touched the GPU
    cpu_rate, n_cpu, el_cpu, cpu_idx = cpu_sample(specs, preps) if not args.no_cpu else (0.0, 0, 0.0, [])
    progress(f"cpu sample: {n_cpu} requests in {el_cpu:.1f} s")
    workers = max(1, min(16, os.cpu_count() or 1))
    if n_cpu:
        slowdown, wall_all = cpu_all_cores(specs, preps, cpu_idx, el_cpu, workers)
        progress(f"cpu sample x{workers} processes: {wall_all:.1f} s, slowdown per copy {slowdown:.2f}")
    import torch
    assert torch.cuda.is_available()
    torch.cuda.init()
    torch.zeros(1, device="cuda:0")
    from pnepma_amd import serving
    free, total = torch.cuda.mem_get_info()
    budget = int(min(free * 0.9, 270e9))
    max_active = 4096
    # capacity probe on the head of the stream, all arriving at once; its tail
    # is poorly batched, so it underestimates the sustained capacity
    head = preps[:min(len(preps), 512)]
    w0, _, _ = serve_run(head, 0.0, budget, max_active, args.tape_fraction)
    capacity = len(head) / w0
    progress(f"capacity probe: {len(head)} requests in {w0:.2f} s")
    # the timed passes offer more than the server sustains, so it stays saturated
    # and `value` is its throughput, not the arrival rate
    rate = 1.5 * capacity
    for _ in range(max(0, args.warmup - 1)):
        serve_run(preps[:64], rate, budget, max_active, args.tape_fraction)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    walls, lats, stats = [], [], None
    for _ in range(args.steps):
        w, lat, stats = serve_run(preps, rate, budget, max_active, args.tape_fraction)
        walls.append(w)
        lats.append(lat)
        progress(f"served pass: {len(preps)} requests in {w:.2f} s")
    torch.cuda.synchronize()
    total_wall = time.perf_counter() - t0
    lat_over = np.concatenate(lats)
    served = len(preps) * args.steps / sum(walls)
    # latency at a load the server sustains: the same stream at 80 % of the
    # measured throughput (untimed)
    rate_lat = 0.8 * served
    w_lat, lat, stats_lat = serve_run(preps, rate_lat, budget, max_active, args.tape_fraction)
    progress(f"latency pass at {rate_lat:.1f}/s: {len(preps)} requests in {w_lat:.2f} s, "
             f"p50 {np.percentile(lat, 50) * 1e3:.0f} ms p99 {np.percentile(lat, 99) * 1e3:.0f} ms")
    # results of the served pass vs the same requests alone (plain loop)
    seq_t, seq_I = sequential_run(specs, preps)
    seq_rate = len(seq_t) / seq_t.sum()
    progress(f"plain loop: {len(seq_t)} requests, {seq_rate:.2f} solves/s")
    # the plain-loop sample is the head of the stream: compare with the served results
    alone_err = max(float(np.max(np.abs(a - pr.I_out)) / np.max(np.abs(a))) for a, pr in zip(seq_I, preps))
    roof = roofline_probe()
    out = {
        "metric": METRIC, "value": served, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * total_wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serving stream of {len(preps)} forward+gradient requests: PN 3/5/7/9, "
                               "1D 100-400 cells or 2D 64^2-192^2, 10-25 keV, 2-4 elements, CFL-0.5 step counts, "
                               "Poisson arrivals at 1.5x the all-at-once capacity of its first 512 requests",
                   "requests": len(preps), "arrival_rate_per_s": rate, "hbm_budget_bytes": budget,
                   "tape_limit": "0.95 of the budget (server default)" if args.tape_fraction is None
                   else f"{args.tape_fraction} of the budget",
                   "mean_energy_steps": float(np.mean([s.problem.n_steps for s in specs]))},
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "latency_load": {"arrival_rate_per_s": rate_lat, "note": "p50/p99 from an untimed pass of the same stream "
                         "at 0.8x the measured throughput", "taped": stats_lat["taped"]},
        "latency_ms_by_grid": latency_split(specs, lat, 1),
        "saturated_pass_latency_ms": {"p50": float(np.percentile(lat_over, 50) * 1e3),
                                      "p99": float(np.percentile(lat_over, 99) * 1e3)},
        "plain_gpu_loop": {"value": seq_rate, "unit": "solves/s", "sample": f"first {len(seq_t)} requests of "
                           "the stream, one at a time, same kernels, no batching, no graph",
                           "mean_ms": float(seq_t.mean() * 1e3)},
        "speedup_vs_plain_gpu_loop": served / seq_rate,
        "served_vs_alone_max_rel_err": alone_err,
        "server_stats": stats,
        "roofline": roof,
    }
    if n_cpu:
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"C oracle forward+adjoint of {n_cpu} 1D and 2D requests from the head of "
                                         f"the stream ({el_cpu:.1f} s on one core); the whole stream's oracle time "
                                         "is predicted from each request's algorithmic bytes (bench.request_cost) "
                                         "and its dimension's measured seconds per byte"}
        out["speedup_vs_cpu"] = served / cpu_rate
        all_rate = cpu_rate * workers / slowdown
        out["cpu_baseline_all_cores"] = {"value": all_rate, "unit": "solves/s", "cores": workers, "kind": "port",
                                         "sample": f"the same sample as {workers} concurrent processes "
                                                   f"({wall_all:.1f} s wall; each copy {slowdown:.2f}x its "
                                                   "one-core time); rate = one-core rate x cores / slowdown"}
        out["speedup_vs_cpu_all_cores"] = served / all_rate
    print(json.dumps(out), flush=True)


def main_raster(args):
    import numpy as np
    import torch
    from pnepma_amd import workloads
    from pnepma_amd.problem import solve_inputs
    from pnepma_amd.solver import DeviceProblem
    assert torch.cuda.is_available()
    torch.cuda.init()
    p, mat, stds, det = workloads.raster2d(n_steps=400)
    si = solve_inputs(p, mat, det)
    dp = DeviceProblem(p, si, ckpt_every=16)
    ib = np.random.default_rng(2).uniform(0.5, 1.5, (p.n_beams, len(p.x_rays)))
    I, rb, gen = dp.forward_adjoint(ib)
    for _ in range(max(0, args.warmup - 1)):
        dp.forward_adjoint_timed(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dp.forward_adjoint_timed(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = dp.forward_timed(3, True)
    dp.close()
    out = {
        "metric": METRIC, "value": p.n_beams * args.steps / wall, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "raster2d_grad: 2D P7 128x96, 32 beams, Fe/Cr/Ni, 400 energy steps (CFL 0.45), "
                               "checkpoints every 16, one forward+adjoint of the 32-beam batch per step",
                   "n_beams": p.n_beams, "n_steps": p.n_steps},
        "forward_ms_per_batch": fwd_ms,
        "roofline": roofline_probe(),
    }
    if not args.no_cpu:
        from oracle import oracle
        I_ref = oracle.forward(p, si, beam=0)
        out["parity_beam0_intensity_rel_err"] = float(np.max(np.abs(I[0] - I_ref)) / np.max(np.abs(I_ref)))
    print(json.dumps(out), flush=True)


def main_raster3d(args):
    """The 3D workload (P5 128x128x96, Al/Si inclusion) at its full shape for one
    beam, on a sample of its energy grid: S steps of the CFL-0.5 step length from
    the top of the energy range, forward+adjoint with checkpoints every 8.  The
    whole solve (all CFL-0.5 steps, 64 beams) is extrapolated from the per-step
    time, since one beam alone is minutes of GPU time."""
    import numpy as np
    import torch
    from pnepma_amd import workloads, serving
    from pnepma_amd.problem import solve_inputs
    from pnepma_amd.solver import DeviceProblem
    assert torch.cuda.is_available()
    torch.cuda.init()
    S = 16
    full, mat, stds, det = workloads.inclusion3d(n_steps=16, n_beams=1)
    n_cfl = serving.cfl_steps(full.N, full.gspec, mat, full.elements, full.E_initial, full.E_cutoff)
    E1 = full.E_initial - (full.E_initial - full.E_cutoff) * S / n_cfl
    p, mat, stds, det = workloads.inclusion3d(n_steps=S, n_beams=1, E0=full.E_initial, E1=E1)
    si = solve_inputs(p, mat, det)
    progress(f"3D problem prepared: {p.gspec.M}, {S} of {n_cfl} steps")
    dp = DeviceProblem(p, si, ckpt_every=8)
    ib = np.ones((1, len(p.x_rays)))
    dp.forward_adjoint(ib)
    torch.cuda.synchronize()
    ms_fa = min(dp.forward_adjoint_timed(1) for _ in range(max(1, args.steps)))
    ms_f = dp.forward_timed(2, True)
    dp.close()
    # state bytes per step: forward (even half step, fused odd pair), recompute, reverse
    per_step = request_cost(p) / S
    from pnepma_amd.grid import class_moments, odd_in
    cm, M = class_moments(p.N), [a.M for a in p.gspec.axes]
    vol = {0: 0.0, 1: 0.0}
    for c in p.gspec.active_classes():
        pts = 1
        for a in range(3):
            pts *= M[a] - 1 if odd_in(c, a) else M[a]
        vol[bin(c).count("1") % 2] += pts * len(cm[c]) * 8.0
    fwd_part = 3 * (vol[0] + vol[1])
    bytes_step = per_step + fwd_part  # checkpointed: the forward runs twice
    out = {
        "metric": METRIC, "value": 1e3 * S / ms_fa, "unit": "energy steps/s (forward+adjoint, one beam)",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": ms_fa, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D P5 128x128x96, Al/Si with an ellipsoidal inclusion, one beam, "
                               f"{S} of the {n_cfl} CFL-0.5 energy steps of 25 -> 1 keV, checkpoints every 8"},
        "ms_per_energy_step_fwd_adj": ms_fa / S, "ms_per_energy_step_fwd": ms_f / S,
        "state_bytes_per_energy_step_fwd_adj": bytes_step,
        "achieved_state_GBps": bytes_step / (ms_fa / S * 1e-3) / 1e9,
        "extrapolated_s_per_beam_fwd_adj": ms_fa / S * n_cfl / 1e3,
        "extrapolated_s_64_beams_fwd_adj": 64 * ms_fa / S * n_cfl / 1e3,
    }
    print(json.dumps(out), flush=True)


def main_kernel(args):
    """Only the dominant kernel (and a plain 256 MiB device copy for counter
    calibration): used under rocprofv3 to read per-dispatch HBM counters."""
    import torch
    assert torch.cuda.is_available()
    x = torch.empty(128 * 1024 * 1024, dtype=torch.float64, device="cuda:0").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    roof = roofline_probe()
    out = {"metric": METRIC, "value": None, "unit": "solves/s", "n_gpus": 1, "steps": 0, "warmup": 0,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "config": {"workload": f"kernel probe: fused even half step and fused reverse pass of "
                                                      f"2D P7 128x96, {PROBE_BEAMS} beams, plus 5 copies of 1 GiB"},
           "roofline": roof, "copy_bytes": 2 * x.numel() * 8}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="serve", choices=["serve", "raster2d_grad", "raster3d_step", "kernel"])
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tape-fraction", type=float, default=None,
                    help="share of the HBM budget the server may hold in taped requests (0: always checkpoint)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X "
              "(how the path would shard over GPUs is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serve":
        main_serve(args)
    elif args.workload == "kernel":
        main_kernel(args)
    elif args.workload == "raster3d_step":
        main_raster3d(args)
    else:
        main_raster(args)


if __name__ == "__main__":
    main()
