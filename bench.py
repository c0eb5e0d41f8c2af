"""Benchmark of the PN-CSD forward+adjoint hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload serve|raster2d_grad|raster3d_step|raster3d_full|kernel]

serve (default): the mixed serving stream -- R = 4096 forward+gradient requests
(PN order 3-9, 1D 100-400 cells or 2D 64^2-192^2, 10-25 keV, 2-4 elements,
stable step counts) with Poisson arrivals, served by the continuous-batching
server under an HBM budget.  The timed region is ONE continuous pass of the
whole stream; a "step" is one window of R/K requests of it, so ms_per_step is
the pass's wall time over K.  The W warm-up windows are the head of the stream
served untimed (all at once; this also measures the capacity from which the
arrival rate of the timed pass is set, 1.5x, so the server stays saturated and
`value` is its own throughput).  Request inputs are prepared in host memory
before the timed region; their preparation time is reported beside `value`.
Reported with it: p50/p99 latency from an untimed pass of the first half of
the stream at 0.8x the measured throughput, the plain one-request-at-a-time
GPU loop (same kernels, taped like the server, no batching), a stratified
sample of 64 served results checked against the C oracle (intensities 1e-10,
density gradients 1e-8), and the C oracle's CPU throughput on the box's host
cores.

raster2d_grad: 2D P7 128x96, 32 beams, forward+adjoint of the batch per step.
raster3d_step: 3D P5 128x128x96, one beam, a sample of its energy steps.
raster3d_full: the same workload whole: 64 beams, every energy step (minutes).

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
    return serving.prepare(spec, want_rho_bar=True)


WORKERS = 16  # the host cores of one GPU's share of the box


def host_workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(WORKERS, n))


def prepare_stream(n, seed=3, grids="all"):
    """Draw the request stream and build its requests as the library takes them
    (cell fractions, element tables, absorption inputs: serving.SpecRequest).
    grids "1d" / "2d" keeps only the stream's requests of that dimension (in
    stream order), to measure what each kind takes of the GPU.  Returns the
    specs, the requests and the mean seconds to build one."""
    from pnepma_amd import serving, physics
    for e in physics.ELEMENTS.values():
        physics.transport_coefficients(e, 1000.0, 9)
    specs = serving.serving_requests(n, seed=seed)
    if grids != "all":
        specs = [sp for sp in specs if (sp.problem.gspec.axes[1].M > 1) == (grids == "2d")]
        n = len(specs)
    t = time.perf_counter()
    reqs = [serving.SpecRequest(sp) for sp in specs]
    return specs, reqs, (time.perf_counter() - t) / n


def prepare_host(specs, idx):
    """Host-side inputs (densities, absorption, records: serving.prepare) of the
    requests idx, for the oracle (process pool, before the GPU is touched)."""
    import multiprocessing as mp
    with mp.get_context("fork").Pool(host_workers()) as pool:
        out = pool.map(_prep_one, [specs[i] for i in idx], chunksize=1)
    return dict(zip(idx, out))


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


def serve_run(srv, preps, rate):
    """Submit every request (serving.SpecRequest: the library builds its records,
    beam data, densities and absorption) to the running server with Poisson
    arrivals of the given rate (0: all at once) and wait for all of them;
    returns (wall seconds from the first arrival to the last completion,
    latencies in seconds, requests admitted taped, seconds spent submitting)."""
    import numpy as np
    rng = np.random.default_rng(7)
    gaps = rng.exponential(1.0 / rate, len(preps)) if rate > 0 else np.zeros(len(preps))
    arr = np.cumsum(gaps)
    arr -= arr[0]
    t0 = srv.clock() + 0.05
    ts = time.perf_counter()
    tickets = [srv.submit_spec(rq, arrival_s=t0 + a) for rq, a in zip(preps, arr)]
    submit_s = time.perf_counter() - ts
    progress(f"submitted {len(tickets)} requests in {submit_s:.3f} s")
    srv.drain()
    tm = [srv.wait(t) for t in tickets]
    lat = np.array([x.done_s - x.arrival_s for x in tm])
    wall = max(x.done_s for x in tm) - t0
    bad = sum(1 for x in tm if x.status != 0)
    if bad:
        raise RuntimeError(f"{bad} served requests failed")
    return wall, lat, sum(x.taped for x in tm), submit_s


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


def sequential_run(specs, max_seconds=15.0, max_requests=64):
    """The plain loop: one request at a time with the same kernels, no batching.
    Each request is set up like a served one (densities and absorption on the
    device from its cell fractions) and taped like one (its whole forward march
    kept, no recompute: checkpoint interval = step count), so the ratio to the
    served rate measures batching alone.  Timed per request: forward_adjoint
    and the material gradient, on a device problem whose buffers are allocated
    beforehand."""
    import numpy as np
    from pnepma_amd.solver import DeviceProblem
    times, outs = [], []
    t_all = time.perf_counter()
    for spec in specs:
        p = spec.problem
        dp = DeviceProblem(p, None, ckpt_every=p.n_steps)
        dp.set_cell_fractions(spec.material, None)
        dp.set_absorption(spec.material, spec.detector)
        dp.reserve_adjoint()  # the server's arena is allocated once too
        t = time.perf_counter()
        I, _, _ = dp.forward_adjoint(spec.I_bar[None, :], want_rho_bar=False)
        dp.material_gradient()
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


def check_sample(specs, per_stratum=8):
    """The served results checked against the oracle: the per_stratum cheapest
    requests (algorithmic bytes) of every (PN order, dimension) stratum."""
    import numpy as np
    cost = np.array([request_cost(s.problem) for s in specs])
    key = [(s.problem.N, s.problem.gspec.axes[1].M > 1) for s in specs]
    out = []
    for k in sorted(set(key)):
        idx = [i for i in range(len(specs)) if key[i] == k]
        out += sorted(idx, key=lambda i: cost[i])[:per_stratum]
    return out


_CPU = None


def _oracle_one(i):
    """Oracle forward+adjoint of request i (in a pool worker):
    (i, I, rho_bar, ac_bar, seconds, noa_bar)."""
    specs, preps = _CPU
    from oracle import oracle
    t = time.perf_counter()
    I, rb, acb, noab = oracle.forward_adjoint(specs[i].problem, preps[i].si, specs[i].I_bar)
    return i, I, rb, acb, time.perf_counter() - t, noab


def _median_time(fn, reps=5):
    import numpy as np
    fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_phase(specs, preps, check_idx, tmpdir):
    """Everything that runs the C oracle, before the GPU is touched:
      * the 1D slab (P5, 200 cells, Cu/Ni, 300 stable steps) forward on one core:
        median latency of one solve and of its k-ratios (sample + two standards);
      * one-core seconds per algorithmic byte of the stream's 1D and 2D requests,
        on a subset of the check sample run alone;
      * the whole check sample as concurrent solves on every host core of the
        share (one process each), whose outputs also check the served results.
    The stream's CPU throughput is its request count over the predicted time of
    all its requests (each request's bytes times its dimension's measured
    seconds per byte), on one core and on all of them."""
    import multiprocessing as mp
    import numpy as np
    from oracle import oracle
    from pnepma_amd import workloads
    from pnepma_amd.problem import solve_inputs
    global _CPU
    oracle.use_native(tmpdir)
    out = {"model": host_cpu_model(), "nproc": os.cpu_count(), "workers": host_workers()}
    p, mat, stds, det = workloads.slab1d(n_steps=300)
    si = solve_inputs(p, mat, det)
    sis = [solve_inputs(p, m, det) for m in stds]

    out["slab1d_forward_ms"] = 1e3 * _median_time(lambda: oracle.forward(p, si))
    out["slab1d_kratios_ms"] = 1e3 * _median_time(
        lambda: [oracle.forward(p, si)] + [oracle.forward(p, x) for x in sis])
    cost = np.array([request_cost(s.problem) for s in specs])
    is2d = np.array([s.problem.gspec.axes[1].M > 1 for s in specs])
    # one core alone: the cheapest 1D request of every order and the two
    # cheapest 2D requests of every order (about 15 s)
    alone = {False: [0.0, 0.0], True: [0.0, 0.0]}
    picked = {}
    for i in check_idx:
        N, d2 = specs[i].problem.N, bool(is2d[i])
        if picked.get((N, d2), 0) >= (2 if d2 else 1):
            continue
        picked[(N, d2)] = picked.get((N, d2), 0) + 1
        t = time.perf_counter()
        oracle.forward_adjoint(specs[i].problem, preps[i].si, specs[i].I_bar)
        alone[d2][0] += time.perf_counter() - t
        alone[d2][1] += cost[i]
    spc1 = {k: v[0] / v[1] for k, v in alone.items() if v[1] > 0}
    # every core: the check sample, longest first
    _CPU = (specs, preps)
    order = sorted(check_idx, key=lambda i: -cost[i])
    t = time.perf_counter()
    with mp.get_context("fork").Pool(out["workers"]) as pool:
        res = pool.map(_oracle_one, order, chunksize=1)
    wall_all = time.perf_counter() - t
    _CPU = None
    busy = {False: [0.0, 0.0], True: [0.0, 0.0]}
    for i, _, _, _, el, _ in res:
        busy[bool(is2d[i])][0] += el
        busy[bool(is2d[i])][1] += cost[i]
    spcn = {k: v[0] / v[1] for k, v in busy.items() if v[1] > 0}
    t1 = float(sum(cost[i] * spc1[bool(is2d[i])] for i in range(len(specs))))
    tn = float(sum(cost[i] * spcn[bool(is2d[i])] for i in range(len(specs))))
    out["one_core_rate"] = len(specs) / t1
    out["all_core_rate"] = len(specs) / tn * out["workers"]
    out["one_core_sample"] = {"requests": sum(picked.values()), "seconds": float(sum(v[0] for v in alone.values()))}
    out["all_core_sample"] = {"requests": len(res), "wall_s": wall_all,
                              "busy_s": float(sum(v[0] for v in busy.values()))}
    out["contention"] = {("2d" if k else "1d"): spcn[k] / spc1[k] for k in spc1 if k in spcn}
    return out, {r[0]: r for r in res}


def served_vs_oracle(specs, reqs, preps, ref):
    """Largest relative differences of the served results from the oracle's:
    intensities, density gradients of every updated class, absorption-factor
    gradients (from the generation sums) and the whole material gradient."""
    import numpy as np
    from oracle import oracle
    from pnepma_amd import physics

    def rel(a, b):
        sc = float(np.max(np.abs(b)))
        return float(np.max(np.abs(a - b))) / sc if sc > 0 else float(np.max(np.abs(a)))
    eI = erb = eac = eg = 0.0
    for i, (_, I, rb, acb, _, noab) in ref.items():
        rq, pr, sp = reqs[i], preps[i], specs[i]
        p = sp.problem
        eI = max(eI, rel(rq.I_out, I))
        for c in p.gspec.active_classes():
            if rb[c].size:
                erb = max(erb, rel(rq.rho_bar[c], rb[c]))
        eac = max(eac, rel(pr.si.noa[:, pr.si.el_idx] * rq.gen, acb))
        mac = np.array([[physics.mass_absorption(x, e) for e in sp.material.elements] for x in p.x_rays])
        g_ref = oracle.material_gradient(p, sp.material, sp.detector, mac, rb, acb, noab)
        eg = max(eg, rel(rq.material_gradient(), g_ref))
    return {"requests": len(ref), "max_rel_err_intensity": eI, "max_rel_err_rho_bar": erb,
            "max_rel_err_ac_bar": eac, "max_rel_err_material_gradient": eg,
            "tolerance": {"intensity": 1e-10, "gradients": 1e-8},
            "pass": bool(eI <= 1e-10 and erb <= 1e-8 and eac <= 1e-8 and eg <= 1e-8),
            "sample": "the 8 cheapest requests of each (PN order 3/5/7/9, 1D/2D) stratum of the stream"}


def progress(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main_serve(args):
    import tempfile
    import numpy as np
    t_start = time.perf_counter()
    K, W = max(1, args.steps), max(0, args.warmup)
    specs, reqs, build_s = prepare_stream(args.requests, grids=args.grids)
    R = len(specs)
    progress(f"drew {R} requests ({build_s * 1e3:.2f} ms each to build as the library takes them)")
    check_idx = check_sample(specs) if not args.no_cpu else []
    for i in check_idx:
        reqs[i] = type(reqs[i])(specs[i], want_gradients=True)
    cpu, ref, preps = None, {}, {}
    if not args.no_cpu:
        # host inputs of the check sample for the oracle, and the C oracle
        # itself, before any process touches the GPU
        preps = prepare_host(specs, check_idx)
        with tempfile.TemporaryDirectory() as tmp:
            cpu, ref = cpu_phase(specs, preps, check_idx, tmp)
        progress(f"cpu phase: one core {cpu['one_core_rate']:.4f} solves/s, {cpu['workers']} cores "
                 f"{cpu['all_core_rate']:.3f} solves/s; check sample {cpu['all_core_sample']['wall_s']:.1f} s wall")
    # torch starts the GPU runtime before this process loads libpnx.so: torch
    # ships its own HIP runtime, and a process that loaded the system one first
    # saw no device after torch had started
    from pnepma_amd import _lib
    assert _lib._lib is None, "libpnx.so was loaded before the GPU phase"
    import torch
    assert torch.cuda.is_available()
    torch.cuda.init()
    torch.zeros(1, device="cuda:0")
    from pnepma_amd import serving
    free, total = torch.cuda.mem_get_info()
    budget = int(min(free * 0.9, 270e9))
    # one server for every pass: its arena is allocated once, outside the timed region
    srv = serving.Server(budget, max_active=4096)
    if args.tape_fraction is not None:
        srv.set_tape_limit(int(args.tape_fraction * budget))
    # warm-up: the head of the stream, W windows of R/K requests (at least 512)
    # all at once; its rate also sets the arrival rate of the timed pass
    n_head = min(R, max(512, W * R // K))
    w0, _, _, _ = serve_run(srv, reqs[:n_head], 0.0)
    capacity = n_head / w0
    rate = 1.5 * capacity
    progress(f"warm-up / capacity probe: {n_head} requests in {w0:.2f} s")
    st0 = srv.stats()
    # timed: one continuous pass of the whole stream (K windows of R/K
    # requests); submitting a request runs its C++ set-up in this thread
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w, lat_sat, taped, submit_s = serve_run(srv, reqs, rate)
    torch.cuda.synchronize()
    total_wall = time.perf_counter() - t0
    st1 = srv.stats()
    stats = {"ticks": st1["ticks"] - st0["ticks"], "completed": st1["completed"] - st0["completed"],
             "peak_active": st1["peak_active"], "peak_bytes": st1["peak_bytes"], "taped": taped}
    served = R / total_wall
    progress(f"timed pass: {R} requests in {total_wall:.2f} s ({served:.1f} solves/s; server {w:.2f} s)")
    check = served_vs_oracle(specs, reqs, preps, ref) if ref else None
    if check:
        progress(f"served vs oracle ({check['requests']} requests): I {check['max_rel_err_intensity']:.2e} "
                 f"rho_bar {check['max_rel_err_rho_bar']:.2e} ac_bar {check['max_rel_err_ac_bar']:.2e} "
                 f"material gradient {check['max_rel_err_material_gradient']:.2e}")
    # latency at a load the server sustains: the first half of the stream at
    # 80 % of the measured throughput (untimed)
    n_lat = R // 2
    rate_lat = 0.8 * served
    w_lat, lat, taped_lat, _ = serve_run(srv, reqs[:n_lat], rate_lat)
    srv.close()
    progress(f"latency pass at {rate_lat:.1f}/s: {n_lat} requests in {w_lat:.2f} s, "
             f"p50 {np.percentile(lat, 50) * 1e3:.0f} ms p99 {np.percentile(lat, 99) * 1e3:.0f} ms")
    seq_t, seq_I = sequential_run(specs)
    seq_rate = len(seq_t) / seq_t.sum()
    progress(f"plain loop: {len(seq_t)} requests, {seq_rate:.2f} solves/s")
    alone_err = max(float(np.max(np.abs(a - rq.I_out)) / np.max(np.abs(a))) for a, rq in zip(seq_I, reqs))
    roof = roofline_probe()
    progress(f"roofline probes done; {time.perf_counter() - t_start:.0f} s since start")
    out = {
        "metric": METRIC, "value": served, "unit": "solves/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": 1e3 * total_wall / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serving stream of {R} forward+gradient requests: PN 3/5/7/9, "
                               + ("" if args.grids == "all" else
                                  f"only the {args.grids.upper()} requests of a {args.requests}-request stream: ") +
                               "1D 100-400 cells or 2D 64^2-192^2, 10-25 keV, 2-4 elements, stable (CFL 0.5) step "
                               "counts, Poisson arrivals at 1.5x the all-at-once capacity of the stream's head; "
                               "each request is submitted as cell fractions and returns its intensities and its "
                               "material gradient; one timed pass, a step = one window of "
                               f"{R // K} requests",
                   "requests": R, "arrival_rate_per_s": rate, "hbm_budget_bytes": budget,
                   "tape_limit": "0.95 of the budget (server default)" if args.tape_fraction is None
                   else f"{args.tape_fraction} of the budget",
                   "mean_energy_steps": float(np.mean([s.problem.n_steps for s in specs]))},
        "server_wall_s": w,
        "host_prep": {"submit_ms_per_request": 1e3 * submit_s / R, "build_ms_per_request": 1e3 * build_s,
                      "note": "submit = the library's C++ set-up (records, beam moments and profiles) and "
                              "staging, inside the timed region; build = the Python arrays of a request "
                              "(cell fractions, element tables), before it; densities and absorption run on "
                              "the device at admission"},
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "latency_load": {"arrival_rate_per_s": rate_lat, "requests": n_lat, "taped": taped_lat,
                         "note": "p50/p99 from an untimed pass of the first half of the stream at 0.8x the "
                                 "measured throughput"},
        "latency_ms_by_grid": latency_split(specs[:n_lat], lat, 1),
        "saturated_pass_latency_ms": {"p50": float(np.percentile(lat_sat, 50) * 1e3),
                                      "p99": float(np.percentile(lat_sat, 99) * 1e3)},
        "plain_gpu_loop": {"value": seq_rate, "unit": "solves/s", "taped": True,
                           "sample": f"first {len(seq_t)} requests of the stream, one at a time, same kernels, "
                                     "taped like the server, material on the device, no batching",
                           "mean_ms": float(seq_t.mean() * 1e3)},
        "speedup_vs_plain_gpu_loop": served / seq_rate,
        "served_vs_alone_max_rel_err": alone_err,
        "served_vs_oracle": check,
        "server_stats": stats,
        "roofline": roof,
    }
    if cpu:
        out["cpu_baseline"] = {
            "value": cpu["one_core_rate"], "unit": "solves/s", "cores": 1, "kind": "port",
            "host": {"model": cpu["model"], "nproc": cpu["nproc"]},
            "sample": f"C oracle (gcc -O3 -march=native) forward+adjoint of {cpu['one_core_sample']['requests']} "
                      f"requests of the stream alone on one core ({cpu['one_core_sample']['seconds']:.1f} s); the "
                      "stream's time is predicted from each request's algorithmic bytes (request_cost) and its "
                      "dimension's measured seconds per byte"}
        out["speedup_vs_cpu"] = served / cpu["one_core_rate"]
        out["cpu_baseline_all_cores"] = {
            "value": cpu["all_core_rate"], "unit": "solves/s", "cores": cpu["workers"], "kind": "port",
            "sample": f"the {cpu['all_core_sample']['requests']} requests of the oracle check as concurrent "
                      f"solves, one process per core ({cpu['all_core_sample']['wall_s']:.1f} s wall); "
                      "seconds per byte under contention: " +
                      ", ".join(f"{k} x{v:.2f}" for k, v in cpu["contention"].items()),
        }
        out["speedup_vs_cpu_all_cores"] = served / cpu["all_core_rate"]
        out["cpu_slab1d_p5_200cells"] = {"forward_ms": cpu["slab1d_forward_ms"],
                                         "kratios_ms": cpu["slab1d_kratios_ms"], "cores": 1,
                                         "note": "C oracle, 300 stable steps, median of 5 after one warm-up; "
                                                 "k-ratios = the sample and the two pure standards"}
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


def main_raster3d_full(args):
    """The whole 3D workload: P5 128x128x96 Al/Si inclusion, the 8 x 8 beam
    raster, all CFL-0.5 energy steps from 25 to 1 keV, forward + reference
    adjoint with checkpoints every 32, material and absorption set on the
    device from the cell fractions, the material gradient of every beam summed.
    Beams run in batches that fit the HBM; about five minutes of GPU time, so it
    is not part of the default run."""
    import copy
    import numpy as np
    import torch
    from pnepma_amd import workloads, serving
    from pnepma_amd.solver import DeviceProblem
    assert torch.cuda.is_available()
    torch.cuda.init()
    p0, mat, stds, det = workloads.inclusion3d(n_steps=8)
    n_cfl = serving.cfl_steps(p0.N, p0.gspec, mat, p0.elements, p0.E_initial, p0.E_cutoff)
    p, mat, stds, det = workloads.inclusion3d(n_steps=n_cfl, n_beams=64)

    def beams(idx):
        q = copy.copy(p)
        q._cache = {}
        q.beam_positions = [p.beam_positions[b] for b in idx]
        return q

    ib = np.ones(len(p.x_rays))
    probe = DeviceProblem(beams([0]), None, ckpt_every=32)
    probe.reserve_adjoint()
    per_beam = probe.device_bytes
    probe.close()
    free, total = torch.cuda.mem_get_info()
    batch = max(1, min(16, int(0.85 * free // per_beam)))
    progress(f"3D P5 {p.gspec.M}: {n_cfl} steps, 64 beams, checkpoints every 32; "
             f"{per_beam / 1e9:.1f} GB per beam, batches of {batch}")
    I_all, grad, t_gpu = [], None, 0.0
    t0 = time.perf_counter()
    for b0 in range(0, 64, batch):
        idx = list(range(b0, min(64, b0 + batch)))
        dp = DeviceProblem(beams(idx), None, ckpt_every=32)
        dp.set_cell_fractions(mat, None)
        dp.set_absorption(mat, det)
        dp.reserve_adjoint()
        torch.cuda.synchronize()
        t = time.perf_counter()
        I, _, _ = dp.forward_adjoint(ib, want_rho_bar=False)
        g = dp.material_gradient()
        t_gpu += time.perf_counter() - t
        dp.close()
        I_all.append(I)
        grad = g if grad is None else grad + g
        progress(f"beams {idx[0]}-{idx[-1]}: {time.perf_counter() - t:.1f} s")
    wall = time.perf_counter() - t0
    I_all = np.concatenate(I_all)
    ok = bool(np.all(np.isfinite(I_all)) and np.all(np.isfinite(grad)) and
              all(np.all(I_all[:, r] > 0) or np.all(I_all[:, r] < 0) for r in range(I_all.shape[1])))
    out = {
        "metric": METRIC, "value": 64 / t_gpu, "unit": "beams/s (forward+adjoint of the whole energy range)",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * t_gpu, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D P5 128x128x96, Al/Si with an ellipsoidal inclusion, 8x8 beam raster, "
                               f"all {n_cfl} CFL-0.5 energy steps of 25 -> 1 keV, checkpoints every 32, "
                               "material gradient summed over the beams"},
        "solve_s": t_gpu, "wall_s_with_setup": wall, "beams_per_batch": batch, "device_bytes_per_beam": per_beam,
        "intensities_finite_one_sign": ok,
        "intensity_range": [float(np.abs(I_all).min()), float(np.abs(I_all).max())],
        "gradient_norm": float(np.linalg.norm(grad)),
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
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="serve",
                    choices=["serve", "raster2d_grad", "raster3d_step", "raster3d_full", "kernel"])
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grids", default="all", choices=["all", "1d", "2d"],
                    help="serve only the stream's 1D or 2D requests (what each kind takes of the GPU)")
    ap.add_argument("--tape-fraction", type=float, default=None,
                    help="share of the HBM budget the server may hold in taped requests (0: always checkpoint)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the served path runs on a single MI355X "
              "(beams and requests are independent solves; over several GPUs each process would serve "
              "its own share with no exchange during the march)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serve":
        main_serve(args)
    elif args.workload == "kernel":
        main_kernel(args)
    elif args.workload == "raster3d_step":
        main_raster3d(args)
    elif args.workload == "raster3d_full":
        main_raster3d_full(args)
    else:
        main_raster(args)


if __name__ == "__main__":
    main()
