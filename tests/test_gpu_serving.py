"""GPU serving: every result of the continuous-batching server equals the same
request solved alone (pnx_forward_adjoint with one beam), and matches the
oracle, while requests of mixed PN order, dimension and size share the device
under an HBM budget that forces queueing."""
import math

import numpy as np
import pytest

from pnepma_amd import serving
from pnepma_amd.problem import solve_inputs
from pnepma_amd.solver import DeviceProblem

from .helpers import assert_physical, rel_err

pytestmark = pytest.mark.gpu


def _specs(n=10, seed=11):
    return serving.serving_requests(n, seed=seed, sizes_1d=(30, 70), sizes_2d=(10, 18), max_steps=40)


def _alone(spec):
    si = solve_inputs(spec.problem, spec.material, spec.detector)
    K = max(1, math.ceil(math.sqrt(spec.problem.n_steps)))
    dp = DeviceProblem(spec.problem, si, ckpt_every=K)
    I, rb, gen = dp.forward_adjoint(spec.I_bar[None, :])
    dp.close()
    return I[0], rb, gen


@pytest.mark.parametrize("tape", ["default", "none", "all"])
def test_served_equals_alone(tape):
    """Checkpointed (recompute per segment) and taped (every forward state kept)
    requests, alone and mixed under a budget that forces queueing."""
    specs = _specs()
    preps = [serving.prepare(s, want_rho_bar=True) for s in specs]
    need = max(serving.request_device_bytes(s.problem) for s in specs)
    budget = 3 * need if tape != "all" else 1 << 30
    srv = serving.Server(hbm_budget_bytes=budget, max_active=64)
    if tape == "none":
        srv.set_tape_limit(0)
    tickets = []
    for n, pr in enumerate(preps):
        tickets.append(srv.submit(pr, arrival_s=0.002 * n))
    taped = []
    for t in tickets:
        tm = srv.wait(t)
        assert tm.status == 0
        taped.append(tm.taped)
    st = srv.stats()
    print("server stats", st, "taped", taped, flush=True)
    if tape == "none":
        assert not any(taped)
    if tape == "all":
        assert all(taped)
    assert st["completed"] == len(specs)
    assert st["peak_bytes"] <= budget
    srv.close()
    for spec, pr in zip(specs, preps):
        I, rb, gen = _alone(spec)
        p = spec.problem
        print(f"N={p.N} M={p.gspec.M} steps={p.n_steps} I={pr.I_out}", flush=True)
        assert np.array_equal(pr.I_out, I) or rel_err(pr.I_out, I) < 1e-12
        for c in p.gspec.active_classes():
            assert rel_err(pr.rho_bar[c], rb[c]) < 1e-12
        assert rel_err(pr.gen, gen) < 1e-12


@pytest.mark.parametrize("taped", [True, False])
def test_served_matches_oracle(oracle_lib, taped):
    specs = _specs(n=4, seed=5)
    preps = [serving.prepare(s, want_rho_bar=True) for s in specs]
    srv = serving.Server(hbm_budget_bytes=1 << 30, max_active=16)
    if not taped:
        srv.set_tape_limit(0)
    for t in [srv.submit(pr) for pr in preps]:
        assert srv.wait(t).taped == int(taped)
    srv.close()
    for spec, pr in zip(specs, preps):
        I, rb, acb, noab = oracle_lib.forward_adjoint(spec.problem, pr.si, spec.I_bar)
        assert rel_err(pr.I_out, I) < 1e-10
        for c in spec.problem.gspec.active_classes():
            assert rel_err(pr.rho_bar[c], rb[c]) < 1e-8
        assert rel_err(pr.si.noa[:, pr.si.el_idx] * pr.gen, acb) < 1e-8


def _spec_3d(N, M, n_steps, seed):
    from pnepma_amd import workloads
    from pnepma_amd.grid import GridSpec
    from pnepma_amd.material import CellFractions, MassFractionMaterial
    from pnepma_amd.physics import XRay, element
    from pnepma_amd.problem import ForwardPNProblem
    rng = np.random.default_rng(seed)
    names = ["Fe", "Cr", "Ni"]
    els = [element(s) for s in names]
    g = GridSpec.make(M, (-1.5e-4, 0.0), (-0.5e-4, 0.5e-4), (-0.4e-4, 0.4e-4))
    p = ForwardPNProblem(N=N, gspec=g, elements=els, beam_energy=(15e3, workloads.SIGMA_E),
                         beam_positions=[(0.0, 0.0, 0.0)], beam_sigma=50 * workloads.NM,
                         beam_direction=((-1.0, 0.0, 0.0), workloads.KAPPA), E_initial=15.5e3, E_cutoff=5e3,
                         n_steps=n_steps, source_mode="lateral")
    p.x_rays = [XRay(s) for s in names]
    vals = workloads.inclusion_fractions(tuple(M), len(els), 4, rng)
    mat = MassFractionMaterial(els, CellFractions(vals, *[workloads.cell_edges(a) for a in g.axes],
                                                  embedding=vals.reshape(-1, len(els))[0]))
    return serving.Spec(p, mat, workloads.detector_40deg(), rng.uniform(0.5, 1.5, len(els)))


@pytest.mark.parametrize("tape", [True, False])
def test_served_3d_equals_alone(tape):
    """3D requests (the grid lanes' DIM = 3 kernels) mixed with 2D ones."""
    specs = [_spec_3d(3, (10, 8, 6), 26, 1), _spec_3d(5, (8, 6, 6), 23, 2), _spec_3d(3, (12, 6, 5), 24, 3)]
    specs += _specs(n=4, seed=13)
    preps = [serving.prepare(s, want_rho_bar=True) for s in specs]
    srv = serving.Server(hbm_budget_bytes=1 << 30, max_active=64)
    if not tape:
        srv.set_tape_limit(0)
    tickets = [srv.submit(pr, arrival_s=0.001 * n) for n, pr in enumerate(preps)]
    for t in tickets:
        tm = srv.wait(t)
        assert tm.status == 0 and tm.taped == int(tape)
    srv.close()
    for spec, pr in zip(specs, preps):
        I, rb, gen = _alone(spec)
        p = spec.problem
        print(f"N={p.N} M={p.gspec.M} steps={p.n_steps} I={pr.I_out}", flush=True)
        assert rel_err(pr.I_out, I) < 1e-12
        for c in p.gspec.active_classes():
            assert rel_err(pr.rho_bar[c], rb[c]) < 1e-12
        assert rel_err(pr.gen, gen) < 1e-12


def test_cancel_pending_and_admitted():
    """A cancelled request completes with status PNX_ERR_CANCELLED and writes no
    output; the requests served beside it are unaffected."""
    import time
    from pnepma_amd._lib import PNX_ERR_CANCELLED
    specs = _specs(n=6, seed=17)
    preps = [serving.prepare(s, want_rho_bar=True) for s in specs]
    srv = serving.Server(hbm_budget_bytes=1 << 30, max_active=64)
    t_pending = srv.submit(preps[0], arrival_s=srv.clock() + 3600.0)  # still pending when cancelled
    tickets = [srv.submit(pr) for pr in preps[1:]]
    t_end = time.perf_counter() + 10.0
    while time.perf_counter() < t_end:  # cancel request 1 once it is admitted
        done, tm = srv.poll(tickets[0])
        if done or tm.admit_s > 0:
            break
    srv.cancel(tickets[0])
    srv.cancel(t_pending)
    tm = srv.wait(t_pending)
    assert tm.status == PNX_ERR_CANCELLED
    assert not np.any(preps[0].I_out)
    tm = srv.wait(tickets[0])
    print("request 1 after cancel: status", tm.status, "ticks", tm.ticks, flush=True)
    assert tm.status in (0, PNX_ERR_CANCELLED)
    if tm.status == PNX_ERR_CANCELLED:
        assert not np.any(preps[1].I_out)
    for t in tickets[1:]:
        assert srv.wait(t).status == 0
    srv.drain()
    srv.close()
    for spec, pr in zip(specs[2:], preps[2:]):
        I, rb, gen = _alone(spec)
        assert rel_err(pr.I_out, I) < 1e-12


def test_spec_requests_match_prepared_requests_and_the_oracle(oracle_lib):
    """Requests given as cell fractions (pnx_server_submit_spec: records, beam
    moments and profiles built by the library's C++ set-up, densities and
    absorption on the device at admission, the material gradient at
    completion) against the same requests prepared on the host, and their
    material gradients against the device problem's and the oracle's."""
    specs = _specs(n=8, seed=23) + [_spec_3d(3, (10, 8, 6), 26, 4)]
    preps = [serving.prepare(s, want_rho_bar=True) for s in specs]
    reqs = [serving.SpecRequest(s, want_gradients=True) for s in specs]
    srv = serving.Server(hbm_budget_bytes=1 << 30, max_active=64)
    tk = [srv.submit(pr, arrival_s=0.001 * n) for n, pr in enumerate(preps)]
    tk += [srv.submit_spec(rq, arrival_s=0.001 * n) for n, rq in enumerate(reqs)]
    for t in tk:
        assert srv.wait(t).status == 0
    srv.close()
    for spec, pr, rq in zip(specs, preps, reqs):
        p = spec.problem
        e = [rel_err(rq.I_out, pr.I_out)] + [rel_err(rq.rho_bar[c], pr.rho_bar[c]) for c in p.gspec.active_classes()]
        e.append(rel_err(rq.gen, pr.gen))
        dp = DeviceProblem(p, None, ckpt_every=0)
        dp.set_cell_fractions(spec.material, None)
        dp.set_absorption(spec.material, spec.detector)
        dp.forward_adjoint(spec.I_bar[None, :], want_rho_bar=False)
        g_dev = dp.material_gradient()
        dp.close()
        e_g = rel_err(rq.material_gradient(), g_dev)
        mac = np.array([[m for m in row] for row in rq.mac])
        I_ref, rb, acb, noab = oracle_lib.forward_adjoint(p, pr.si, spec.I_bar)
        g_ref = oracle_lib.material_gradient(p, spec.material, spec.detector, mac, rb, acb, noab)
        e_o = rel_err(rq.material_gradient(), g_ref)
        print(f"N={p.N} M={p.gspec.M}: spec vs prepared {max(e):.2e}, gradient vs device problem {e_g:.2e}, "
              f"vs oracle {e_o:.2e}, I vs oracle {rel_err(rq.I_out, I_ref):.2e}", flush=True)
        assert max(e) < 1e-12 and e_g < 1e-11 and e_o < 1e-8 and rel_err(rq.I_out, I_ref) < 1e-10


@pytest.mark.parametrize("taped", [True, False], ids=["taped", "checkpointed"])
def test_spec_requests_per_cell_density_gradient(oracle_lib, taped):
    """2D requests given as cell fractions whose caller takes no per-class
    density gradient: their reverse passes add the four classes of a cell into
    one per-cell rho_bar in the block (DensArgs::rb_cell), which only the
    material pullback reads.  Intensities equal those of the same requests
    returning per-class gradients bit for bit (the forward is the same), the
    material gradients agree to rounding, and both agree with the oracle's."""
    specs = [s for s in serving.serving_requests(24, seed=29, sizes_1d=(30, 70), sizes_2d=(10, 40), max_steps=40)
             if s.problem.gspec.axes[1].M > 1][:6]
    assert len(specs) >= 4 and all(len(s.material.elements) <= 4 for s in specs)
    per_class = [serving.SpecRequest(s, want_gradients=True) for s in specs]
    per_cell = [serving.SpecRequest(s) for s in specs]
    srv = serving.Server(hbm_budget_bytes=1 << 30, max_active=64)
    if not taped:
        srv.set_tape_limit(0)
    tk = [srv.submit_spec(rq, arrival_s=0.001 * n) for n, rq in enumerate(per_class + per_cell)]
    for t in tk:
        assert srv.wait(t).status == 0
    srv.close()
    for spec, a, b in zip(specs, per_class, per_cell):
        p = spec.problem
        pr = serving.prepare(spec, want_rho_bar=True)
        I_ref, rb, acb, noab = oracle_lib.forward_adjoint(p, pr.si, spec.I_bar)
        mac = np.array([[m for m in row] for row in a.mac])
        g_ref = oracle_lib.material_gradient(p, spec.material, spec.detector, mac, rb, acb, noab)
        e_cells = rel_err(b.material_gradient(), a.material_gradient())
        e_o = rel_err(b.material_gradient(), g_ref)
        print(f"N={p.N} M={p.gspec.M} ne={len(spec.material.elements)}: per-cell vs per-class gradient {e_cells:.2e}, "
              f"vs oracle {e_o:.2e}", flush=True)
        assert np.array_equal(a.I_out, b.I_out)
        assert_physical(b.I_out)
        assert e_cells < 1e-12 and e_o < 1e-8


def _serve_1d(preps, max_active, together):
    srv = serving.Server(hbm_budget_bytes=4 << 30, max_active=max_active)
    # with one arrival time for all, the first tick admits every request, so
    # each march launch holds all of them
    at = srv.clock() + 0.3 if together else 0.0
    for t in [srv.submit(pr, arrival_s=at) for pr in preps]:
        assert srv.wait(t).status == 0
    srv.close()


def test_served_1d_multi_tile_requests_do_not_depend_on_the_workgroup_width(oracle_lib):
    """1D requests of two and three tiles (257-600 cells): a march launch of at
    most 16 requests runs a 512-thread workgroup per request (two tiles of a
    pass side by side), one of more requests a 256-thread workgroup (tiles in
    turn).  The results are bitwise the same, equal the request solved alone
    and match the oracle."""
    drawn = serving.serving_requests(300, seed=41, sizes_1d=(257, 600), sizes_2d=(10, 12), max_steps=24)
    specs = [s for s in drawn if s.problem.gspec.dims == 1][:72]
    assert len(specs) == 72
    assert {(s.problem.gspec.axes[0].M + 255) // 256 for s in specs} == {2, 3}
    preps = [serving.prepare(s, want_rho_bar=True) for s in specs]
    for k in range(0, len(preps), 8):
        _serve_1d(preps[k:k + 8], 8, together=False)
    wide = [(pr.I_out.copy(), [None if x is None else x.copy() for x in pr.rho_bar], pr.gen.copy()) for pr in preps]
    _serve_1d(preps, 128, together=True)
    for spec, pr, (I, rb, gen) in zip(specs, preps, wide):
        p = spec.problem
        assert np.array_equal(pr.I_out, I), (p.N, p.gspec.M)
        for c in p.gspec.active_classes():
            assert np.array_equal(pr.rho_bar[c], rb[c]), (p.N, p.gspec.M, c)
        assert np.array_equal(pr.gen, gen)
    by_order = {}
    for spec, pr in zip(specs, preps):
        by_order.setdefault(spec.problem.N, (spec, pr))
    for spec, pr in by_order.values():
        p = spec.problem
        print(f"N={p.N} M={p.gspec.M} steps={p.n_steps} I={pr.I_out}", flush=True)
        I, rb, gen = _alone(spec)
        assert np.array_equal(pr.I_out, I) or rel_err(pr.I_out, I) < 1e-12
        for c in p.gspec.active_classes():
            assert rel_err(pr.rho_bar[c], rb[c]) < 1e-12
        Io, rbo, acb, _ = oracle_lib.forward_adjoint(p, pr.si, spec.I_bar)
        assert_physical(pr.I_out)
        assert rel_err(pr.I_out, Io) < 1e-10
        for c in p.gspec.active_classes():
            assert rel_err(pr.rho_bar[c], rbo[c]) < 1e-8
