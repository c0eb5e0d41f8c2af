"""Device material layer: class densities from per-cell mass fractions, and the
pullback of the density gradient to those fractions, against the host
restatement (pnepma_amd.material, pn_material.jl / pn_parametrization_boxed.jl).

* densities: a problem whose densities were computed on the device gives the
  same intensities as one given the host densities (bitwise: the two compute
  w_e / sum(w / rho_pure) in the same order);
* pullback: the device gradient of sum(rho_bar * rho) (plus the eee
  number-of-atoms term) equals the host densities_pullback summed over the
  classes, to 1e-12;
* calc_intensities_pullback takes the device path for grid-aligned box
  materials and agrees with the host path to 1e-12."""
import numpy as np
import pytest

from pnepma_amd import physics, solver, workloads
from pnepma_amd.grid import GridSpec
from pnepma_amd.material import CellFractions, MassFractionMaterial
from pnepma_amd.physics import XRay, element
from pnepma_amd import serving
from pnepma_amd.problem import ForwardPNProblem, solve_inputs
from pnepma_amd.solver import DeviceProblem, cell_aligned

from .helpers import rel_err

pytestmark = pytest.mark.gpu


def _problem(M, N=3, n_steps=None, seed=4, n_beams=2):
    """n_steps None: the stable step count (nu <= 0.5, at least 30)."""
    rng = np.random.default_rng(seed)
    names = ["Fe", "Cr", "Ni"]
    els = [element(s) for s in names]
    dims = sum(m > 1 for m in M)
    lo = [(-1.5e-4, 0.0), (-0.5e-4, 0.5e-4), (-0.4e-4, 0.4e-4)]
    g = GridSpec.make(M, lo[0], lo[1] if dims >= 2 else 0.0, lo[2] if dims >= 3 else 0.0)
    ys = np.linspace(-0.3e-4, 0.3e-4, n_beams) if dims >= 2 else np.zeros(n_beams)
    p = ForwardPNProblem(N=N, gspec=g, elements=els, beam_energy=(15e3, workloads.SIGMA_E),
                         beam_positions=[(0.0, float(y), 0.0) for y in ys], beam_sigma=50 * workloads.NM,
                         beam_direction=((-1.0, 0.0, 0.0), workloads.KAPPA), E_initial=15.5e3, E_cutoff=5e3,
                         n_steps=n_steps or 30, source_mode="lateral")
    p.x_rays = [XRay(s) for s in names]
    vals = workloads.inclusion_fractions(tuple(M), len(els), 4, rng)
    mat = MassFractionMaterial(els, CellFractions(vals, *[workloads.cell_edges(a) for a in g.axes],
                                                  embedding=vals.reshape(-1, len(els))[0]))
    if n_steps is None:
        p.n_steps = max(30, serving.cfl_steps(N, g, mat, els, p.E_initial, p.E_cutoff))
    return p, mat


@pytest.mark.parametrize("M", [(40, 1, 1), (20, 14, 1), (10, 8, 6)], ids=["1d", "2d", "3d"])
def test_device_densities_give_the_same_intensities(M):
    p, mat = _problem(M)
    assert cell_aligned(mat, p.gspec)
    si = solve_inputs(p, mat, workloads.detector_40deg())
    I_host = DeviceProblem(p, si, adjoint=False).forward()
    dp = DeviceProblem(p, si, adjoint=False)
    dp.set_cell_fractions(mat, si)
    I_dev = dp.forward()
    print("intensities host", I_host[0], "device", I_dev[0], flush=True)
    assert np.array_equal(I_host, I_dev)


@pytest.mark.parametrize("M", [(40, 1, 1), (20, 14, 1), (10, 8, 6)], ids=["1d", "2d", "3d"])
def test_device_fraction_gradient_matches_host(M):
    p, mat = _problem(M)
    si = solve_inputs(p, mat, workloads.detector_40deg())
    dp = DeviceProblem(p, si, ckpt_every=8)
    dp.set_cell_fractions(mat, si)
    ib = np.random.default_rng(5).uniform(0.5, 1.5, (p.n_beams, len(p.x_rays)))
    _, rb, gen = dp.forward_adjoint(ib)
    extra = np.random.default_rng(6).normal(size=rb[0].shape)
    g_dev = dp.cell_fraction_gradient(extra_eee=extra)
    g_host = np.zeros(mat.param.n_params)
    for c in range(8):
        if np.any(rb[c]):
            g_host += mat.densities_pullback(p.gspec.class_points(c), rb[c])
    g_host += mat.densities_pullback(p.gspec.class_points(0), extra)
    e = rel_err(g_dev, g_host)
    print(f"cell-fraction gradient rel err {e:.3e}", flush=True)
    assert e < 1e-12


def test_calc_intensities_pullback_device_path():
    p, mat = _problem((20, 14, 1), n_beams=3)
    det = workloads.detector_40deg()
    ib = np.random.default_rng(7).uniform(0.5, 1.5, (p.n_beams, len(p.x_rays)))
    I, back = solver.calc_intensities_pullback(p, mat, det, ckpt_every=8)
    g_dev = back(ib)
    si = solve_inputs(p, mat, det)
    dp = DeviceProblem(p, si, ckpt_every=8)
    _, rb, gen = dp.forward_adjoint(ib)
    g_host = solver.material_gradient(p, mat, det, si, rb, gen)
    e = rel_err(g_dev, g_host)
    print(f"material gradient device vs host rel err {e:.3e}", flush=True)
    assert e < 1e-12


@pytest.mark.parametrize("M", [(40, 1, 1), (20, 14, 1), (10, 8, 6)], ids=["1d", "2d", "3d"])
def test_device_absorption_factors_match_host(M):
    p, mat = _problem(M)
    det = workloads.detector_40deg()
    si = solve_inputs(p, mat, det)
    dp = DeviceProblem(p, si, adjoint=False)
    dp.set_cell_fractions(mat, si)
    ac = dp.set_absorption(mat, det)
    e = rel_err(ac, si.ac)
    print(f"absorption factors device vs host rel err {e:.3e} (min factor {si.ac.min():.3f})", flush=True)
    assert e < 1e-13
    # intensities with device line weights agree with the host-weighted solve
    I_dev = dp.forward()
    I_host = DeviceProblem(p, si, adjoint=False).forward()
    assert rel_err(I_dev, I_host) < 1e-13


@pytest.mark.parametrize("M", [(40, 1, 1), (20, 14, 1), (10, 8, 6)], ids=["1d", "2d", "3d"])
def test_device_material_gradient_matches_host(M):
    p, mat = _problem(M, n_beams=2)
    det = workloads.detector_40deg()
    si = solve_inputs(p, mat, det)
    ib = np.random.default_rng(9).uniform(0.5, 1.5, (p.n_beams, len(p.x_rays)))
    dp = DeviceProblem(p, si, ckpt_every=8)
    dp.set_cell_fractions(mat, si)
    dp.set_absorption(mat, det)
    _, _, gen = dp.forward_adjoint(ib, want_rho_bar=False)
    g_dev = dp.material_gradient()
    hp = DeviceProblem(p, si, ckpt_every=8)
    _, rb, gen_h = hp.forward_adjoint(ib)
    g_host = solver.material_gradient(p, mat, det, si, rb, gen_h)
    e = rel_err(g_dev, g_host)
    print(f"whole material gradient device vs host rel err {e:.3e}", flush=True)
    assert e < 1e-11


def _oracle_inputs(p, mat, det):
    """SolveInputs built from the oracle's material layer alone."""
    from oracle import oracle
    from pnepma_amd.problem import SolveInputs, el_index
    g = p.gspec
    mac = np.array([[physics.mass_absorption(x, e) for e in mat.elements] for x in p.x_rays])
    am = np.array([e.atom_mass_g for e in mat.elements])
    rho = [oracle.component_densities(mat, g.class_points(c)) for c in range(8)]
    noa = oracle.number_of_atoms(mat, g.class_points(0), am)
    ac = oracle.absorption(mat, g.class_points(0), mac, det, p.beam_positions[0][0])
    ei = el_index(mat.elements, p.x_rays)
    return SolveInputs(rho, noa, ac, ei, np.ascontiguousarray(ac * noa[:, ei])), mac


@pytest.mark.parametrize("M", [(40, 1, 1), (20, 14, 1), (10, 8, 6)], ids=["1d", "2d", "3d"])
def test_device_material_layer_matches_oracle(M, oracle_lib):
    """Densities, absorption factors and the whole material gradient of the
    reference adjoint computed on the device (pnx_problem_set_cell_fractions,
    pnx_problem_set_absorption, pnx_problem_material_gradient) against the
    oracle's restatement of pn_material.jl:236-256, the box parametrization
    and x_ray_intensities.jl:43-153, composed with the oracle's reference
    adjoint: intensities 1e-10, absorption factors 1e-13, gradient 1e-8."""
    p, mat = _problem(M, n_beams=2)
    det = workloads.detector_40deg()
    si_o, mac = _oracle_inputs(p, mat, det)
    ib = np.random.default_rng(11).uniform(0.5, 1.5, (p.n_beams, len(p.x_rays)))
    dp = DeviceProblem(p, None, ckpt_every=8)
    dp.set_cell_fractions(mat, None)
    ac = dp.set_absorption(mat, det)
    e_ac = rel_err(ac, si_o.ac)
    I, _, _ = dp.forward_adjoint(ib, want_rho_bar=False)
    g_dev = dp.material_gradient()
    dp.close()
    g_ref = np.zeros_like(g_dev)
    for b in range(p.n_beams):
        I_ref, rb, acb, noab = oracle_lib.forward_adjoint(p, si_o, ib[b], beam=b)
        assert rel_err(I[b], I_ref) < 1e-10
        g_ref += oracle_lib.material_gradient(p, mat, det, mac, rb, acb, noab)
    e_g = rel_err(g_dev, g_ref)
    print(f"{M}: absorption factors {e_ac:.3e}, material gradient {e_g:.3e}", flush=True)
    assert e_ac < 1e-13
    assert e_g < 1e-8
