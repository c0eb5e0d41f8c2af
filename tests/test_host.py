"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol include/pnx.h declares, its matrices equal the oracle's, argument
validation, the host tables and the material layer."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pnepma_amd import _lib
from pnepma_amd import grid as G
from pnepma_amd import physics, workloads
from pnepma_amd.material import CellFractions, MassFractionMaterial

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "pnx.h")).read()
    return sorted(set(re.findall(r"\b(pnx_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"


def test_matrices_match_oracle(oracle_lib):
    L = _lib.lib()
    for N in (1, 3, 5, 7, 9):
        cm = G.class_moments(N)
        for c in range(8):
            for d in range(3):
                f = G.switch_axis(c, d)
                nc, nf = len(cm[c]), len(cm[f])
                if nc == 0 or nf == 0:
                    continue
                a = np.zeros((nc, nf))
                _lib.check(L.pnx_advection_block(N, c, d, _lib.dptr(a)))
                assert np.array_equal(a, oracle_lib.advection_dense(N, c, d, nc, nf))
                if G.odd_in(c, d):
                    for adj in (0, 1):
                        b = np.zeros((nc, nf))
                        _lib.check(L.pnx_boundary_block(N, c, d, adj, _lib.dptr(b)))
                        assert np.abs(b - oracle_lib.boundary_dense(N, c, d, adj, nc, nf)).max() < 1e-14


def test_coefficient_buffer_layout():
    """All advection nonzeros come first in column-major (CSC) order per block."""
    L = _lib.lib()
    N = 5
    n = L.pnx_coefficients(N, 0, None, 0)
    buf = np.zeros(n)
    assert L.pnx_coefficients(N, 0, _lib.dptr(buf), n) == n
    adj = np.zeros(n)
    L.pnx_coefficients(N, 1, _lib.dptr(adj), n)
    cm = G.class_moments(N)
    idx = 0
    for c in range(8):
        for d in range(3):
            f = G.switch_axis(c, d)
            a = np.zeros((len(cm[c]), len(cm[f])))
            L.pnx_advection_block(N, c, d, _lib.dptr(a))
            for j in range(len(cm[f])):
                for i in range(len(cm[c])):
                    (l, k), (lf, kf) = cm[c][i], cm[f][j]
                    if abs(l - lf) == 1:
                        pat = {0: (kf == (k - (1 if k > 0 else -1) if k else -1) and k != -1)
                                  or kf == (k + (1 if k > 0 else -1) if k else 1),
                               1: (kf == -(k - (1 if k > 0 else -1) if k else -1) and k != 1)
                                  or kf == -(k + (1 if k > 0 else -1) if k else 1),
                               2: kf == k}[d]
                        if pat:
                            assert buf[idx] == a[i, j]
                            assert adj[idx] == -a[i, j]
                            idx += 1
                    else:
                        assert a[i, j] == 0.0


def test_order_validation():
    L = _lib.lib()
    assert L.pnx_advection_block(11, 0, 0, None) < 0
    assert "bad argument" in L.pnx_last_error().decode()


def test_stopping_power_table():
    el = physics.element("Cu")
    sp = physics.StoppingPower(el)
    k = sp.knots[40]
    assert sp(k) == pytest.approx(sp.values[40], rel=1e-15)
    mid = 0.5 * (sp.knots[40] + sp.knots[41])
    slope = (sp.values[41] - sp.values[40]) / (sp.knots[41] - sp.knots[40])
    assert sp.derivative(mid) == pytest.approx(slope, rel=1e-14)
    assert sp.derivative(k) == pytest.approx(slope, rel=1e-14)  # right segment at a knot


def test_energy_recurrences_and_records():
    p, mat, stds, det = workloads.slab1d(n_steps=30)
    fe = p.forward_energies()
    assert fe[0][0] == pytest.approx(p.E_initial - p.dE / 2)
    assert fe[-1][1] == pytest.approx(p.E_cutoff, abs=1e-9)
    re_ = p.reverse_energies()
    assert re_[p.n_steps] == p.E_cutoff
    assert re_[0] == pytest.approx(p.E_initial, abs=1e-9)
    fwd, adj, dens = p.records()
    assert fwd.shape == (31, p.record_len) and adj.shape == fwd.shape and dens.shape == fwd.shape
    ne = len(p.elements)
    # adjoint records carry no stopping-power derivative and no beam factor
    assert np.all(adj[:, ne + ne * (p.N + 1)] == 0.0)
    assert np.all(adj[:, ne:ne + ne * (p.N + 1)].reshape(-1, ne, p.N + 1)[:, :, 0] == 0.0)


def test_mass_fraction_pullback_fd():
    els = [physics.element(s) for s in ("Fe", "Cr", "Ni")]
    rng = np.random.default_rng(0)
    vals = rng.uniform(0.1, 1.0, (3, 2, 1, 3))
    param = CellFractions(vals, np.linspace(0, 3, 4), np.linspace(0, 2, 3), np.array([-1.0, 1.0]), [0.3, 0.3, 0.4])
    m = MassFractionMaterial(els, param)
    pts = np.column_stack([rng.uniform(0, 3, 50), rng.uniform(0, 2, 50), np.zeros(50)])
    rbar = rng.standard_normal((50, 3))
    g = m.densities_pullback(pts, rbar)
    p0 = param.params()
    for q in rng.choice(p0.size, 6, replace=False):
        h = 1e-6
        pp, pm = p0.copy(), p0.copy()
        pp[q] += h
        pm[q] -= h
        fd = ((m.with_params(pp).densities(pts) - m.with_params(pm).densities(pts)) * rbar).sum() / (2 * h)
        assert g[q] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def _expected_sequence(n, K):
    """The operation order of the reference pullback with checkpoints
    (pn_intensities.jl:366-404 with the tape replaced by checkpoints every K
    steps), as the solver's loops run it, with the fused-pair flags of the
    forward runs: a run is one checkpoint segment of the forward pass or one
    recompute; its first step starts with the first odd half step, every step
    but its last ends with the fused odd pair."""
    K = min(K, n)
    nck = -(-n // K)
    ops = []
    for k in range(1, n + 1):
        seg_start = (k - 1) % K == 0
        ops.append((1, k, (k - 1) // K if seg_start else -1, 0, 0, 0, 0, int(seg_start), int(k < n and k % K != 0)))
    for s in range(nck - 1, -1, -1):
        k0, k1 = s * K, min(n, s * K + K)
        for k in range(k0 + 1, k1 + 1):
            src = (1, s) if k == k0 + 1 else (2, k - k0 - 2)
            ops.append((2, k, -1, *src, 2, k - k0 - 1, int(k == k0 + 1), int(k < k1)))
        for i in range(k1, k0, -1):
            ops.append((3, i, -1, 2, i - k0 - 1, 0, 0, 0, 0))
    ops.append((4, 0, -1, 1, 0, 0, 0, 0, 0))
    return ops


def _expected_taped_sequence(n):
    """The reference pullback's own order (pn_intensities.jl:366-404), every
    forward state kept: one forward run from the zero state (buffer 0) into
    slots 0..n-1 of buffer 2, the reverse steps on those slots, the final
    pullback on the zero state."""
    ops = []
    for k in range(1, n + 1):
        src = (0, 0) if k == 1 else (2, k - 2)
        ops.append((1, k, -1, *src, 2, k - 1, int(k == 1), int(k < n)))
    for i in range(n, 0, -1):
        ops.append((3, i, -1, 2, i - 1, 0, 0, 0, 0))
    ops.append((4, 0, -1, 0, 0, 0, 0, 0, 0))
    return ops


@pytest.mark.parametrize("n", [1, 2, 7, 16, 33, 60])
def test_served_operation_sequence(n):
    L = _lib.lib()
    for K in sorted({0, 1, 2, 3, 5, max(1, int(np.ceil(np.sqrt(n)))), n, n + 3}):
        m = L.pnx_march_sequence(n, K, None, 0)
        assert m == (2 * n + 1 if K == 0 else 3 * n + 1)
        buf = (C.c_int32 * (9 * m))()
        assert L.pnx_march_sequence(n, K, buf, 9 * m) == m
        got = [tuple(buf[9 * t + i] for i in range(9)) for t in range(m)]
        # destinations of reverse/final ops are unused: compare what the op reads and writes
        exp = _expected_taped_sequence(n) if K == 0 else _expected_sequence(n, K)
        assert len(got) == len(exp)
        for t, (a, b) in enumerate(zip(got, exp)):
            if a[0] in (3, 4):
                assert a[:5] == b[:5], (n, K, t, a, b)
            else:
                assert a == b, (n, K, t, a, b)
