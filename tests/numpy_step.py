"""Independent dense numpy formulation of the StaRMAP energy step and of the
reference pullback loop, used to pin the C oracle on tiny grids.

It shares no code with the oracle: states are (nz, ny, nx, n_moments) arrays,
the advection and boundary blocks are applied as dense matrix products over
whole slabs of points, and the boundary ghosts are written slab-wise.
"""
from __future__ import annotations

import numpy as np

from pnepma_amd import grid as G


def phi1(c):
    c = np.asarray(c, dtype=np.float64)
    small = np.abs(c) <= 0.0002
    with np.errstate(divide="ignore", invalid="ignore"):
        big = (np.exp(c) - 1.0) / c
    return np.where(small, 1.0 + 0.5 * c + c * c / 6.0, big)


class DenseProblem:
    def __init__(self, p, si, A, B, adjoint=False, beam=0):
        """A[c][d]: dense advection (forward sign); B[(c, d)]: boundary block of the
        requested direction; si: SolveInputs."""
        self.p, self.g, self.N = p, p.gspec, p.N
        self.cm = G.class_moments(p.N)
        self.adjoint = adjoint
        self.sign = -1.0 if adjoint else 1.0
        self.A, self.B = A, B
        self.size = [self.g.class_size(c)[::-1] for c in range(8)]  # (nz, ny, nx)
        self.rho = [si.rho[c].reshape(self.size[c] + (-1,)) for c in range(8)]
        self.h = self.g.h
        mms = p.beam_moments()
        prof = p.beam_profiles(beam)
        self.gset = [None] * 8
        for c in range(8):
            if G.odd_in(c, 0) and not adjoint:
                nz, ny, _ = self.size[c]
                self.gset[c] = prof[c].reshape(nz, ny)[..., None] * mms[c][None, None, :]

    def set_record(self, rec):
        ne, N = len(self.p.elements), self.N
        s, cl, gfac = rec[:ne], rec[ne:ne + ne * (N + 1)].reshape(ne, N + 1), rec[ne + ne * (N + 1)]
        self.s, self.cl, self.gfac = s, cl, gfac
        self.S, self.C = [], []
        for c in range(8):
            ls = np.array([l for l, _ in self.cm[c]], dtype=int)
            S = np.repeat((self.rho[c] @ s)[..., None], len(ls), axis=-1)
            C = self.rho[c] @ cl[:, ls] if len(ls) else np.zeros(self.size[c] + (0,))
            self.S.append(S)
            self.C.append(C)

    def zeros(self):
        return [np.zeros(self.size[c] + (len(self.cm[c]),)) for c in range(8)]

    @staticmethod
    def _sl(axis, idx):
        """index tuple selecting position idx along spatial axis (0 = x) of a (z, y, x, m) array"""
        t = [slice(None)] * 4
        t[2 - axis] = idx
        return tuple(t)

    def boundary(self, u, parity):
        for a in range(3):
            if self.g.axes[a].singleton:
                continue
            for c in range(8):
                if not (G.odd_in(c, a) and G.parity_of(c) == parity) or len(self.cm[c]) == 0:
                    continue
                v = u[G.switch_axis(c, a)]
                Bm = self.B[(c, a)]
                n = u[c].shape[2 - a]
                src = (a == 0 and not self.adjoint)
                left = v[self._sl(a, 0)] @ Bm.T
                right = v[self._sl(a, n - 2)] @ Bm.T
                g2 = self.gset[c] * self.gfac if src else 0.0
                u[c][self._sl(a, 0)] = 2.0 * (-left + 0.0) - u[c][self._sl(a, 1)]
                u[c][self._sl(a, n - 1)] = 2.0 * (right + g2) - u[c][self._sl(a, n - 2)]

    def calc_du(self, u, parity, dt, use_phi, Q=None):
        du = {}
        for c in range(8):
            if G.parity_of(c) != parity or len(self.cm[c]) == 0:
                continue
            n = self.size[c]
            rng = []
            for a in range(3):
                m = n[2 - a]
                rng.append(slice(1, m - 1) if G.odd_in(c, a) else slice(0, m))
            inner = (rng[2], rng[1], rng[0])
            d = np.zeros(u[c].shape)
            acc = np.zeros(u[c][inner].shape)
            for a in range(3):
                if n[2 - a] == 1:
                    continue
                f = G.switch_axis(c, a)
                if len(self.cm[f]) == 0:
                    continue
                R, Lq = list(inner), list(inner)
                r = rng[a]
                if G.odd_in(c, a):
                    R[2 - a] = slice(r.start, r.stop)
                    Lq[2 - a] = slice(r.start - 1, r.stop - 1)
                else:
                    R[2 - a] = slice(r.start + 1, r.stop + 1)
                    Lq[2 - a] = slice(r.start, r.stop)
                diff = (u[f][tuple(R)] - u[f][tuple(Lq)]) / self.h[a]
                acc = acc + diff @ (self.sign * self.A[c][a]).T
            q = Q[c][inner] if Q is not None else 0.0
            t = (q - acc - self.C[c][inner] * u[c][inner]) / self.S[c][inner]
            if use_phi:
                t = t * phi1(-self.C[c][inner] / self.S[c][inner] * dt)
            d[inner] = t
            du[c] = d
        return du

    def step(self, u, dt, Q=None):
        h = dt / 2
        for par, tau in ((0, h), (1, dt), (0, h)):
            self.boundary(u, par)
            du = self.calc_du(u, 1 - par, tau, True, Q)
            for c, d in du.items():
                u[c] += tau * d


def dense_blocks(N, adjoint, adv, bnd):
    """adv(N,c,d,nc,nf), bnd(N,c,d,adjoint,nc,nf) -> dict inputs of DenseProblem."""
    cm = G.class_moments(N)
    A = [[None] * 3 for _ in range(8)]
    B = {}
    for c in range(8):
        for d in range(3):
            f = G.switch_axis(c, d)
            A[c][d] = adv(N, c, d, len(cm[c]), len(cm[f]))
            if G.odd_in(c, d):
                B[(c, d)] = bnd(N, c, d, adjoint, len(cm[c]), len(cm[f]))
    return A, B


def pullback(p, si, adv, bnd, I_bar, beam=0):
    """The reference pullback (pn_intensities.jl:353-404) with a full tape, in
    dense numpy.  Returns (I, rho_bar[8] as (P_c, n_e), ac_bar, noa_bar)."""
    fwd, adj, dens = p.records()
    ne, nl = len(p.elements), len(p.x_rays)
    sig0 = ne + ne * (p.N + 1) + 1
    A, Bf = dense_blocks(p.N, False, adv, bnd)
    _, Ba = dense_blocks(p.N, True, adv, bnd)
    F = DenseProblem(p, si, A, Bf, False, beam)
    Ad = DenseProblem(p, si, A, Ba, True, beam)
    dE, dvol = p.dE, p.gspec.dvol
    u = F.zeros()
    tape = [[x.copy() for x in u]]
    I = np.zeros(nl)
    Wl = si.ac * si.noa[:, si.el_idx]

    def eee0(st):
        return st[0][..., 0].ravel()

    for k in range(1, p.n_steps + 1):
        F.set_record(fwd[k])
        F.step(u, dE)
        tape.append([x.copy() for x in u])
        I += dE * dvol * fwd[k][sig0:] * (Wl * eee0(u)[:, None]).sum(axis=0) * 1.0
    lam = F.zeros()
    rb = [np.zeros(F.rho[c].shape) for c in range(8)]
    acb = np.zeros_like(si.ac)
    noab = np.zeros_like(si.noa)

    def densities(uu, rec, w):
        F.set_record(rec)
        uu = [x.copy() for x in uu]
        F.boundary(uu, 0)
        F.boundary(uu, 1)
        du = {}
        du.update(F.calc_du(uu, 0, dE, False))
        du.update(F.calc_du(uu, 1, dE, False))
        for c, d in du.items():
            ls = [l for l, _ in F.cm[c]]
            for e in range(ne):
                term = lam[c] * (F.s[e] * d + F.cl[e, ls] * uu[c]) * dvol * dE * w
                n = F.size[c]
                mask = np.ones(n, dtype=bool)
                for a in range(3):
                    if G.odd_in(c, a):
                        sl = [slice(None)] * 3
                        sl[2 - a] = 0
                        mask[tuple(sl)] = False
                        sl[2 - a] = n[2 - a] - 1
                        mask[tuple(sl)] = False
                rb[c][..., e] += np.where(mask, term.sum(axis=-1), 0.0)

    def intensity_pb(uu, rec, w):
        integ = dE * dvol * rec[sig0:] * w * I_bar
        u0 = eee0(uu)
        acb[:] += si.noa[:, si.el_idx] * u0[:, None] * integ[None, :]
        np.add.at(noab.T, si.el_idx, (si.ac * u0[:, None] * integ[None, :]).T)

    for i in range(p.n_steps, 0, -1):
        densities(tape[i], dens[i], 1.0)
        intensity_pb(tape[i], dens[i], 1.0)
        q = Ad.zeros()
        q[0][..., 0] = (-(adj[i][sig0:][None, :] * Wl * I_bar[None, :]).sum(axis=1)).reshape(q[0].shape[:3])
        Ad.set_record(adj[i])
        Ad.step(lam, dE, q)
    densities(tape[0], dens[0], 0.5)
    intensity_pb(tape[0], dens[0], 0.5)
    return I, [r.reshape(-1, ne) for r in rb], acb, noab
