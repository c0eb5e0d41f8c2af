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


# ----------------------------------------------------------------------------
# Exact (discrete) adjoint of the forward march, dense numpy formulation.
def dphi1(x):
    """Derivative of the forward's phi1 (Taylor branch |x| <= 2e-4, else (e^x - 1)/x)."""
    x = np.asarray(x, dtype=np.float64)
    small = np.abs(x) <= 0.0002
    mid = (~small) & (np.abs(x) < 0.05)
    ser = np.zeros_like(x)
    fact = 1.0
    for n in range(1, 12):  # sum_n n x^(n-1) / (n+1)!
        fact *= (n + 1)
        ser = ser + n * x ** (n - 1) / fact
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        big = (np.exp(x) * (x - 1.0) + 1.0) / (x * x)
    return np.where(small, 0.5 + x / 3.0, np.where(mid, ser, big))


def _combine(F, even_src, odd_src):
    return [ (even_src[c] if G.parity_of(c) == 0 else odd_src[c]).copy() for c in range(8)]


def _inner(F, c):
    n = F.size[c]
    rng = []
    for a in range(3):
        m = n[2 - a]
        rng.append(slice(1, m - 1) if G.odd_in(c, a) else slice(0, m))
    return (rng[2], rng[1], rng[0])


def _adjoint_half(F, lam, s, P, tau, rb, rec):
    """Backpropagate lam through the half step of parity P applied to state s."""
    F.set_record(rec)
    s = [x.copy() for x in s]
    F.boundary(s, 1 - P)          # ghosts of the from-classes, as the forward half step saw them
    ne = len(F.p.elements)
    ubar = [np.zeros_like(x) for x in s]
    for c in range(8):
        if G.parity_of(c) != P or not F.cm[c]:
            continue
        inner = _inner(F, c)
        n = F.size[c]
        S = F.S[c][inner]
        Cm = F.C[c][inner]
        u = s[c][inner]
        # r = -d - C u with d from s (ghosts included)
        d = np.zeros(u.shape)
        for a in range(3):
            if n[2 - a] == 1:
                continue
            f = G.switch_axis(c, a)
            if not F.cm[f]:
                continue
            R, L = list(inner), list(inner)
            r_ = inner[2 - a]
            if G.odd_in(c, a):
                R[2 - a], L[2 - a] = slice(r_.start, r_.stop), slice(r_.start - 1, r_.stop - 1)
            else:
                R[2 - a], L[2 - a] = slice(r_.start + 1, r_.stop + 1), slice(r_.start, r_.stop)
            d = d + ((s[f][tuple(R)] - s[f][tuple(L)]) / F.h[a]) @ F.A[c][a].T
        r = -d - Cm * u
        x = -Cm / S * tau
        ph = phi1(x)
        yb = lam[c][inner]
        w = yb * tau * ph / S
        # densities: y = u + tau phi(x) r / S, x = -C tau / S
        dS = (tau * r / S ** 2) * (-x * dphi1(x) - ph)
        dC = -(tau / S) * (dphi1(x) * tau * r / S + ph * u)
        ls = [l for l, _ in F.cm[c]]
        for e in range(ne):
            rb[c][inner + (e,)] += (yb * (dS * F.s[e] + dC * F.cl[e, ls])).sum(axis=-1)
        # own values
        ubar[c][inner] += yb - w * Cm
        # from-classes through the stencil (scatter form, ghosts included)
        for a in range(3):
            if n[2 - a] == 1:
                continue
            f = G.switch_axis(c, a)
            if not F.cm[f]:
                continue
            gbar = -(w @ F.A[c][a]) / F.h[a]       # d(y)/d(delta) contribution
            R, L = list(inner), list(inner)
            r_ = inner[2 - a]
            if G.odd_in(c, a):
                R[2 - a], L[2 - a] = slice(r_.start, r_.stop), slice(r_.start - 1, r_.stop - 1)
            else:
                R[2 - a], L[2 - a] = slice(r_.start + 1, r_.stop + 1), slice(r_.start, r_.stop)
            ubar[f][tuple(R)] += gbar
            ubar[f][tuple(L)] -= gbar
    # chain the from-classes' ghost adjoints through apply_boundary (order reversed)
    for a in reversed(range(3)):
        if F.g.axes[a].singleton:
            continue
        for f in reversed(range(8)):
            if not (G.odd_in(f, a) and G.parity_of(f) == 1 - P) or not F.cm[f]:
                continue
            v = G.switch_axis(f, a)     # class of parity P (the face values)
            Bm = F.B[(f, a)]
            m = s[f].shape[2 - a]
            gl = ubar[f][F._sl(a, 0)].copy()
            gr = ubar[f][F._sl(a, m - 1)].copy()
            ubar[f][F._sl(a, 0)] = 0.0
            ubar[f][F._sl(a, m - 1)] = 0.0
            ubar[f][F._sl(a, 1)] -= gl
            ubar[f][F._sl(a, m - 2)] -= gr
            ubar[v][F._sl(a, 0)] += -2.0 * (gl @ Bm)
            ubar[v][F._sl(a, m - 2 if False else s[v].shape[2 - a] - 1)] += 2.0 * (gr @ Bm)
    for c in range(8):
        if G.parity_of(c) == P:
            lam[c] = ubar[c]
        else:
            lam[c] = lam[c] + ubar[c]
    # adjoints at ghost slots are not state
    for c in range(8):
        if not F.cm[c]:
            continue
        mask = np.zeros(F.size[c], dtype=bool)
        mask[_inner(F, c)] = True
        lam[c][~mask] = 0.0


def discrete_pullback(p, si, adv, bnd, I_bar, beam=0):
    """Exact gradient of sum_r I_bar_r I_r of the forward march: returns
    (I, rho_bar[8] (P_c, n_e), gen (P_eee, n_lines)) with gen the sensitivity to W."""
    fwd, _, _ = p.records()
    ne, nl = len(p.elements), len(p.x_rays)
    sig0 = ne + ne * (p.N + 1) + 1
    A, Bf = dense_blocks(p.N, False, adv, bnd)
    F = DenseProblem(p, si, A, Bf, False, beam)
    dE, dvol = p.dE, p.gspec.dvol
    Wl = si.W
    u = F.zeros()
    tape = [[x.copy() for x in u]]
    mids = [None]
    I = np.zeros(nl)
    for k in range(1, p.n_steps + 1):
        F.set_record(fwd[k])
        h = dE / 2
        F.boundary(u, 0)
        du = F.calc_du(u, 1, h, True)
        for c, dd in du.items():
            u[c] += h * dd
        mids.append([x.copy() for x in u])
        F.boundary(u, 1)
        du = F.calc_du(u, 0, dE, True)
        for c, dd in du.items():
            u[c] += dE * dd
        F.boundary(u, 0)
        du = F.calc_du(u, 1, h, True)
        for c, dd in du.items():
            u[c] += h * dd
        tape.append([x.copy() for x in u])
        I += dE * dvol * fwd[k][sig0:] * (Wl * u[0][..., 0].ravel()[:, None]).sum(axis=0)
    lam = F.zeros()
    rb = [np.zeros(F.rho[c].shape) for c in range(8)]
    gen = np.zeros((tape[0][0][..., 0].size, nl))
    for k in range(p.n_steps, 0, -1):
        wk = dE * dvol * fwd[k][sig0:] * I_bar
        lam[0][..., 0] += (Wl @ wk).reshape(lam[0].shape[:3])
        gen += tape[k][0][..., 0].ravel()[:, None] * wk[None, :]
        even_k, mid, prev = tape[k], mids[k], tape[k - 1]
        _adjoint_half(F, lam, _combine(F, even_k, mid), 1, dE / 2, rb, fwd[k])
        _adjoint_half(F, lam, _combine(F, prev, mid), 0, dE, rb, fwd[k])
        _adjoint_half(F, lam, prev, 1, dE / 2, rb, fwd[k])
    return I, [r.reshape(-1, ne) for r in rb], gen
