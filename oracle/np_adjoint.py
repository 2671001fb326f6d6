"""Reverse-mode derivative of one canonical density evaluation, restated in numpy --
TEST INFRASTRUCTURE ONLY (checks adscf_adjoint; no reference counterpart).

Loss L = sum_r (phi_t(r) - phi_target(r))^2 + lam * F with F of
Calculate_Energy (main_multiblock.f90:548-557) and phi, Q from
Propagate_Multiblock (mod_multiblock.f90:112-188), canonical ensemble, monomer
volumes v, Helfand kappa and chi (mod_FloryHuggins.f90:74-129).  All propagators
are stored (no checkpoints); pinned by central finite differences in
tests/test_adjoint_math.py.  Arrays use the (k, nz, ny, nx) layout.
"""
from __future__ import annotations

import numpy as np

from oracle import np_ref


def _lop(x, L, ds):
    nz, ny, nx = x.shape
    K = np.exp(-ds * np_ref.k2(nx, ny, nz, *L))
    return np_ref.c2r(K * np_ref.r2c(x), x.shape)


def _dop(x, L, a):
    nz, ny, nx = x.shape
    q = np_ref.wavenumbers(nx, ny, nz, *L)[a] ** 2 * np.ones((nz, ny, nx // 2 + 1))
    return np_ref.c2r(q * np_ref.r2c(x), x.shape)


def _types(block_types, block_segs, chain):
    order = range(len(block_types)) if chain == 0 else range(len(block_types) - 1, -1, -1)
    t = [None]
    for b in order:
        t += [block_types[b] - 1] * block_segs[b]
    return t


def _weights(block_segs, ds, s):
    out, s0 = [], 0
    for b, n in enumerate(block_segs):
        if s0 <= s <= s0 + n:
            out.append((b, ds / 2 if s in (s0, s0 + n) else ds))
        s0 += n
    return out


def forward(w, L, block_types, block_segs, ds, kappa, chi, vol, phit, lam, rho=1.0, target=0):
    nt, nz, ny, nx = w.shape
    ns = int(sum(block_segs))
    E = np.exp(-0.5 * ds * w)
    chains = []
    for c in (0, 1):
        t = _types(block_types, block_segs, c)
        q = [np.ones((nz, ny, nx))]
        for s in range(1, ns + 1):
            q.append(E[t[s]] * _lop(E[t[s]] * q[-1], L, ds))
        chains.append(q)
    q, p = chains
    V = float(np.prod(L))
    ng = nx * ny * nz
    dvol = V / ng
    S = q[ns].sum()
    A = rho * ng / S
    phi = np.zeros_like(w)
    for s in range(ns + 1):
        for b, wt in _weights(block_segs, ds, s):
            phi[block_types[b] - 1] += A * wt * q[s] * p[ns - s]
    vf = phi * np.asarray(vol)[:, None, None, None]
    tot = vf.sum(0)
    eint = 0.5 * kappa * dvol * ((tot - 1) ** 2).sum()
    for i in range(nt):
        for j in range(i + 1, nt):
            eint += chi[i][j] * dvol * (vf[i] * vf[j]).sum()
    F = rho * V * (np.log(rho * V / (dvol * S)) - 1) - dvol * (phi * w).sum() + eint
    loss = ((phi[target] - phit) ** 2).sum() + lam * F
    return loss, dict(q=q, p=p, E=E, S=S, A=A, phi=phi, F=F, dvol=dvol, V=V, ns=ns)


def adjoint(w, L, block_types, block_segs, ds, kappa, chi, vol, phit, lam, rho=1.0, target=0):
    """Returns (loss, F, dL/dw, dL/dL_box)."""
    loss, f = forward(w, L, block_types, block_segs, ds, kappa, chi, vol, phit, lam, rho, target)
    q, p, E, S, A, phi, F, dvol, V, ns = (f[k] for k in ("q", "p", "E", "S", "A", "phi", "F", "dvol", "V", "ns"))
    nt = w.shape[0]
    vol = np.asarray(vol)
    vf = phi * vol[:, None, None, None]
    tot = vf.sum(0)
    g = np.zeros_like(phi)
    for t in range(nt):
        de = kappa * (tot - 1) * vol[t] + sum(chi[t][j] * vol[t] * vf[j] for j in range(nt) if j != t)
        g[t] = lam * dvol * (de - w[t])
    g[target] += 2 * (phi[target] - phit)
    dLdS = -(g * phi).sum() / S - lam * rho * V / S

    def seed(s):
        a = np.zeros(w.shape[1:])
        for b, wt in _weights(block_segs, ds, s):
            a += wt * g[block_types[b] - 1]
        return A * a

    tq, tp = _types(block_types, block_segs, 0), _types(block_types, block_segs, 1)
    mu = seed(ns) * p[0] + dLdS
    nu = seed(0) * q[0]
    gE = np.zeros_like(w)
    stress = np.zeros(3)
    for s in range(ns, 0, -1):
        Eq, Ep = E[tq[s]], E[tp[s]]
        y, yp = _lop(Eq * mu, L, ds), _lop(Ep * nu, L, ds)
        gE[tq[s]] += mu * q[s] / Eq + q[s - 1] * y
        gE[tp[s]] += nu * p[s] / Ep + p[s - 1] * yp
        for a in range(3):
            stress[a] += ((Eq * mu) * _dop(q[s] / Eq, L, a)).sum() + ((Ep * nu) * _dop(p[s] / Ep, L, a)).sum()
        mu, nu = seed(s - 1) * p[ns - s + 1] + Eq * y, seed(ns - s + 1) * q[ns - s + 1] + Ep * yp
    gw = -0.5 * ds * E * gE - lam * dvol * phi
    gb = np.array([2 * ds / L[a] * stress[a] + lam * F / L[a] for a in range(3)])
    return loss, F, gw, gb
