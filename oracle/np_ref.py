"""Second, independent restatement of the reference path in numpy -- TEST INFRASTRUCTURE ONLY.

It shares no code with ``scft_oracle.c``: the FFT is numpy's pocketfft
(``rfftn``/``irfftn`` with ``norm="forward"``, i.e. forward divided by Ng and
an unnormalised inverse, as in mod_fourier.f90:75-105), the loops are
vectorised, and Anderson mixing solves its system with LAPACK.  Agreement of
the two restatements pins each against a plausible slip in the other.
Arrays use the (k, nz, ny, nx) layout described in ``oracle/__init__.py``.
"""
from __future__ import annotations

import numpy as np


def wavenumbers(nx, ny, nz, lx, ly, lz):
    """k-vectors per Calculate_Laplace (mod_fourier.f90:175-186): x never folded,
    y/z folded only for 1-based index > n/2+1 (so a Nyquist index stays positive)."""
    ix = np.arange(nx // 2 + 1)
    iy = np.arange(ny)
    iz = np.arange(nz)
    qx = 2 * np.pi * ix / lx
    qy = 2 * np.pi * np.where(iy + 1 > ny // 2 + 1, iy - ny, iy) / ly
    qz = 2 * np.pi * np.where(iz + 1 > nz // 2 + 1, iz - nz, iz) / lz
    return qx[None, None, :], qy[None, :, None], qz[:, None, None]


def k2(nx, ny, nz, lx, ly, lz):
    qx, qy, qz = wavenumbers(nx, ny, nz, lx, ly, lz)
    return qx * qx + qy * qy + qz * qz


def r2c(a):
    return np.fft.rfftn(a, norm="forward")


def c2r(c, shape):
    return np.fft.irfftn(c, s=shape, axes=(0, 1, 2), norm="forward")


def step(q, E, K):
    """Propagate_Step (mod_propagate.f90:56-64)."""
    return E * c2r(K * r2c(E * q), q.shape)


def chain(field, block_types, block_segs, L, ds, ensemble=1, rho=1.0, mu=0.0, keep_q=False):
    """Calculate_Expfield + Propagate_Multiblock (mod_multiblock.f90:112-188)."""
    ntypes, nz, ny, nx = field.shape
    lx, ly, lz = L
    K = np.exp(-ds * k2(nx, ny, nz, lx, ly, lz))
    E = np.exp(-0.5 * ds * field)
    ns = int(np.sum(block_segs))
    vol = lx * ly * lz
    dvol = vol / (nx * ny * nz)
    qq = np.empty((ns + 1, nz, ny, nx))
    qd = np.empty_like(qq)
    for store, order in ((qq, range(len(block_types))), (qd, range(len(block_types) - 1, -1, -1))):
        p = np.ones((nz, ny, nx))
        store[0] = p
        s = 0
        for b in order:
            for _ in range(block_segs[b]):
                p = step(p, E[block_types[b] - 1], K)
                s += 1
                store[s] = p
    Q = dvol * qq[ns].sum()
    dist = np.empty((len(block_types), nz, ny, nx))
    s0 = 0
    for b in range(len(block_types)):
        n = block_segs[b]
        w = np.full(n + 1, ds)
        w[0] = w[-1] = 0.5 * ds
        idx = np.arange(s0, s0 + n + 1)
        dist[b] = np.tensordot(w, qq[idx] * qd[ns - idx], axes=1)
        s0 += n
    if ensemble == 1:
        dens = rho * vol * dist / Q
        mu = np.log(rho * vol / Q)
        energy = rho * vol * (mu - 1.0)
    elif ensemble == 2:
        dens = np.exp(mu) * dist
        rho = dens.sum() * dvol / vol / (ds * ns)
        energy = -rho * vol
    else:
        raise ValueError("Invalid ensemble")
    out = dict(distribution=dist, density=dens, Qsum=Q, mu=mu, rho=rho, energy=energy)
    if keep_q:
        out.update(q=qq, qdag=qd)
    return out


def species(dens_blocks, block_types, ntypes):
    out = np.zeros((ntypes,) + dens_blocks.shape[1:])
    for b, t in enumerate(block_types):
        out[t - 1] += dens_blocks[b]
    return out


def mean_fields(density, kappa, chi, vol):
    """Calculate_Mean_Fields (mod_FloryHuggins.f90:104-129)."""
    chi = np.asarray(chi)
    vf = density * np.asarray(vol)[:, None, None, None]
    tot = vf.sum(axis=0)
    return np.stack([(kappa * (tot - 1) + np.tensordot(chi[c], vf, axes=1)) * vol[c] for c in range(len(vol))])


def interaction_energy(density, kappa, chi, vol, dvol):
    """Interaction_Energy (mod_FloryHuggins.f90:74-96)."""
    vf = density * np.asarray(vol)[:, None, None, None]
    e = 0.5 * kappa * dvol * ((vf.sum(axis=0) - 1.0) ** 2).sum()
    for i in range(len(vol)):
        for j in range(i + 1, len(vol)):
            e += chi[i][j] * dvol * (vf[i] * vf[j]).sum()
    return e


def anderson(field, dfield, lam_mix, it, lam00=0.001):
    """Mixing_Anderson (mod_iterate.f90:92-150) on (m+1, n) history arrays, in place."""
    m = field.shape[0] - 1
    if it > m:
        dd = dfield[0][None, :] - dfield[1:]
        v = dd @ dfield[0]
        u = dd @ dd.T
        x = np.linalg.solve(u, v)
        fmix = field[0] - x @ (field[0][None, :] - field[1:])
        dmix = dfield[0] - x @ dd
        lam = lam_mix
    else:
        fmix = field[0].copy()
        dmix = dfield[0].copy()
        lam = lam00
    field[1:] = field[:-1].copy()
    dfield[1:] = dfield[:-1].copy()
    field[0] = fmix + lam * dmix
