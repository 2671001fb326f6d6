"""Seeded synthetic fields for the benchmark and parity workloads.

Random numbers come from MT19937 seeded the way the reference seeds it
(sgrnd: mt[0] = seed, mt[i] = 69069*mt[i-1] mod 2^32,
General_Routines/mod_mt.f90:71-87), drawn through numpy's MT19937 whose state
is set to that key.  grnd() = y / (2^32 - 1) (mod_mt.f90:142-146).  Normal
deviates use Box-Muller on u = (y + 0.5) / 2^32, which is never 0.  Grid point
i (0-based) sits at i*dx, as in Calculate_Laplace's phase convention
(Geometry_Modules/Bulk/mod_fourier.f90:178).  Arrays are (k, nz, ny, nx),
x fastest.
"""
from __future__ import annotations

import numpy as np


def sgrnd_key(seed: int) -> np.ndarray:
    key = np.empty(624, dtype=np.uint64)
    key[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        key[i] = (69069 * int(key[i - 1])) & 0xFFFFFFFF
    return key.astype(np.uint32)


class MTStream:
    """The reference's MT19937 word stream for a seed."""

    def __init__(self, seed: int):
        self.bg = np.random.MT19937()
        self.bg.state = {"bit_generator": "MT19937", "state": {"key": sgrnd_key(seed), "pos": 624}}

    def words(self, n: int) -> np.ndarray:
        return self.bg.random_raw(n).astype(np.uint64)

    def grnd(self, n: int) -> np.ndarray:
        return self.words(n).astype(np.float64) / (2.0 ** 32 - 1.0)

    def normals(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        y = self.words(2 * m).astype(np.float64)
        u1 = (y[0::2] + 0.5) / 2.0 ** 32
        u2 = (y[1::2] + 0.5) / 2.0 ** 32
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.empty(2 * m)
        z[0::2] = r * np.cos(2.0 * np.pi * u2)
        z[1::2] = r * np.sin(2.0 * np.pi * u2)
        return z[:n]


def _coords(n, L):
    return np.arange(n) * (L / n)


def lamellar_seed(n, L, chiN, sigma=0.1, seed=2007):
    """w_A = -w_B = 0.5*chiN*(cos(2*pi*x/L) + eta), eta ~ N(0, sigma^2) per point."""
    x = _coords(n, L)
    eta = MTStream(seed).normals(n ** 3).reshape(n, n, n) * sigma
    wa = 0.5 * chiN * (np.cos(2 * np.pi * x / L)[None, None, :] + eta)
    return np.stack([wa, -wa])


def cylinder_seed(n, L, chiN, sigma=0.2, seed=2007):
    """w_A - w_B = chiN*[(cos(k.x) + cos(k(x/2+sqrt3 y/2)) + cos(k(x/2-sqrt3 y/2)))/3 + eta],
    split as w_A = -w_B; k = 2*pi/L, eta ~ N(0, sigma^2)."""
    k = 2 * np.pi / L
    x = _coords(n, L)[None, None, :]
    y = _coords(n, L)[None, :, None]
    hexa = (np.cos(k * x) + np.cos(k * (x / 2 + np.sqrt(3) * y / 2)) + np.cos(k * (x / 2 - np.sqrt(3) * y / 2))) / 3
    eta = MTStream(seed).normals(n ** 3).reshape(n, n, n) * sigma
    diff = chiN * (hexa + eta)
    return np.stack([0.5 * diff, -0.5 * diff])


def bcc_seed(n, L, chiN, sigma=0.1, seed=2007):
    """BCC sphere seed: sum of cos(k.r) over the 12 {110} wavevectors (2*pi/L)(+-1,+-1,0) and
    permutations, normalised by 12, plus eta ~ N(0, sigma^2); split as w_A = -w_B."""
    k = 2 * np.pi / L
    x = _coords(n, L)[None, None, :]
    y = _coords(n, L)[None, :, None]
    z = _coords(n, L)[:, None, None]
    s = np.zeros((n, n, n))
    for a, b in ((x, y), (y, z), (x, z)):
        for sa in (1, -1):
            for sb in (1, -1):
                s = s + np.cos(k * (sa * a + sb * b))
    eta = MTStream(seed).normals(n ** 3).reshape(n, n, n) * sigma
    diff = chiN * (s / 12.0 + eta)
    return np.stack([0.5 * diff, -0.5 * diff])


def block_segments(fractions, ns):
    """nint(block_length/ds) with ds = 1/ns (Fortran NINT, mod_multiblock.f90:85)."""
    ds = 1.0 / ns
    return [int(np.floor(f / ds + 0.5)) for f in fractions]


def smooth_target(n, sigma=2.0, lo=0.1, hi=0.9, seed=2007):
    """Seeded N(0,1) noise, Gaussian-filtered with a periodic kernel of sigma grid cells,
    rescaled to [lo, hi]; shape (n, n, n)."""
    eta = MTStream(seed).normals(n ** 3).reshape(n, n, n)
    k = np.fft.fftfreq(n) * 2 * np.pi
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    filt = np.exp(-0.5 * sigma ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    sm = np.fft.ifftn(np.fft.fftn(eta) * filt).real
    sm = (sm - sm.min()) / (sm.max() - sm.min())
    return lo + (hi - lo) * sm


def unit_directions(count, sizes, seed=2007):
    """count seeded unit vectors over a concatenated parameter vector of the given sizes."""
    st = MTStream(seed)
    n = int(sum(sizes))
    out = st.normals(count * n).reshape(count, n)
    return out / np.linalg.norm(out, axis=1, keepdims=True)


def batched_instances(seed=2007):
    """The 64-instance parameter sweep: instances 0-15 on 32^3, 16-47 on 48^3, 48-63 on 64^3
    grids.  From one MT19937 stream seeded with `seed`, per instance in index order:
    Ns = [100, 150, 200][floor(3u)], chiN = 12 + 18u, f_A = 0.2 + 0.3u, box = 3.5 + 1.5u
    (four draws).  Fields: lamellar seed with noise N(0, 0.1^2) from seed + 1 + i.
    Returns (instances, fields) for adscf_amd.Batch."""
    st = MTStream(seed)
    inst, w = [], []
    for i in range(64):
        u = st.grnd(4)
        ns = [100, 150, 200][min(int(3 * u[0]), 2)]
        chiN = 12.0 + 18.0 * u[1]
        fa = 0.2 + 0.3 * u[2]
        L = 3.5 + 1.5 * u[3]
        n = 32 if i < 16 else (48 if i < 48 else 64)
        segs = block_segments([fa, 1.0 - fa], ns)
        inst.append(dict(grid=(n, n, n), box=(L, L, L), ds=1.0 / ns, blocks=[1, 2], segments=segs, kappa=25.0,
                         chi=[[0.0, chiN], [chiN, 0.0]], vol=[1.0, 1.0]))
        w.append(lamellar_seed(n, L, chiN, sigma=0.1, seed=seed + 1 + i))
    return inst, w
