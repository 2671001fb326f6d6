"""Pins the numpy restatement of the adjoint (oracle/np_adjoint.py) by central
finite differences of its own forward model, on a small grid with odd and even
axes and a three-block chain (two junctions)."""
import numpy as np

from oracle import np_adjoint


def test_np_adjoint_matches_central_differences():
    rng = np.random.default_rng(0)
    nx, ny, nz = 6, 5, 4
    bt, bs, ds = [1, 2, 1], [3, 4, 2], 0.05
    chi = [[0.0, 12.0], [12.0, 0.0]]
    w = rng.standard_normal((2, nz, ny, nx))
    phit = rng.uniform(0.1, 0.9, (nz, ny, nx))
    L = np.array([2.0, 1.7, 1.5])
    args = (bt, bs, ds, 25.0, chi, [1.0, 0.8], phit, 0.1)
    loss, F, gw, gb = np_adjoint.adjoint(w, L, *args)
    h = 1e-6
    for k in range(12):
        d = rng.standard_normal(w.size + 3)
        dw, db = d[: w.size].reshape(w.shape), d[w.size:]
        fd = (np_adjoint.forward(w + h * dw, L + h * db, *args)[0] -
              np_adjoint.forward(w - h * dw, L - h * db, *args)[0]) / (2 * h)
        ad = (gw * dw).sum() + gb @ db
        assert abs(fd - ad) < 1e-6 * abs(ad)
    for a in range(3):
        e = np.zeros(3)
        e[a] = 1.0
        fd = (np_adjoint.forward(w, L + h * e, *args)[0] - np_adjoint.forward(w, L - h * e, *args)[0]) / (2 * h)
        assert abs(fd - gb[a]) < 1e-6 * abs(gb[a])
