"""The oracle's EPD field steps of Main_Blend (main_blend.f90:580-596, 678-701, 709-740),
pinned without retyping them:

* Calculate_Field_Derivative is -laplacian(newfield - field) (laplace = -q^2,
  mod_fourier.f90:187, and :735 multiplies by -laplace): checked against numpy's spectral
  Laplacian of the oracle's own mean fields (pinned in test_oracle.py) of its blend densities;
* EPD_Timestep_Euler is one explicit step of that derivative;
* EPD_Timestep_RungeKutta is fourth-order accurate: halving the step divides its error
  against a fine-step solution by ~16, where Euler's falls by ~2 (a wrong stage point or
  weight drops the order and fails this).
"""
import numpy as np
import pytest


def _setup(orc):
    g = orc.Grid(32, 8, 1, 6.4, 1.6, 1.0, 0.1)
    bl = orc.Blend([1, 2], [10, 10], [0.5, 0.5])
    fh = orc.FH(20.0, [[0.0, 2.0], [2.0, 0.0]], [1.0, 1.0])
    x = np.arange(32) / 32
    rng = np.random.default_rng(1)
    w = np.zeros((2,) + g.shape)
    w[0] += 0.3 * np.cos(2 * np.pi * x)[None, None, :] + 0.05 * rng.standard_normal(g.shape)
    w[1] -= 0.3 * np.cos(2 * np.pi * x)[None, None, :]
    return g, bl, fh, w


def _minus_laplacian(a, lengths):
    """-laplacian of a (nz, ny, nx) periodic array, spectrally with numpy."""
    nz, ny, nx = a.shape
    lx, ly, lz = lengths
    kx = 2 * np.pi * np.fft.rfftfreq(nx, d=lx / nx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=ly / ny)
    kz = 2 * np.pi * np.fft.fftfreq(nz, d=lz / nz)
    k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    return np.fft.irfftn(k2 * np.fft.rfftn(a), s=a.shape, axes=(0, 1, 2))


def test_field_derivative_is_minus_laplacian_of_mean_field_change(orc):
    g, bl, fh, w = _setup(orc)
    d = orc.field_derivative(g, bl, fh, w)
    mu = orc.mean_fields(g, fh, orc.blend_densities(g, bl, w)) - w
    ref = np.stack([_minus_laplacian(mu[t], (g.lx, g.ly, g.lz)) for t in range(2)])
    assert np.max(np.abs(d - ref)) < 1e-11 * np.max(np.abs(ref))


def test_field_derivative_of_uniform_field_vanishes(orc):
    g, bl, fh, _ = _setup(orc)
    w = np.stack([np.full(g.shape, 0.4), np.full(g.shape, -0.2)])
    assert np.max(np.abs(orc.field_derivative(g, bl, fh, w))) < 1e-12


def test_euler_step_is_one_explicit_step(orc):
    g, bl, fh, w = _setup(orc)
    dt = 1e-4
    got = orc.epd_timestep(g, bl, fh, w, dt, "euler")
    np.testing.assert_allclose(got, w + dt * orc.field_derivative(g, bl, fh, w), rtol=0, atol=1e-15)


def _run(orc, g, bl, fh, w, method, dt, n):
    f = w.copy()
    for _ in range(n):
        f = orc.epd_timestep(g, bl, fh, f, dt, method)
    return f


def test_convergence_orders(orc):
    g, bl, fh, w = _setup(orc)
    T = 2e-3
    ref = _run(orc, g, bl, fh, w, "rk4", T / 128, 128)
    err = {m: [np.max(np.abs(_run(orc, g, bl, fh, w, m, T / n, n) - ref)) for n in (4, 8, 16)]
           for m in ("rk4", "euler")}
    r4 = [err["rk4"][i] / err["rk4"][i + 1] for i in range(2)]
    r1 = [err["euler"][i] / err["euler"][i + 1] for i in range(2)]
    print("rk4 errors", err["rk4"], "ratios", r4)
    print("euler errors", err["euler"], "ratios", r1)
    assert 12.0 < r4[1] < 24.0 and r4[0] > r4[1] * 0.9
    assert all(1.7 < r < 2.5 for r in r1)
