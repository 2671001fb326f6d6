"""CPU checks of the C-ABI boundary: the library builds, loads without a GPU and
exports exactly what include/adscf.h declares; host-only Iterate routines agree
with the oracle."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "adscf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(adscf_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    import adscf_amd

    if not os.path.exists(adscf_amd.LIB_PATH):
        from adscf_amd import build

        build.build()
    return adscf_amd.load_library()


def test_header_symbols_exported(lib):
    import adscf_amd

    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding types exactly the declared surface
    assert sorted(adscf_amd.SIGNATURES) == names


def test_init_without_gpu_fails_loudly(lib):
    import adscf_amd
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(adscf_amd.AdscfError):
        adscf_amd.Context(8, 8, 8, 1.0, 1.0, 1.0, 0.01)


def test_host_lu_matches_oracle(lib, orc):
    import adscf_amd

    rng = np.random.default_rng(0)
    for n in (1, 3, 10, 20, 40, 70):
        a = rng.standard_normal((n, n))
        b = rng.standard_normal(n)
        assert np.array_equal(adscf_amd.solve_lu(a, b), orc.solve_lu(a, b))  # same operation order
        lu1, i1, d1 = adscf_amd.ludcmp(a)
        lu2, i2, d2, _ = orc.ludcmp(a)
        assert np.array_equal(lu1, lu2) and np.array_equal(i1, i2) and d1 == d2
        spd = a @ a.T + n * np.eye(n)
        assert np.allclose(adscf_amd.solve_cholesky(spd, b), orc.solve_cholesky(spd, b), rtol=1e-13)
    # duplicate rows: tiny-pivot substitution, same result as the oracle
    a = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0], [0.5, 1.0, 4.0]])
    lu1, _, _ = adscf_amd.ludcmp(a)
    lu2, _, _, _ = orc.ludcmp(a)
    assert np.array_equal(lu1, lu2)
