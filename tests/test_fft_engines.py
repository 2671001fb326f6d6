"""CPU host emulation of the device line-FFT engines: the LDS Stockham passes
(adscf_amd/csrc/fft_line.hpp), the register-resident four-step operator
(adscf_amd/csrc/fft_reg.hpp) and the R^3 cube operator (adscf_amd/csrc/fft_cube.hpp)
are compiled for the host with g++ and checked against direct long-double DFTs for
every supported length."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("src", ["fft_line_check.cpp", "fft_reg_check.cpp", "fft_cube_check.cpp"])
def test_engine_host_emulation(src, tmp_path):
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "cpu", src), "-o", str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "BAD" not in res.stdout
