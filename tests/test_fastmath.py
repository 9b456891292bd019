"""The kernel's FP64 logarithm (hpc_ode_amd/csrc/fastmath.hpp) against the C library.

fast_log replaces the library log in the RHS (ln T), the Troe falloff (log10 Fc, log10 Pr,
the falloff form of pyJac's rxn_rates_pres_mod.c:5-366) and the BDF step-size powers. The
same header is compiled for the host here; the GPU parity tests then pin the device RHS and
Jacobian that use it against the oracle and pyJac.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def results():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "test_fastmath")
        subprocess.run([gxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "hpc_ode_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "test_fastmath.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe, "2000000"], check=True, capture_output=True, text=True).stdout
    return {k: float(v) for k, v in (line.split() for line in out.splitlines())}


def test_log_within_two_ulp(results):
    assert results["log_bits"] <= 2.0
    assert results["log_reduced"] <= 2.0


def test_log10_and_pow_accuracy(results):
    # log10 = log * (1/ln 10): one more rounding; pow = exp(y log x) with |y log x| < 7
    assert results["log10"] <= 3.0
    assert results["pow"] <= 8.0


def test_special_values(results):
    assert results["log0"] == float("-inf")
    assert results["loginf"] == float("inf")
    assert results["logneg"] != results["logneg"]      # NaN
    assert results["lognan"] != results["lognan"]
    assert results["pow0"] == 0.0
    assert results["tab_pow0"] == 0.0   # a zero error estimate: eta = 1 / (0 + ADDON), as CVODE


def test_exp_within_two_ulp(results):
    """tab_exp, the exponential of the rate constants, equilibrium constants, Troe terms and
    step-size powers, against the C library's exp."""
    assert results["exp"] <= 2.0
    assert results["exp_small"] <= 2.0
    assert results["exp_subnormal"] <= 1.0      # in units of the smallest subnormal
    assert results["tab_pow"] <= 8.0


def test_exp_special_values(results):
    assert results["exp0"] == 1.0
    assert results["exp_big"] == float("inf") and results["exp_tiny"] == 0.0
    assert 1.79e308 < results["exp_max"] < float("inf")
    assert results["exp_nan"] != results["exp_nan"]
    # the largest arguments the kernels pass (|x| <= 1e13: bounded Troe coefficients)
    assert results["exp_neg_huge"] == 0.0 and results["exp_pos_huge"] == float("inf")
    # infinite arguments are outside the domain
    assert results["exp_neginf"] != results["exp_neginf"] and results["exp_posinf"] != results["exp_posinf"]
