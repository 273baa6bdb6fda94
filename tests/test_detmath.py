"""The portable exp / log / log10 / pow of the rate constants
(tchem-atm_amd/csrc/detmath.h, compiled into the oracle library for the
host) against the C library through numpy: within a few units in the last
place over the ranges rate constants use and far beyond, with the special
values of the C functions.  The device build of the same source is checked
bit for bit against the host through the rate constants
(tests/test_gpu_parity.py)."""
import numpy as np
import pytest

pytestmark = []


def ulps(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


@pytest.fixture(scope="module")
def dm(oracle_mod):
    return oracle_mod.detmath


def test_exp(dm):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-745, 709, 100000), rng.uniform(-60, 60, 100000), rng.uniform(-1e-3, 1e-3, 10000),
                        [0.0, 1.0, -1.0, 0.5 * np.log(2.0), 700.0, -700.0]])
    ref = np.exp(x)
    ok = ref > np.finfo(float).tiny  # normal results
    assert ulps(dm("exp", x[ok]), ref[ok]).max() <= 1.0
    assert dm("exp", [0.0])[0] == 1.0
    out = dm("exp", [800.0, -800.0, np.inf, -np.inf, np.nan])
    assert out[0] == np.inf and out[1] == 0.0 and out[2] == np.inf and out[3] == 0.0 and np.isnan(out[4])


def test_log_and_log10(dm):
    rng = np.random.default_rng(2)
    x = np.concatenate([10.0 ** rng.uniform(-307, 307, 100000), rng.uniform(0.5, 2.0, 100000),
                        rng.uniform(0.999, 1.001, 10000), [1e-300, 3e-5, 0.6, 1.0 + 2 ** -40]])
    ref = np.log(x)
    ok = ref != 0.0
    assert ulps(dm("log", x[ok]), ref[ok]).max() <= 2.0
    assert ulps(dm("log10", x[ok]), np.log10(x[ok])).max() <= 3.0
    assert dm("log", [1.0])[0] == 0.0
    out = dm("log", [0.0, -1.0, np.inf, np.nan])
    assert out[0] == -np.inf and np.isnan(out[1]) and out[2] == np.inf and np.isnan(out[3])


def test_pow(dm):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(220 / 300, 320 / 300, 50000), rng.uniform(0.1, 1.0, 50000)])  # T/300, Fc
    y = np.concatenate([rng.uniform(-4, 4, 50000), rng.uniform(0.05, 1.0, 50000)])
    assert ulps(dm("pow", x, y), x ** y).max() <= 4.0
    assert dm("pow", [2.0], [0.0])[0] == 1.0 and dm("pow", [0.0], [2.0])[0] == 0.0
    assert dm("pow", [0.0], [-1.0])[0] == np.inf
