"""Seeded workloads: the (seed, cell, slot)-keyed generator.

Philox4x32-10 is pinned by the known-answer vectors published with the
Random123 library (Salmon et al., SC'11, kat_vectors); the keyed draws make
any subset of a workload hold exactly the rows of the full workload."""
import importlib

import numpy as np
import pytest

from conftest import golden

S = importlib.import_module("tchem-atm_amd.scenarios")


@pytest.mark.parametrize("ctr,key,out", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox4x32_known_answers(ctr, key, out):
    got = tuple(int(x) for x in S.philox4x32(*ctr, *key))
    assert got == out


def test_keyed_uniform_range_and_independence():
    u = S.keyed_uniform(7, np.arange(4096), np.arange(8))
    assert u.shape == (4096, 8) and u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    # different seeds, cells and slots give different streams
    assert not np.array_equal(u, S.keyed_uniform(8, np.arange(4096), np.arange(8)))
    assert np.corrcoef(u[:, 0], u[:, 1])[0, 1] < 0.05


@pytest.fixture(scope="module")
def kmd_base(pkg):
    kmd = pkg.KineticModelData(golden("cb05cl_ae5.yaml"))
    return kmd, pkg.setScenarioConditions(golden("cb05cl_ae5.yaml"), kmd)


@pytest.mark.parametrize("which", ["config0", "config1"])
def test_subsets_are_rows_of_the_full_workload(kmd_base, which):
    kmd, base = kmd_base
    make = getattr(S, which)
    full = make(kmd, base[0], n=2048)
    idx = np.array([0, 5, 17, 1024, 2047])
    sub = make(kmd, base[0], cells=idx)
    np.testing.assert_array_equal(sub.state, full.state[idx])
    np.testing.assert_array_equal(sub.photo, full.photo[idx])
    strided = make(kmd, base[0], cells=np.arange(0, 2048, 64))
    np.testing.assert_array_equal(strided.state, full.state[::64])


def test_config1_distributions(kmd_base):
    kmd, base = kmd_base
    m = kmd.nActive
    w = S.config1(kmd, base[0], n=8192)
    p, T = w.state[:, 1], w.state[:, 2]
    assert 5.0e4 <= p.min() and p.max() <= 1.013e5 and 220.0 <= T.min() and T.max() <= 320.0
    ppb = w.state[:, 3:3 + m] * 1e3
    assert 1e-3 <= ppb.min() and ppb.max() <= 1e2
    assert abs(np.log10(ppb).mean() - (-0.5)) < 0.02
    night = S.night_flags(8192, 1, m, kmd.nPhoto)
    assert np.all(w.photo[night] == 0.0) and np.all(w.photo[~night] > 0.0)
    assert abs(night.mean() - 0.5) < 0.02
    np.testing.assert_array_equal(w.state[:, 3 + m:], np.broadcast_to(base[0, 3 + m:], (8192, kmd.nConstSpec)))
