"""Pins the Philox random streams.  The reference's CURAND / Mersenne Twister
streams (src/QED2dHMC.jl:25,42,69-70,95) cannot be reproduced; the library uses
Philox4x32-10 (Salmon et al., SC'11), checked against the Random123
known-answer vectors, and Box-Muller with fixed-FMA log / sincos, checked
against libm and by distribution tests."""
import math

import numpy as np
import pytest
from scipy import stats

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(O, ctr, key, out):
    assert O.philox(ctr, key) == out


def test_det_log_accuracy(O):
    rng = np.random.default_rng(0)
    us = np.concatenate([rng.random(20000), 10.0 ** rng.uniform(-300, 0, 2000), [1.0, 0.5]])
    for u in us:
        ref = math.log(u)
        assert abs(O.det_log(u) - ref) <= 4e-16 * max(1.0, abs(ref))


def test_det_sincos_accuracy(O):
    import mpmath as mp
    mp.mp.dps = 40
    rng = np.random.default_rng(1)
    for u in np.concatenate([rng.random(3000), [0.0, 0.125, 0.25, 0.5, 0.75, -0.5, -0.25]]):
        s, c = O.det_sincos2pi(u)
        a = 2 * mp.pi * mp.mpf(float(u))
        assert abs(s - float(mp.sin(a))) < 2.3e-16
        assert abs(c - float(mp.cos(a))) < 2.3e-16


def test_gaussian_momenta_distribution(O):
    mom = O.momenta(512, 512, 1234, 0, 0).ravel()
    assert abs(mom.mean()) < 5 * 1 / math.sqrt(mom.size)
    assert abs(mom.var() - 1.0) < 5 * math.sqrt(2.0 / mom.size)
    assert stats.kstest(mom[:200000], "norm").pvalue > 1e-3
    assert abs(stats.kurtosis(mom)) < 0.02


def test_pseudofermion_distribution(O):
    X = O.pseudofermion(256, 256, 77, 0, 3).ravel()
    # (N + iN) / sqrt 2: each part variance 1/2, |X|^2 mean 1
    assert abs(np.mean(np.abs(X) ** 2) - 1.0) < 0.01
    assert abs(np.var(X.real) - 0.5) < 0.01 and abs(np.var(X.imag) - 0.5) < 0.01
    assert abs(np.mean(X.real * X.imag)) < 0.005


def test_streams_are_distinct(O):
    a = O.momenta(64, 64, 1, 0, 0)
    b = O.momenta(64, 64, 1, 0, 1)      # next trajectory
    c = O.momenta(64, 64, 1, 1, 0)      # another chain
    d = O.momenta(64, 64, 2, 0, 0)      # another seed
    for x in (b, c, d):
        assert abs(np.corrcoef(a.ravel(), x.ravel())[0, 1]) < 0.03
    us = np.array([O.metropolis_uniform(5, 0, t) for t in range(5000)])
    assert stats.kstest(us, "uniform").pvalue > 1e-3


def test_hot_start_uniform_angles(O):
    U = O.hot_start(256, 128, 1234)
    assert np.allclose(np.abs(U), 1.0, atol=1e-15)
    th = np.angle(U).ravel()
    assert stats.kstest((th + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
