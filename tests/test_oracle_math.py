"""Pins the elementary functions and the random stream shared by the oracle
and the kernels (voxeltoy_amd/csrc/vtmath.h) against libm in double and
against PCG32's published known-answer vector."""
import math

import numpy as np
import pytest


def ulp_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    spacing = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    spacing = np.maximum(spacing, np.spacing(np.float32(1e-30)))
    return np.abs(got - ref) / spacing


def test_pcg32_known_answer(orc):
    # pcg32-demo (pcg-c-basic), pcg32_srandom_r(&rng, 42u, 54u)
    ref = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert list(orc.pcg_raw(42, 54, 6)) == ref


def test_pcg32_stream_is_uniform_and_keyed(orc):
    a = orc.pcg_stream(7, 100, 3, 4096)
    b = orc.pcg_stream(7, 101, 3, 4096)
    c = orc.pcg_stream(7, 100, 4, 4096)
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.02
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
    # (x >> 8) * 2^-24: every value is a multiple of 2^-24
    assert np.all((a * 2 ** 24) == np.floor(a * 2 ** 24))


@pytest.mark.parametrize("lo,hi", [(-7.0, 7.0), (0.0, 2 * math.pi), (-100.0, 100.0)])
def test_sincos_accuracy(orc, lo, hi):
    xs = np.linspace(lo, hi, 20001, dtype=np.float32)
    s = np.array([orc.sincos(float(x))[0] for x in xs])
    c = np.array([orc.sincos(float(x))[1] for x in xs])
    x64 = xs.astype(np.float64)
    assert np.max(np.abs(s - np.sin(x64))) < 3e-7 * max(1.0, hi / 10)
    assert np.max(np.abs(c - np.cos(x64))) < 3e-7 * max(1.0, hi / 10)


def test_atan2_accuracy(orc):
    rng = np.random.default_rng(0)
    y = rng.uniform(-1, 1, 20000).astype(np.float32)
    x = rng.uniform(-1, 1, 20000).astype(np.float32)
    got = np.array([orc.lib().orc_atan2f(float(a), float(b)) for a, b in zip(y, x)])
    assert np.max(ulp_err(got, np.arctan2(y.astype(np.float64), x.astype(np.float64)))) <= 4
    assert orc.lib().orc_atan2f(0.0, 1.0) == 0.0
    assert abs(orc.lib().orc_atan2f(1.0, 0.0) - math.pi / 2) < 1e-7
    assert abs(orc.lib().orc_atan2f(0.0, -1.0) - math.pi) < 1e-6


def test_acos_accuracy(orc):
    xs = np.linspace(-1, 1, 20001, dtype=np.float32)
    got = np.array([orc.lib().orc_acosf(float(x)) for x in xs])
    ref = np.arccos(xs.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 5e-7
    assert math.isnan(orc.lib().orc_acosf(1.5))


def test_log2_exp2_pow_accuracy(orc):
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-30, 30, 20000)).astype(np.float32)
    got = np.array([orc.lib().orc_log2f(float(v)) for v in x])
    assert np.max(np.abs(got - np.log2(x.astype(np.float64)))) < 4e-6
    e = rng.uniform(-60, 60, 20000).astype(np.float32)
    got = np.array([orc.lib().orc_exp2f(float(v)) for v in e])
    assert np.max(ulp_err(got, np.exp2(e.astype(np.float64)))) <= 4
    b = rng.uniform(0, 4, 5000).astype(np.float32)
    p = rng.uniform(0.01, 5, 5000).astype(np.float32)
    got = np.array([orc.lib().orc_powf(float(u), float(v)) for u, v in zip(b, p)])
    ref = np.power(b.astype(np.float64), p.astype(np.float64))
    assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-30)) < 2e-5
    # the cases the integrator relies on
    assert orc.lib().orc_powf(0.0, 0.5) == 0.0
    assert orc.lib().orc_powf(3.0, 0.0) == 1.0
    assert math.isnan(orc.lib().orc_powf(-0.5, 1 / 2.2))
    assert orc.lib().orc_powf(2.0, 5.0) == pytest.approx(32.0, rel=1e-6)
