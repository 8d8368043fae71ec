"""Pins the oracle's FP32 definitions of the OpenCL built-ins against libm.

OpenCL 1.2 (section 7.4) only bounds these built-ins: atan2 <= 6 ulp,
atan <= 5 ulp, acos <= 4 ulp, and C99 Annex F special values.  The oracle's
definitions (oracle/ofp32.h) must stay inside those bounds; a dropped
polynomial term, a wrong reduction constant or a wrong quadrant fails here.
"""
import math

import numpy as np
import pytest

from oracle import bridge


def _ulp(got: np.ndarray, ref: np.ndarray) -> np.ndarray:
    ref32 = ref.astype(np.float32)
    sp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / sp


@pytest.fixture(scope="module")
def rng():
    return np.random.default_rng(2024)


def _mixed(rng, n):
    return (rng.standard_normal(n) * rng.choice([1e-6, 1e-3, 1.0, 50.0, 1e4, 1e20], n)).astype(np.float32)


def test_atan2_accuracy(rng):
    n = 1_000_000
    y, x = _mixed(rng, n), _mixed(rng, n)
    got = bridge.eval_f32("atan2", y, x)
    err = _ulp(got, np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert err.max() <= 3.0, err.max()


def test_atan_accuracy(rng):
    a = _mixed(rng, 1_000_000)
    got = bridge.eval_f32("atan", a)
    err = _ulp(got, np.arctan(a.astype(np.float64)))
    assert err.max() <= 3.0, err.max()


def test_acos_accuracy(rng):
    a = np.concatenate([rng.uniform(-1, 1, 1_000_000), rng.uniform(0.99, 1.0, 100_000),
                        rng.uniform(-1.0, -0.99, 100_000)]).astype(np.float32)
    got = bridge.eval_f32("acos", a)
    err = _ulp(got, np.arccos(a.astype(np.float64)))
    assert err.max() <= 2.0, err.max()


def test_atan2_quadrants_and_dense_angles():
    # points on a circle at known angles: every quadrant and both reductions
    ang = np.linspace(-math.pi, math.pi, 200_001)
    y = np.sin(ang).astype(np.float32)
    x = np.cos(ang).astype(np.float32)
    got = bridge.eval_f32("atan2", y, x)
    err = _ulp(got, np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert err.max() <= 3.0


def test_atan2_special_values():
    pi = np.float32(math.pi)
    cases = [
        (0.0, 1.0, 0.0), (-0.0, 1.0, -0.0), (0.0, -1.0, pi), (-0.0, -1.0, -pi),
        (0.0, 0.0, 0.0), (-0.0, 0.0, -0.0), (0.0, -0.0, pi), (-0.0, -0.0, -pi),
        (1.0, 0.0, np.float32(math.pi / 2)), (-1.0, 0.0, -np.float32(math.pi / 2)),
        (1.0, np.inf, 0.0), (-1.0, np.inf, -0.0), (1.0, -np.inf, pi), (-1.0, -np.inf, -pi),
        (np.inf, 1.0, np.float32(math.pi / 2)), (np.inf, np.inf, np.float32(math.pi / 4)),
        (np.inf, -np.inf, np.float32(3 * math.pi / 4)),
    ]
    y = np.array([c[0] for c in cases], np.float32)
    x = np.array([c[1] for c in cases], np.float32)
    want = np.array([c[2] for c in cases], np.float32)
    got = bridge.eval_f32("atan2", y, x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (got, want)
    assert np.isnan(bridge.eval_f32("atan2", np.array([np.nan], np.float32),
                                    np.array([1.0], np.float32)))[0]


def test_acos_domain_and_endpoints():
    a = np.array([1.0, -1.0, 0.0, 1.0000001, -1.0000001, np.nan], np.float32)
    got = bridge.eval_f32("acos", a)
    assert got[0] == 0.0
    assert got[1] == np.float32(math.pi)
    assert got[2] == np.float32(math.pi / 2)
    assert np.isnan(got[3]) and np.isnan(got[4]) and np.isnan(got[5])


def test_sign_opencl_semantics():
    a = np.array([3.0, -2.0, 0.0, -0.0, np.nan, 1e-45, -1e-45], np.float32)
    got = bridge.eval_f32("sign", a)
    want = np.array([1.0, -1.0, 0.0, -0.0, 0.0, 1.0, -1.0], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_div_sqrt_are_ieee(rng):
    a, b = _mixed(rng, 200_000), _mixed(rng, 200_000)
    assert np.array_equal(bridge.eval_f32("div", a, b).view(np.uint32),
                          (a / b).view(np.uint32))
    s = np.abs(a)
    assert np.array_equal(bridge.eval_f32("sqrt", s).view(np.uint32), np.sqrt(s).view(np.uint32))
