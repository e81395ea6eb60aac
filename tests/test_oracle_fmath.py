"""Pins of the shared sine / cosine / x^5 sequences (barney_amd/csrc/fmath.h)
that the kernels and the oracle both use in the shading step, against
float64 truth: a wrong coefficient, sign or quadrant fails these bounds.
  * sin and cos within 2 ulp of float64 sin / cos over the arguments the
    shading step makes ([-pi, 2.5 pi]: cosine-hemisphere, GGX and phase
    azimuths, concentric-disk angles) and over random |x| < 2^16;
  * exact odd / even symmetry and the quarter-turn identities at multiples
    of pi/2 (sin = 0, +-1), NaN outside the domain;
  * x^5 within 3 ulp of float64 x^5 on [0, 1] (the Schlick weight's
    domain), exact at 0, 1 and powers of two."""
import numpy as np

from oracle import refcpu as R


def _ulps(got, want):
    want32 = np.abs(want).astype(np.float32)
    unit = np.spacing(np.maximum(want32, np.float32(2.0 ** -24)))
    return np.abs(got.astype(np.float64) - want) / unit.astype(np.float64)


def test_sin_cos_within_2_ulp_on_the_shading_arguments():
    x = np.linspace(-np.pi, 2.5 * np.pi, 2_000_001).astype(np.float32)
    s, c, _ = R.fmath(x)
    xd = x.astype(np.float64)
    es, ec = _ulps(s, np.sin(xd)), _ulps(c, np.cos(xd))
    print(f"max ulp sin {es.max():.3f} cos {ec.max():.3f}")
    assert es.max() <= 2.0 and ec.max() <= 2.0


def test_sin_cos_within_2_ulp_on_the_domain():
    x = np.random.default_rng(3).uniform(-65535.0, 65535.0, 1_000_000).astype(np.float32)
    s, c, _ = R.fmath(x)
    xd = x.astype(np.float64)
    assert _ulps(s, np.sin(xd)).max() <= 2.0 and _ulps(c, np.cos(xd)).max() <= 2.0


def test_symmetries_and_special_values():
    x = np.random.default_rng(4).uniform(0.0, 40.0, 100_000).astype(np.float32)
    s, c, _ = R.fmath(x)
    sn, cn, _ = R.fmath(-x)
    assert np.array_equal(sn, -s) and np.array_equal(cn, c)
    # 0 and the float32 nearest multiples of pi/2: the reduced argument is
    # tiny, so sin/cos are exactly 0/+-1 in the right quadrant
    k = np.arange(-8, 9)
    s, c, _ = R.fmath((k * np.float32(np.pi / 2)).astype(np.float32))
    want_s = np.array([0, 1, 0, -1])[k % 4]
    want_c = np.array([1, 0, -1, 0])[k % 4]
    assert np.all(np.abs(s - want_s) < 1e-6) and np.all(np.abs(c - want_c) < 1e-6)
    s, c, _ = R.fmath(np.array([0.0], np.float32))
    assert s[0] == 0.0 and c[0] == 1.0
    s, c, _ = R.fmath(np.array([np.nan, np.inf, -np.inf, 65536.0, 1e30], np.float32))
    assert np.isnan(s).all() and np.isnan(c).all()


def test_pow5_within_3_ulp():
    x = np.linspace(0.0, 1.0, 1_000_001).astype(np.float32)
    _, _, p = R.fmath(x)
    err = _ulps(p, x.astype(np.float64) ** 5)
    print(f"max ulp pow5 {err.max():.3f}")
    assert err.max() <= 3.0
    _, _, p = R.fmath(np.array([0.0, 1.0, 0.5, 0.25, 2.0], np.float32))
    assert list(p) == [0.0, 1.0, 0.5 ** 5, 0.25 ** 5, 32.0]


def test_acos_within_3_ulp():
    x = np.concatenate([np.linspace(-1.0, 1.0, 2_000_001), np.random.default_rng(5).uniform(-1, 1, 500_000),
                        1.0 - np.logspace(-8, -1, 1000), -1.0 + np.logspace(-8, -1, 1000)]).astype(np.float32)
    _, c = R.fmath_inv(np.zeros_like(x), x)
    err = _ulps(c, np.arccos(x.astype(np.float64)))
    print(f"max ulp acos {err.max():.3f}")
    assert err.max() <= 3.0
    _, c = R.fmath_inv(np.zeros(6, np.float32), np.array([1.0, -1.0, 0.0, 1.0000001, -2.0, np.nan], np.float32))
    assert c[0] == 0.0 and c[1] == np.float32(np.pi) and c[2] == np.float32(np.pi / 2)
    assert np.isnan(c[3:]).all()


def test_atan2_within_3_ulp_and_quadrants():
    rng = np.random.default_rng(6)
    # directions on the unit sphere (the environment map's use) and wide ranges
    v = rng.normal(size=(1_000_000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    ys = [v[:, 1], rng.uniform(-1e3, 1e3, 500_000), np.sin(np.linspace(-np.pi, np.pi, 100_001))]
    xs = [v[:, 0], rng.uniform(-1e-3, 1e-3, 500_000), np.cos(np.linspace(-np.pi, np.pi, 100_001))]
    for y, x in zip(ys, xs):
        y, x = y.astype(np.float32), x.astype(np.float32)
        a, _ = R.fmath_inv(y, x)
        err = _ulps(a, np.arctan2(y.astype(np.float64), x.astype(np.float64)))
        print(f"max ulp atan2 {err.max():.3f}")
        assert err.max() <= 3.0
    # the signed zero and infinity cases of atan2
    y = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, np.inf, np.inf, 1.0, 1.0, 0.0], np.float32)
    x = np.array([0.0, 0.0, -0.0, -0.0, 0.0, -0.0, np.inf, 1.0, np.inf, -np.inf, -1.0], np.float32)
    a, _ = R.fmath_inv(y, x)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(np.float32)
    assert np.array_equal(a, want) and np.array_equal(np.signbit(a), np.signbit(want))
    a, _ = R.fmath_inv(np.array([np.nan, 1.0], np.float32), np.array([1.0, np.nan], np.float32))
    assert np.isnan(a).all()
