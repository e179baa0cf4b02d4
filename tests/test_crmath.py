"""The library's correctly rounded trigonometry (the sin, cos, asin and atan2 behind GeoMath's positions,
GeoMath.cs:49-80), evaluated on the host through the C ABI, against libquadmath's 113-bit functions
rounded to double (the oracle's or_sin / or_cos / or_asin / or_atan2): equal on every argument. The
platform's own functions (glibc) differ from both on a fraction of the same arguments, so the test tells
a correctly rounded implementation from the C library's."""
import numpy as np
import pytest


def _arguments(rng, n):
    """Angles of regional grids (radians of lon 13-14, lat 43-44 degrees), wide ranges, small values,
    and the doubles nearest to k pi/2 (the reduction's hardest cases)."""
    near = []
    for k in range(1, 651):
        x = k * (np.pi / 2)
        for d in range(-3, 4):
            near.append(x + d * np.spacing(x))
    return {
        "regional": np.concatenate([np.deg2rad(rng.uniform(13, 14, n)), np.deg2rad(rng.uniform(43, 44, n))]),
        "wide": rng.uniform(-7, 7, n),
        "reduced": rng.uniform(-1024, 1024, n),
        "small": np.ldexp(rng.uniform(0.5, 1, n), -rng.integers(0, 40, n)),
        "near_k_pi_2": np.array(near),
    }


def _asin_arguments(rng, n):
    return {
        "regional": np.sin(np.deg2rad(rng.uniform(43, 44, n))),
        "wide": rng.uniform(-1, 1, n),
        "near_one": 1 - np.ldexp(rng.uniform(0, 1, n), -rng.integers(1, 52, n)),
        "small": np.ldexp(rng.uniform(0.5, 1, n), -rng.integers(0, 40, n)),
        "exact": np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5]),
    }


def _atan2_arguments(rng, n):
    return {
        "regional": (rng.uniform(1000, 1500, n), rng.uniform(4500, 4800, n)),
        "wide": (rng.uniform(-5000, 5000, n), rng.uniform(-5000, 5000, n)),
        "skewed": (np.ldexp(rng.uniform(-1, 1, n), -rng.integers(0, 60, n)), rng.uniform(-1, 1, n)),
        "diagonal": (np.full(n, 3.0), 3.0 * (1 + np.ldexp(rng.uniform(-1, 1, n), -rng.integers(10, 52, n)))),
    }


def test_sin_cos_are_correctly_rounded(oracle):
    from tomostar_amd.grid import trig

    rng = np.random.default_rng(1)
    libm_differs = 0
    for name, x in _arguments(rng, 60_000).items():
        for f in ("sin", "cos"):
            got, want = trig(f, x), oracle.trig(f, x)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, f"{f} {name}: {bad.size} of {x.size}, e.g. x={x[bad[0]]!r}"
            libm_differs += int(np.sum(oracle.trig(f, x, correct=False) != want))
    assert libm_differs > 0


def test_asin_is_correctly_rounded(oracle):
    from tomostar_amd.grid import trig

    rng = np.random.default_rng(2)
    libm_differs = 0
    for name, z in _asin_arguments(rng, 60_000).items():
        got, want = trig("asin", z), oracle.trig("asin", z)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, f"asin {name}: {bad.size} of {z.size}, e.g. z={z[bad[0]]!r}"
        assert np.array_equal(np.signbit(got), np.signbit(want))
        libm_differs += int(np.sum(oracle.trig("asin", z, correct=False) != want))
    assert libm_differs > 0


def test_atan2_is_correctly_rounded(oracle):
    from tomostar_amd.grid import trig

    rng = np.random.default_rng(3)
    libm_differs = 0
    for name, (y, x) in _atan2_arguments(rng, 60_000).items():
        got, want = trig("atan2", y, x), oracle.trig("atan2", y, x)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, f"atan2 {name}: {bad.size} of {y.size}, e.g. y={y[bad[0]]!r} x={x[bad[0]]!r}"
        libm_differs += int(np.sum(oracle.trig("atan2", y, x, correct=False) != want))
    assert libm_differs > 0


def test_special_values_follow_the_c_library():
    from tomostar_amd.grid import trig

    y = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, np.inf, 1.0, np.nan])
    x = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, np.inf, np.inf, 1.0])
    got, want = trig("atan2", y, x), np.arctan2(y, x)
    assert np.array_equal(got, want, equal_nan=True) and np.array_equal(np.signbit(got), np.signbit(want))
    assert np.isnan(trig("asin", np.array([1.5, np.nan]))).all()
    assert np.isnan(trig("sin", np.array([np.inf]))).all()


@pytest.mark.gpu
def test_device_trigonometry_equals_the_host(device, oracle):
    """The same functions compiled for gfx950: equal to the correctly rounded values on the device."""
    import ctypes as C

    from tomostar_amd import _lib
    from tomostar_amd.grid import trig

    trig("sin", np.zeros(1))  # declares the entry point
    L = _lib.load()
    rng = np.random.default_rng(4)

    def dev(func, a, b=None):
        a = np.ascontiguousarray(a, np.float64)
        out = np.full_like(a, np.nan)
        bb = None if b is None else np.ascontiguousarray(b, np.float64)
        _lib.check(L.ts_crmath_eval(device.handle, func, a.size, _lib.ptr(a), _lib.ptr(bb), _lib.ptr(out)))
        return out

    for name, x in _arguments(rng, 200_000).items():
        for f, code in (("sin", 0), ("cos", 1)):
            assert np.array_equal(dev(code, x), oracle.trig(f, x)), (f, name)
    for name, z in _asin_arguments(rng, 200_000).items():
        assert np.array_equal(dev(2, z), oracle.trig("asin", z)), ("asin", name)
    for name, (y, x) in _atan2_arguments(rng, 200_000).items():
        assert np.array_equal(dev(3, y, x), oracle.trig("atan2", y, x)), ("atan2", name)
