"""Pins the binary32 transcendentals that the kernels and the oracle share
(strelka_amd/csrc/fmath.h) against binary64 numpy: every result within one
binary32 ulp of the correctly rounded value (x^5 within three: it rounds three
products), NaN outside the domain, and the identities a wrong quadrant, a
dropped series term or a swapped branch would break."""
import numpy as np

from oracle.pyoracle import fmath

SIN, COS, ACOS, POW5 = 0, 1, 2, 3


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance in binary32 steps between two float32 arrays (same sign
    ordering as the integers of their bit patterns)."""
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def sample_angles(rng):
    parts = [
        rng.uniform(-4 * np.pi, 4 * np.pi, 1_500_000),
        rng.uniform(0, 2 * np.pi, 500_000),  # the BSDF's phi = 2 pi u
        rng.uniform(-1e5, 1e5, 200_000),
        np.ldexp(rng.uniform(0.5, 1.0, 100_000), rng.integers(-40, 0, 100_000)),
    ]
    # floats next to multiples of pi/4, where a reduction error shows most
    k = np.arange(-400, 401)
    c = (k * np.pi / 4).astype(np.float32)
    near = np.concatenate([np.nextafter(c, np.float32(np.inf)), c, np.nextafter(c, -np.float32(np.inf))])
    x = np.concatenate([*parts, near]).astype(np.float32)
    return np.concatenate([x, -x, np.float32([0.0, -0.0])])


def test_sin_cos_within_one_ulp():
    rng = np.random.default_rng(11)
    x = sample_angles(rng)
    xd = x.astype(np.float64)
    for op, ref in ((SIN, np.sin), (COS, np.cos)):
        got = fmath(op, x)
        exp = ref(xd).astype(np.float32)
        d = ulp_distance(got, exp)
        # near a zero of the function the relative ulp blows up; there the
        # absolute error must stay below 2^-30 (the reduction's accuracy)
        small = np.abs(exp) < 1e-6
        assert d[~small].max() <= 1, (op, x[~small][np.argmax(d[~small])])
        assert np.abs(got[small].astype(np.float64) - ref(xd[small])).max() < 2.0 ** -30


def test_sin_cos_identities_and_special_values():
    rng = np.random.default_rng(12)
    x = rng.uniform(-20, 20, 100_000).astype(np.float32)
    s = fmath(SIN, x).astype(np.float64)
    c = fmath(COS, x).astype(np.float64)
    assert np.abs(s * s + c * c - 1).max() < 3e-7
    # odd / even symmetry holds bit for bit
    assert np.array_equal(fmath(SIN, -x), -fmath(SIN, x))
    assert np.array_equal(fmath(COS, -x), fmath(COS, x))
    assert fmath(SIN, np.float32([0.0]))[0] == 0.0 and fmath(COS, np.float32([0.0]))[0] == 1.0
    half_pi = np.float32(np.pi / 2)
    assert abs(float(fmath(SIN, np.float32([half_pi]))[0]) - 1.0) < 1e-7
    assert abs(float(fmath(COS, np.float32([np.pi]))[0]) + 1.0) < 1e-7
    bad = np.float32([np.inf, -np.inf, np.nan])
    assert np.all(np.isnan(fmath(SIN, bad))) and np.all(np.isnan(fmath(COS, bad)))


def test_acos_within_one_ulp():
    rng = np.random.default_rng(13)
    edges = np.float32([-1.0, -0.5, 0.0, 0.5, 1.0])
    near = np.concatenate([edges] + [np.nextafter(edges, np.float32(s * 2)) for s in (1, -1)])
    # every float in [1 - 2^-12, 1] and its negative (where acos is steepest)
    top = np.arange(np.float32(1 - 2.0 ** -12).view(np.int32), np.float32(1.0).view(np.int32) + 1,
                    dtype=np.int32).view(np.float32)
    x = np.concatenate([rng.uniform(-1, 1, 3_000_000).astype(np.float32), near, top, -top])
    x = x[np.abs(x) <= 1]
    got = fmath(ACOS, x)
    exp = np.arccos(x.astype(np.float64)).astype(np.float32)
    d = ulp_distance(got, exp)
    assert d.max() <= 1, x[np.argmax(d)]
    assert fmath(ACOS, np.float32([1.0]))[0] == 0.0
    assert fmath(ACOS, np.float32([-1.0]))[0] == np.float32(np.pi)
    assert fmath(ACOS, np.float32([0.0]))[0] == np.float32(np.pi / 2)
    out = np.float32([1.0000001, -1.0000001, 2.0, np.inf, np.nan])
    assert np.all(np.isnan(fmath(ACOS, out)))


def test_pow5():
    rng = np.random.default_rng(14)
    x = np.concatenate([rng.uniform(0, 1, 1_000_000), rng.uniform(-3, 3, 100_000)]).astype(np.float32)
    got = fmath(POW5, x)
    exp = (x.astype(np.float64) ** 5).astype(np.float32)
    nz = np.abs(exp) > 1e-30
    assert ulp_distance(got[nz], exp[nz]).max() <= 3
    assert fmath(POW5, np.float32([0.0, 1.0, 2.0, -1.0])).tolist() == [0.0, 1.0, 32.0, -1.0]
