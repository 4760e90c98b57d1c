"""adrt_amd.utils.coord_adrt / coord_cart_to_adrt (mirrors the reference's
tests/test_coord_adrt.py and tests/test_coord_cart_to_adrt.py): shapes and
dtypes, angle spacing and stitching, offset structure, quadrant boundaries,
periodicity, corner cases, and the identity coord_adrt -> coord_cart_to_adrt
-> the ADRT's own entries.  Host functions: these run on the CPU."""

import itertools

import numpy as np
import pytest

from adrt_amd import utils as U
from oracle import oracle as O


def test_coord_adrt_return_order_and_errors():
    ret = U.coord_adrt(8)
    offset, angle = ret
    assert offset is ret.offset and angle is ret.angle
    with pytest.raises(ValueError, match="invalid Radon domain size 1, must be at least 2"):
        U.coord_adrt(1)
    with pytest.raises(ValueError, match="must be a power of two"):
        U.coord_adrt(6)


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_coord_adrt_shapes_and_angles(n):
    ret = U.coord_adrt(n)
    assert ret.offset.shape == (4, 2 * n - 1, n) and ret.offset.dtype == np.float64
    assert ret.angle.shape == (4, 1, n) and ret.angle.dtype == np.float64
    angle = ret.angle[:, 0, :]
    assert np.all(angle >= -np.pi / 2) and np.all(angle <= np.pi / 2)
    shift = np.array([np.pi / 2, 0, 0, -np.pi / 2])[:, None]
    assert np.allclose(np.diff(np.tan(angle + shift)), (np.array([1, -1, 1, -1]) / (n - 1))[:, None])
    assert np.allclose(angle[:, [0, -1]], [[-np.pi / 2, -np.pi / 4], [0, -np.pi / 4],
                                           [0, np.pi / 4], [np.pi / 2, np.pi / 4]])


def test_coord_adrt_exact_at_n2():
    """Every entry at n = 2, where the definition fixes each value exactly:
    heights 3/4, 1/4, -1/4 (midpoints of three steps of 1/2 down from 1),
    slopes 0 and 1, so offsets h - 1/2 at slope 0 and ((h + 3/4)/2 - 1/2) *
    (cos + sin)(pi/4) at slope 1; angles 0 and atan(1) with the quadrant
    shifts (quadrant 1's first angle is -0.0)."""
    offset, angle = U.coord_adrt(2)
    c = np.cos(np.arctan(1.0)) + np.sin(np.arctan(1.0))
    q0 = np.array([[0.25, 0.25 * c], [-0.25, 0.0], [-0.75, -0.25 * c]])
    want = np.stack([q0, -q0, q0, -q0])
    assert offset.tobytes() == want.tobytes()
    t = np.arctan(1.0)
    want_angle = np.array([[-np.pi / 2, t - np.pi / 2], [-0.0, -t], [0.0, t],
                           [np.pi / 2, -(t - np.pi / 2)]])[:, None, :]
    assert angle.tobytes() == want_angle.tobytes()


@pytest.mark.parametrize("n", [2 ** k for k in range(1, 13)])
def test_coord_adrt_exact_first_and_last_angle(n):
    """At slope 0 and slope 1 every operation of the definition is exact in
    binary for a power-of-two n except the final product with cos + sin:
    offset(row k, slope 0) = 1/2 - (k + 1/2)/n exactly, and
    offset(row k, slope 1) = (1/2 - (k + 1)/(2n)) * (cos + sin)(pi/4),
    one rounding; the angles are exactly -pi/2, 0, 0, pi/2 at slope 0 and
    +-atan(1), +-(atan(1) - pi/2) at slope 1."""
    offset, angle = U.coord_adrt(n)
    k = np.arange(2 * n - 1, dtype=np.float64)
    first = 0.5 - (k + 0.5) / n
    c = np.cos(np.arctan(1.0)) + np.sin(np.arctan(1.0))
    last = (0.5 - (k + 1) / (2 * n)) * c
    for q, sgn in enumerate((1, -1, 1, -1)):
        assert np.array_equal(offset[q, :, 0], sgn * first)
        assert np.array_equal(offset[q, :, -1], sgn * last)
    t = np.arctan(1.0)
    assert np.array_equal(angle[:, 0, 0], [-np.pi / 2, 0.0, 0.0, np.pi / 2])
    assert np.array_equal(angle[:, 0, -1], [t - np.pi / 2, -t, t, -(t - np.pi / 2)])


def test_coord_adrt_angles_stitch():
    angle = U.coord_adrt(16).angle
    stitched = np.concatenate([angle[0, 0], angle[1, 0, :-1][::-1], angle[2, 0, 1:],
                               angle[3, 0, :-1][::-1]])
    assert np.allclose(np.unique(angle), stitched)
    full = np.concatenate([angle[0, 0], angle[1, 0][::-1], angle[2, 0], angle[3, 0][::-1]])
    assert np.allclose(np.sort(angle.ravel()), full)


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_coord_adrt_offsets(n):
    offset = U.coord_adrt(n).offset
    assert np.all(offset[:2] == offset[2:]) and np.all(offset[0] == -offset[1])
    d = np.diff(offset[0], axis=0)
    assert np.allclose(d, d[:1])
    assert np.isclose(d[0, 0], -1 / n)
    assert np.isclose(offset[0, n - 1, 0], (1 - n) / (2 * n))
    assert np.isclose(offset[0, n - 1, 0], -offset[0, 0, 0])
    last = np.sqrt(2) * (n - 1) / (2 * n)
    assert np.isclose(offset[0, 0, -1], last) and np.isclose(offset[0, -1, -1], -last)


def test_cart_to_adrt_return_order_and_dtypes():
    for td, sd in itertools.product([np.float32, np.float64], repeat=2):
        ret = U.coord_cart_to_adrt(np.array([-np.pi / 4, 0, np.pi / 4]).astype(td),
                                   np.array([-0.25, 0, 0.25]).astype(sd), 8)
        q, h, s, f = ret
        assert q is ret.quadrant and h is ret.height and s is ret.slope and f is ret.factor
        assert (q.dtype, h.dtype, s.dtype, f.dtype) == (np.uint8, np.int64, np.uint64, np.float64)


@pytest.mark.parametrize("theta,quadrant", [(-3 * np.pi / 8, 0), (-np.pi / 8, 1),
                                            (np.pi / 8, 2), (3 * np.pi / 8, 3)])
def test_cart_to_adrt_quadrant_midpoints(theta, quadrant):
    assert U.coord_cart_to_adrt(np.array([theta]), np.zeros(1), 8).quadrant.item() == quadrant


@pytest.mark.parametrize("period", [-2, -1, 1, 2])
def test_cart_to_adrt_periodic(period):
    theta = np.arange(-3, 4, 2) * np.pi / 8
    t = np.linspace(-1 / 16, 1 / 16, 4)
    a = U.coord_cart_to_adrt(theta, t, 8)
    b = U.coord_cart_to_adrt(theta + period * np.pi, t, 8)
    assert np.all(a.quadrant == b.quadrant) and np.all(a.height == b.height)
    assert np.all(a.slope == b.slope) and np.allclose(a.factor, b.factor)


def test_cart_to_adrt_errors():
    with pytest.raises(ValueError, match="mismatched shapes for theta and t"):
        U.coord_cart_to_adrt(np.zeros(3), np.zeros(2), 8)
    with pytest.raises(ValueError, match="invalid Radon domain size 1, must be at least 2"):
        U.coord_cart_to_adrt(np.array([0.25 * np.pi]), np.array([0.5]), 1)


def test_cart_to_adrt_on_adrt_coordinates():
    n = 8
    c = U.coord_adrt(n)
    idx = U.coord_cart_to_adrt(np.broadcast_to(c.angle, c.offset.shape), c.offset, n)
    assert np.all(idx.quadrant[:, :, 1:-1] == np.arange(4)[:, None, None])
    assert np.all(np.isin(idx.quadrant[:2, :, -1], [0, 1]))
    assert np.all(np.isin(idx.quadrant[2:, :, -1], [2, 3]))
    assert np.all(idx.quadrant[0, :, 0] == 0) and np.all(idx.quadrant[3, :, 0] == 3)
    assert np.all(idx.slope == np.arange(n)[None, None, :])
    out = O.adrt(np.arange(n * n, dtype=np.float32).reshape(n, n))
    assert np.allclose(out[idx.quadrant, idx.height, idx.slope], out)


_N = 4
_TH = np.array([-0.5, -0.25, -0.25, 0.0, 0.0, 0.25, 0.25, 0.5]) * np.pi + \
    np.array([1, -1, 1, -1, 1, -1, 1, -1]) * 1e-8
_T = [(1 - 1 / _N) * 0.5, (1 / _N - 1) * 0.5, (1 / _N - 1) / np.sqrt(2), (1 - 1 / _N) / np.sqrt(2)]


@pytest.mark.parametrize("ti,si,expected", [
    (0, 0, (0, 0, 0)), (0, 1, (0, _N - 1, 0)), (1, 2, (0, 2 * _N - 2, _N - 1)), (1, 3, (0, 0, _N - 1)),
    (2, 3, (1, 2 * _N - 2, _N - 1)), (2, 2, (1, 0, _N - 1)), (3, 1, (1, 0, 0)), (3, 0, (1, _N - 1, 0)),
    (4, 0, (2, 0, 0)), (4, 1, (2, _N - 1, 0)), (5, 2, (2, 2 * _N - 2, _N - 1)), (5, 3, (2, 0, _N - 1)),
    (6, 3, (3, 2 * _N - 2, _N - 1)), (6, 2, (3, 0, _N - 1)), (7, 1, (3, 0, 0)), (7, 0, (3, _N - 1, 0)),
])
def test_cart_to_adrt_corners(ti, si, expected):
    c = U.coord_cart_to_adrt(np.array([_TH[ti]]), np.array([_T[si]]), _N)
    assert (c.quadrant.item(), c.height.item(), c.slope.item()) == expected
