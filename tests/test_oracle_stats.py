"""The oracle's turbulence statistics (statistics.h:122-224) against what the mathematics
fixes, on prescribed fields (no GPU):

- moments are the products of the fields;
- the x and y derivatives are the reference's 4th-order central differences: on a sine wave
  their error is the known truncation (k h)^4 / 30;
- the wall-normal derivative uses 5-point weights on each vertical mesh line, which are
  exact for quartics in z, including the one-sided windows at the walls;
- sums add up over samples.
"""
import numpy as np
import pytest

GRID = dict(nx=32, ny=48, nz=40, tau=0.6833, re=50.0, u_ref=0.0, mesh_mode=0)


@pytest.fixture()
def orc(po):
    o = po.Oracle(po.default_params(**GRID))
    yield o
    o.close()


def coords(o):
    nx, ny, nz = GRID["nx"], GRID["ny"], GRID["nz"]
    x = o.table("x")
    y = o.table("y")
    z = o.table("z").reshape(ny, nz)  # [j][k]
    X = np.broadcast_to(x[None, None, :], (nz, ny, nx))
    Y = np.broadcast_to(y[None, :, None], (nz, ny, nx))
    Z = np.broadcast_to(z.T[:, :, None], (nz, ny, nx))
    return X, Y, Z


def test_moments_and_sums(orc):
    X, Y, Z = coords(orc)
    rng = np.random.default_rng(3)
    u, v, w = (rng.uniform(-0.1, 0.1, X.shape) for _ in range(3))
    rho = 1.0 + rng.uniform(-0.01, 0.01, X.shape)
    orc.stats_enable()
    orc.set_macros(rho, u, v, w)
    orc.stats_accumulate()
    orc.stats_accumulate()
    assert orc.stats_count() == 2
    p = rho / 3.0 - 1.0 / 3.0
    want = {"U": u, "V": v, "W": w, "P": p, "UU": u * u, "UV": u * v, "UW": u * w, "VV": v * v, "VW": v * w,
            "WW": w * w, "PU": p * u, "PV": p * v, "PW": p * w, "PP": p * p, "KT": 0.5 * (u * u + v * v + w * w),
            "UUU": u ** 3, "UUV": u * u * v, "UUW": u * u * w, "VVU": v * v * u, "VVV": v ** 3, "VVW": v * v * w,
            "WWU": w * w * u, "WWV": w * w * v, "WWW": w ** 3}
    for name, ref in want.items():
        # P = rho/3 - 1/3 cancels: absolute round-off of a few 1e-17
        np.testing.assert_allclose(orc.stats(name), 2 * ref, rtol=1e-13, atol=1e-15, err_msg=name)


def test_horizontal_derivatives_are_fourth_order(orc):
    X, Y, Z = coords(orc)
    lx, ly = 4.5, 9.0
    kx, ky = 2 * np.pi / lx, 2 * np.pi * 3 / ly
    u = 0.05 * np.sin(kx * X)
    v = 0.05 * np.sin(ky * Y)
    w = np.zeros_like(u)
    orc.stats_enable()
    orc.set_macros(np.ones_like(u), u, v, w)
    orc.stats_accumulate()
    dx, dy = lx / GRID["nx"], ly / GRID["ny"]
    # 4th-order central difference of sin: k sinc-like factor (8 sin(kh) - sin(2kh)) / (6 h)
    fx = (8 * np.sin(kx * dx) - np.sin(2 * kx * dx)) / (6 * dx)
    fy = (8 * np.sin(ky * dy) - np.sin(2 * ky * dy)) / (6 * dy)
    np.testing.assert_allclose(orc.stats("DUDX2"), (0.05 * fx * np.cos(kx * X)) ** 2, rtol=1e-10, atol=1e-16)
    np.testing.assert_allclose(orc.stats("DVDY2"), (0.05 * fy * np.cos(ky * Y)) ** 2, rtol=1e-10, atol=1e-16)
    # the stencil factor differs from the exact derivative by the (k h)^4 / 30 truncation
    assert abs(fx / kx - 1) < (kx * dx) ** 4 / 30 * 1.01 and abs(fx / kx - 1) > (kx * dx) ** 4 / 30 * 0.9
    for nm in ("DUDY2", "DUDZ2", "DVDX2", "DVDZ2", "DWDX2", "DWDY2", "DWDZ2"):
        assert np.max(np.abs(orc.stats(nm))) < 1e-24, nm


def test_wall_normal_derivative_is_exact_for_quartics(orc):
    X, Y, Z = coords(orc)
    u = 0.01 * (Z ** 4 - 2 * Z ** 3 + 0.5 * Z)
    dudz = 0.01 * (4 * Z ** 3 - 6 * Z ** 2 + 0.5)
    w = 0.02 * Z ** 2
    orc.stats_enable()
    orc.set_macros(np.ones_like(u), u, np.zeros_like(u), w)
    orc.stats_accumulate()
    np.testing.assert_allclose(orc.stats("DUDZ2"), dudz ** 2, rtol=1e-9, atol=1e-18)
    np.testing.assert_allclose(orc.stats("DWDZ2"), (0.04 * Z) ** 2, rtol=1e-9, atol=1e-18)
    # fields constant along x and y: no horizontal derivative
    assert np.max(np.abs(orc.stats("DUDX2"))) < 1e-24
    # a quintic is not reproduced: the weights are exactly 5-point (degree 4)
    orc.stats_enable()
    orc.set_macros(np.ones_like(u), Z ** 5, np.zeros_like(u), np.zeros_like(u))
    orc.stats_accumulate()
    assert np.max(np.abs(orc.stats("DUDZ2") - (5 * Z ** 4) ** 2)) > 1e-8
