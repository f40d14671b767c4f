"""The growth mode of the reference scheme on fine wall-normal grids, pinned on the CPU.

With no body force the rest state f = w_q is an exact fixed point of the scheme (every
interpolation, bounce-back and BFL rule reproduces a uniform field), so a 1e-8 perturbation
of it measures the growth of the linearised step.  On the 256 (streamwise) x 128
(wall-normal) hill at the reference's tau = 0.6833 it grows by about 4% per step from the
first two node layers a third of a hill height past the crest; the flat channel and
nz = 64 decay.  Its cause is the xi interpolation next to the hill: a y-moving population
whose departure point lies between the hill surface and the first node layer (not
BFL-flagged, because the surface there slopes by less than 1/2) is extrapolated by the
one-sided 7-point Lagrange stencil, whose weights sum in magnitude to about 24.  Clamping
the departure xi to the node range (ISLBM_XI_CLAMPED_LINEAR) removes the mode.
"""
import numpy as np
import pytest

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)[:, None, None, None]
FINE = dict(nx=2, ny=256, nz=128, tau=0.6833, re=50.0, mesh_mode=0, u_ref=0.0)


def growth(po, steps=200, **kw):
    p = dict(FINE)
    p.update(kw)
    o = po.Oracle(po.default_params(**p))
    o.init_equilibrium(u0=(0, 0, 0), amp=1e-8, seed=20261203)
    o.force = 0.0
    norms = []
    for s in range(steps // 10):
        o.step(10)
        d = o.get_f() - W
        norms.append(float(np.sqrt(np.sum(d * d))))
    rate = (norms[-1] / norms[len(norms) // 2]) ** (1.0 / (10 * (len(norms) - 1 - len(norms) // 2)))
    d = np.abs(o.get_f() - W).sum(axis=(0, 3))  # [k][j]
    k, j = np.unravel_index(np.argmax(d), d.shape)
    return rate, int(k), int(j), o


def test_reference_scheme_grows_at_the_crest_shoulder(po):
    rate, k, j, o = growth(po)
    y = o.table("y")
    print(f"reference xi weights: growth {rate:.5f} per step, peak at k={k}, y={y[j]:.3f}", flush=True)
    assert rate > 1.02
    assert k <= 1
    yc = min(y[j], 9.0 - y[j])  # distance from the crest (y = 0 = LY)
    assert 0.2 < yc < 0.6


def test_flat_channel_and_coarser_wall_normal_grid_decay(po):
    assert growth(po, flat=1)[0] < 1.0
    assert growth(po, nz=64)[0] < 1.0


def test_clamped_near_wall_xi_removes_the_mode(po):
    rate, k, j, _ = growth(po, xi_near_wall=1)
    print(f"clamped-linear xi weights: growth {rate:.5f} per step", flush=True)
    assert rate < 1.0


def test_cause_is_extrapolation_below_the_first_layer(po):
    """At the mode's nodes the departure point of 3, 7, 8 (xi family 3) lies between the
    hill surface and the first node layer, the nodes are not BFL-flagged, and the
    reference weights extrapolate with sum |w| > 10; the clamped weights there are the
    first node's value (sum |w| = 1)."""
    o = po.Oracle(po.default_params(**FINE))
    o1 = po.Oracle(po.default_params(**FINE, xi_near_wall=1))
    sc = o.scalars()
    ms = sc["dt"]
    y, xi = o.table("y"), o.table("xi")
    z = o.table("z").reshape(256, 128)
    w0 = o.table("xiw3").reshape(128, 256, 7)
    w1 = o1.table("xiw3").reshape(128, 256, 7)
    flag = o.table("bfl_flag3").reshape(2, 256)
    hits = []
    for j in range(3, 20):
        H = po.hill(y[j] - ms)
        xs = 10.0 * (z[j, 0] - (H + ms / 2)) / (3.036 - H - ms)
        if xs < xi[0] and not flag[0, j]:
            hits.append((j, float(np.abs(w0[0, j]).sum())))
            assert np.array_equal(w1[0, j], np.eye(7)[0]), j
    assert hits, "no extrapolated departure point on the shoulder"
    assert max(a for _, a in hits) > 10.0, hits


def test_clamped_linear_weights_are_exact_for_linear_fields(po):
    """Near-wall levels: linear in xi between the bracketing nodes (exact for fields linear
    in xi); interior levels: the reference's 7-point weights (exact to degree 6)."""
    p = dict(nx=4, ny=48, nz=32, hill_alpha=0.8, ly=5.142 + 3.858 * 0.8, mesh_mode=1, stretch=2.0, tau=0.8,
             re=50.0, u_ref=0.0)
    o = po.Oracle(po.default_params(**p, xi_near_wall=1))
    ref = po.Oracle(po.default_params(**p))
    sc = o.scalars()
    ms = sc["dt"]
    y, xi = o.table("y"), o.table("xi")
    z = o.table("z").reshape(48, 32)
    fams = {"xiw3": (0, -1), "xiw4": (0, 1), "xiw5": (-1, 0), "xiw6": (1, 0),
            "xiw15": (-1, -1), "xiw16": (-1, 1), "xiw17": (1, -1), "xiw18": (1, 1)}
    for name, (dz, dy) in fams.items():
        w = o.table(name).reshape(32, 48, 7)
        wr = ref.table(name).reshape(32, 48, 7)
        for k in range(32):
            st = 0 if k <= 3 else (32 - 7 if k >= 32 - 4 else k - 3)
            nodes = xi[st:st + 7]
            for j in range(48):
                yd = y[j] + dy * ms
                H = po.hill(yd, p["ly"], p["hill_alpha"])
                xs = 10.0 * (z[j, k] + dz * ms - (H + ms / 2)) / (3.036 - H - ms)
                xc = min(max(xs, xi[0]), xi[-1])
                assert w[k, j].sum() == pytest.approx(1.0, abs=1e-12)
                assert np.dot(w[k, j], nodes) == pytest.approx(xc, abs=1e-11)
                if 4 <= k <= 32 - 5 and xi[0] <= xs <= xi[-1]:
                    assert np.array_equal(w[k, j], wr[k, j])
                if k <= 3 or k >= 32 - 4:
                    assert np.count_nonzero(w[k, j]) <= 2 and np.all(w[k, j] >= 0)
