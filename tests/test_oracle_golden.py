"""The FP64 oracle against the reference's own outputs and against closed forms.

tests/golden/ref_geometry.json holds numbers produced by the reference's host code
(variables.h, model.h, initializationTool.h, MRT_Matrix.h compiled by oracle/Makefile into
oracle/_ref/ref_geom; regenerated by tests/golden/make_ref_golden.py).  Bitwise agreement
is required where the oracle evaluates the same expression.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ref_geometry.json")))


def _dp(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_hill_matches_reference_bitwise(po):
    ys = np.array(GOLD["hill_y"])
    hs = np.array([po.hill(y) for y in ys])
    assert np.array_equal(hs, np.array(GOLD["hill_h"]))
    hb = np.array([po.hill(y) for y in GOLD["hill_break_y"]])
    assert np.array_equal(hb, np.array(GOLD["hill_break_h"]))


def test_hill_invariants(po):
    # crest height 1 at both ends, flat floor between the flanks, mirror symmetry
    assert po.hill(0.0) == 1.0 and po.hill(9.0) == 1.0
    for y in np.linspace(1.93, 9.0 - 1.93, 17):
        assert po.hill(y) == 0.0
    for y in np.linspace(0.0, 4.5, 91):
        assert po.hill(y) == pytest.approx(po.hill(9.0 - y), abs=1e-12)
    # continuity at the six ERCOFTAC breakpoints to the fit's own accuracy
    for b in (9, 14, 20, 30, 40, 54):
        y = b / 28.0
        assert abs(po.hill(y - 1e-9) - po.hill(y + 1e-9)) < 1e-5
    # periodic wrap
    assert po.hill(-0.3) == po.hill(9.0 - 0.3)
    # alpha = 2 stretches the flank: H_2(2y) == H_1(y) on the left flank
    for y in np.linspace(0.0, 1.9, 20):
        assert po.hill(2 * y, 9.0, 2.0) == pytest.approx(po.hill(y), abs=1e-14)


def test_tanh_parameter_and_mesh_match_reference(po):
    L = po.lib()
    a = L.or_bisect_a(GOLD["LZ"], GOLD["minSize"], GOLD["NZ"], 1.0)
    assert a == GOLD["a_bisect"]
    total = GOLD["LZ"] - 1.0 - GOLD["minSize"]
    col = [L.or_tanh_map(total, GOLD["minSize"], a, k, GOLD["NZ"] - 1) for k in range(GOLD["NZ"])]
    assert np.array_equal(np.array(col), np.array(GOLD["tanh_column_y0"]))
    # first spacing equals minSize; the column spans [minSize/2, L + minSize/2]
    assert abs((col[1] - col[0]) - GOLD["minSize"]) <= 1e-14
    assert col[0] == pytest.approx(GOLD["minSize"] / 2, abs=1e-15)
    assert col[-1] == pytest.approx(total + GOLD["minSize"] / 2, abs=1e-14)


def test_lagrange_weights_match_reference(po):
    L = po.lib()
    xi = np.array(GOLD["xi_nodes"])
    got = []
    for st, pos in zip(GOLD["lag_xi_start"], GOLD["lag_xi_pos"]):
        nodes = np.ascontiguousarray(xi[int(st):int(st) + 7])
        got.append([L.or_lagrange6(pos, _dp(nodes), m) for m in range(7)])
    got = np.array(got)
    assert np.array_equal(got.ravel(), np.array(GOLD["lag_xi_w"]))
    # partition of unity and exact reproduction of polynomials up to degree 6
    for (st, pos), w in zip(zip(GOLD["lag_xi_start"], GOLD["lag_xi_pos"]), got):
        nodes = xi[int(st):int(st) + 7]
        for deg in range(7):
            assert np.dot(w, nodes ** deg) == pytest.approx(pos ** deg, rel=1e-10, abs=1e-10)


@pytest.mark.parametrize("fam", ["bfl3", "bfl4", "bfl15", "bfl16"])
def test_bfl_bisections_match_reference(po, fam):
    L = po.lib()
    ms = GOLD["minSize"]
    zs, ys, ds = GOLD[fam + "_z"], GOLD[fam + "_y"], GOLD[fam + "_delta"]
    assert len(ds) > 0
    for z, y, d in zip(zs, ys, ds):
        if fam == "bfl3":
            got = L.or_delta_horizontal(z, y - ms, y, y, 9.0, 1.0, ms)
        elif fam == "bfl4":
            got = L.or_delta_horizontal(z, y + ms, y, y, 9.0, 1.0, ms)
        elif fam == "bfl15":
            got = L.or_delta_45(z, y, y, y - ms, 9.0, 1.0, ms)
        else:
            got = L.or_delta_45(z, y, y, y + ms, 9.0, 1.0, ms)
        assert got == d
    # the horizontal search lands on the hill surface: H(y_wall) = z_node
    for z, y, d in zip(GOLD["bfl3_z"], GOLD["bfl3_y"], GOLD["bfl3_delta"]):
        y_wall = y - (ms - d) / 2.0
        assert po.hill(y_wall) == pytest.approx(z, abs=1e-12)


def test_mrt_tables_match_reference(po):
    M, Mi = po.mrt_matrices()
    assert np.array_equal(M.ravel(), np.array(GOLD["M"]))
    assert np.array_equal(Mi.ravel(), np.array(GOLD["Mi"]))
    assert np.max(np.abs(M @ Mi - np.eye(19))) < 1e-14
    s = np.array(GOLD["relaxation"])
    assert s[[0, 3, 5, 7]].tolist() == [0, 0, 0, 0]
    assert s[9] == pytest.approx(1 / 0.6833)


def test_reference_parameter_set(po):
    o = po.Oracle(po.default_params(nx=32, ny=128, nz=64, tau=0.6833, re=50.0, mesh_mode=0))
    sc = o.scalars()
    assert sc["dt"] == GOLD["minSize"] == GOLD["dt"]
    assert sc["nu"] == GOLD["niu"]
    assert sc["uref"] == GOLD["Uref"]
    assert sc["a"] == GOLD["a_bisect"]


def test_live_reference_driver_when_present():
    exe = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_geom")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/ref_geom not built (reference tree absent)")
    live = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    for k in ("hill_h", "a_bisect", "M", "Mi", "bfl15_delta", "lag_xi_w"):
        assert live[k] == GOLD[k]
