"""The bulk-flux probe (AccumulateUbulk, evolution.h:826-840) and the forcing controller
(Launch_ModifyForcingTerm, evolution.h:1003-1042) on every kernel, against the closed form
on the solver's own fields: after N steps the accumulator holds the sum over steps
m = 0 .. N-1 of sum_k,i v_m[k][jp][i] dx_i dz_k, v_m being the v that step m starts from
(the tiled kernel sums it inside the step; the other kernels in separate launches)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = dict(tau=0.6833, re=50.0, mesh_mode=0, u_ref=0.0)
KINDS = [dict(nx=64, ny=32, nz=48, precision="fp32"),        # tiled, 16-byte staging
         dict(nx=16, ny=32, nz=32, precision="fp32"),        # tiled, one x tile wider than the grid
         dict(nx=48, ny=21, nz=40, precision="bf16", cs_smag=0.1),  # partial row tile
         dict(nx=18, ny=20, nz=24, precision="fp16"),        # tiled, 4-byte staging
         dict(nx=160, ny=36, nz=70, precision="fp32"),       # five 32-column tiles, several z chunks
         dict(nx=32, ny=35, nz=30, precision="fp32"),        # 32-column tiles, the probe row first in its wave
         dict(nx=16, ny=32, nz=32, precision="fp64")]        # two-pass FP64, separate probe


def probe(s, fw, jp):
    v = s.macros()[2]
    return float(np.sum(v[:, jp, :].astype(np.float64) * fw))


@pytest.mark.parametrize("kw", KINDS)
def test_flux_accumulator_and_force_update(pkg, kw):
    p = pkg.default_params(**BASE, ndtfrc=6, **kw)
    s = pkg.Solver(p)
    fw = s.table("fluxw").reshape(p.nz, p.nx)
    jp = (p.ny - 3) % p.ny
    s.init_equilibrium(u0=(0.0, 0.0583, 0.0), amp=1e-3, seed=11)
    sc = s.scalars()
    F = s.force
    for cycle in range(2):
        acc = 0.0
        for m in range(6):
            acc += probe(s, fw, jp)
            s.step(1)
            if m < 5:
                assert s.flux_sum() == pytest.approx(acc, rel=1e-12), (cycle, m)
        # the controller ran after the 6th step and reset the accumulator
        ub = acc / (p.lx * (p.lz - 1.0)) / p.ndtfrc
        beta = max(0.001, 3.0 / p.re)
        F = F + beta * (sc["uref"] - ub) * sc["uref"] / 3.036
        assert s.force == pytest.approx(F, rel=1e-12)
        assert s.flux_sum() == 0.0
    # v set from outside the step: the next step's probe reads the uploaded v
    r, u, v, w = s.macros()
    v2 = v * 1.5
    if kw["precision"] != "fp64":
        v2 = v2.astype(np.float32).astype(np.float64)  # the value the solver stores
    s.upload_macros(r, u, v2, w)
    acc = float(np.sum(v2[:, jp, :] * fw))
    s.step(1)
    assert s.flux_sum() == pytest.approx(acc, rel=1e-12)


@pytest.mark.parametrize("kw", KINDS[:3] + KINDS[5:])
def test_flux_and_force_through_graphs_equal_eager(pkg, kw):
    """Chunks replayed from graphs (and the chunks around a controller update, stepped
    eagerly) give the same accumulator and force bit for bit."""
    p = pkg.default_params(**BASE, ndtfrc=6, **kw)
    a, b = pkg.Solver(p), pkg.Solver(p)
    for s in (a, b):
        s.init_equilibrium(u0=(0.0, 0.0583, 0.0), amp=1e-3, seed=5)
    for chunk in range(8):
        a.step(2)
        b.step_graph(2)
        assert a.flux_sum() == b.flux_sum(), chunk
        assert a.force == b.force, chunk
    assert np.array_equal(a.f(), b.f())
