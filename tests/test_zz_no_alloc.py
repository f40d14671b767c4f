"""islbm_mass_correct_noalloc (the library's allocation-free entry point) against
islbm_mass_correct, through the C ABI: two calls in a row, the second on a smaller grid,
each giving the same mean density and the same corrected state bit for bit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = dict(tau=0.6833, re=50.0, mesh_mode=0, u_ref=0.0)


def correct(pkg, s, fn, rho0):
    m = ctypes.c_double()
    rc = getattr(pkg.lib(), fn)(s.h, ctypes.c_double(rho0), ctypes.byref(m))
    assert rc == 0, pkg.lib().islbm_last_error()
    return m.value


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_noalloc_mass_correction_equals_the_allocating_one(pkg, prec):
    for call, grid in enumerate([dict(nx=64, ny=96, nz=64), dict(nx=16, ny=32, nz=32)]):
        p = pkg.default_params(precision=prec, **BASE, **grid)
        a, b = pkg.Solver(p), pkg.Solver(p)
        for s in (a, b):
            s.init_equilibrium(u0=(0.0, 0.0583, 0.0), amp=1e-3, seed=31 + call)
            s.step(7)
        ma = correct(pkg, a, "islbm_mass_correct", 1.0)
        mb = correct(pkg, b, "islbm_mass_correct_noalloc", 1.0)
        print(f"{prec} call {call} {grid}: mean rho {ma!r} / {mb!r}", flush=True)
        assert np.isfinite(ma) and ma == mb, (call, ma, mb)
        # the same rho_modify enters the next step of both solvers
        a.step(3)
        b.step(3)
        assert np.array_equal(a.f(), b.f()), call
        assert all(np.array_equal(x, y) for x, y in zip(a.macros(), b.macros())), call
        a.close()
        b.close()
