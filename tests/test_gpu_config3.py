"""Reduced-precision storage at the 192 x 384 x 192 shape (14.2M nodes, the reduced-precision
configuration): the tiled kernel with BF16 and FP16 population storage, and FP32, against the
FP64 per-node kernel on the same seeded field.

The physics is the bounded point of the production-shape tests (tau 1.0, Re 50, Smagorinsky
Cs 0.1), from u = (0, U_ref, 0) + a seeded 1e-3 perturbation: the error is the storage
precision's, the work per step the configuration's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C3 = dict(nx=192, ny=384, nz=192, tau=1.0, re=50.0, u_ref=0.0, cs_smag=0.1, mesh_mode=0)
STEPS = 100


@pytest.fixture(scope="module")
def c3_fp64(pkg):
    s = pkg.Solver(pkg.default_params(**C3, precision="fp64", kernel=pkg.KERNEL_PERNODE))
    s.init_equilibrium(u0=(0.0, s.scalars()["uref"], 0.0), amp=1e-3, seed=20261203)
    s.step(STEPS)
    m = s.macros()
    s.close()
    assert all(np.isfinite(x).all() for x in m)
    return m


# bounds as at 128 x 256 x 128 (16-bit storage: FP16 ~1e-3 relative per value, BF16 ~8e-3)
@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("fp16", 6e-3), ("bf16", 2e-2)])
def test_reduced_precision_at_192x384x192(pkg, c3_fp64, prec, tol):
    s = pkg.Solver(pkg.default_params(**C3, precision=prec))
    assert s.kernel_in_use() == pkg.KERNEL_TILED
    s.init_equilibrium(u0=(0.0, s.scalars()["uref"], 0.0), amp=1e-3, seed=20261203)
    s.step(STEPS)
    r, u, v, w = s.macros()
    s.close()
    ro, uo, vo, wo = c3_fp64
    e_u = float(np.sqrt(np.sum((u - uo) ** 2 + (v - vo) ** 2 + (w - wo) ** 2) / np.sum(uo ** 2 + vo ** 2 + wo ** 2)))
    e_rho = float(np.linalg.norm(r - ro) / np.linalg.norm(ro))
    print(f"{prec} 192x384x192 {STEPS} steps: rel L2 u {e_u:.3e} rho {e_rho:.3e}", flush=True)
    assert e_u <= tol and e_rho <= tol, (e_u, e_rho)
