"""The C-ABI library loads, exports every function include/islbm.h declares, and rejects bad
input before touching a GPU.  No kernel runs here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "islbm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|double|const char\*)\s+\**\s*([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("islbm_create", "islbm_step", "Launch_CollisionStreaming", "Launch_ModifyForcingTerm"):
        assert n in names


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_default_params_are_the_reference_set(pkg):
    p = pkg.default_params()
    assert (p.nx, p.ny, p.nz) == (32, 128, 64)
    assert (p.lx, p.ly, p.lz, p.cfl, p.lxi) == (4.5, 9.0, 3.036, 0.6, 10.0)
    assert (p.tau, p.re, p.ndtfrc) == (0.6833, 50.0, 10000)
    assert p.mesh_mode == pkg.MESH_BISECT and p.precision == pkg.FP64


def test_required_bytes(pkg):
    p = pkg.default_params(nx=128, ny=256, nz=128, precision="fp32")
    n = 128 * 256 * 128
    b = pkg.required_bytes(p)
    # population arrays of 4 MiB and more sit an odd multiple of 512 KiB apart: 16 -> 16.5 MiB
    qb = 33 * 2 ** 19
    assert 2 * 19 * qb + 4 * n * 4 <= b < 2 * 19 * qb + 4 * n * 4 + 32 * 2 ** 20
    p16 = pkg.default_params(nx=128, ny=256, nz=128, precision="bf16")
    assert pkg.required_bytes(p16) < b


@pytest.mark.parametrize("grid,q32,q16", [
    ((32, 64, 64), 512 * 2 ** 10, 256 * 2 ** 10),  # under 4 MiB: no padding
    ((128, 256, 128), 33 * 2 ** 19, 17 * 2 ** 19),  # 16 -> 16.5 MiB, 8 -> 8.5 MiB
    ((128, 256, 120), 31 * 2 ** 19, 15 * 2 ** 19),  # 15 -> 15.5 MiB, 7.5 MiB stays
])
def test_population_stride(pkg, grid, q32, q16):
    # FP32 and BF16 storage differ only in the two population buffers (same FP32 macros and
    # tables): 2 x 19 arrays at the stride
    nx, ny, nz = grid
    b32 = pkg.required_bytes(pkg.default_params(nx=nx, ny=ny, nz=nz, precision="fp32"))
    b16 = pkg.required_bytes(pkg.default_params(nx=nx, ny=ny, nz=nz, precision="bf16"))
    assert b32 - b16 == 2 * 19 * (q32 - q16)


@pytest.mark.parametrize("bad", [dict(nx=2), dict(nz=4), dict(re=0.0), dict(tau=0.0, u_ref=0.0),
                                 dict(tau=0.5), dict(ndtfrc=0), dict(precision=7), dict(cs_smag=-1.0),
                                 dict(mesh_mode=5), dict(lz=0.9)])
def test_create_rejects_bad_parameters(pkg, bad):
    p = pkg.default_params(**bad)
    with pytest.raises(pkg.ISLBMError) as e:
        pkg.Solver(p)
    assert "(-1)" in str(e.value)


def test_create_refuses_over_budget(pkg):
    p = pkg.default_params(nx=128, ny=256, nz=128, precision="fp32")
    with pytest.raises(pkg.ISLBMError) as e:
        pkg.Solver(p, hbm_budget=1 << 20)
    assert "(-3)" in str(e.value) and "budget" in str(e.value)


def test_reference_entry_point_without_init_records_error(pkg):
    L = pkg.lib()
    arr = (ctypes.c_void_p * 19)()
    L.islbm_ref_shutdown()
    L.Launch_CollisionStreaming(arr, arr)
    assert b"islbm_ref_init" in L.islbm_last_error()


def test_collision_forms_agree_on_host(pkg):
    """The opposite-pair MRT used by the kernels equals the row-by-row form of
    MRT_Process.h to round-off (host-only check inside the library)."""
    import numpy as np
    out = np.zeros(4)
    assert pkg.lib().islbm_selftest_collision(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    assert out[0] < 1e-15 and out[1] < 1e-15, out
    assert out[2] < 1e-7 and out[3] < 5e-7, out


@pytest.mark.parametrize("kw", [dict(chunk=3), dict(chunk=0), dict(max_active=0)])
def test_batch_rejects_bad_arguments_before_touching_the_gpu(pkg, kw):
    args = dict(device=0, hbm_budget=1 << 30, max_active=4, chunk=10)
    args.update(kw)
    with pytest.raises(pkg.ISLBMError):
        pkg.Batch(**args)


def test_c_driver_rejects_bad_arguments_before_the_gpu(pkg):
    """The C time-loop driver is built beside the library and refuses odd step counts and
    cadences (two steps per loop iteration) and drop-in runs the reference's calls cannot
    express, with exit status 2, before it touches a GPU."""
    import subprocess

    exe = os.path.join(os.path.dirname(pkg.islbm.LIB_PATH), "islbm_run")
    assert os.access(exe, os.X_OK), exe
    for args in (["--steps", "3"], ["--ndtfrc", "7"], ["--ndtmit", "5"], ["--mass", "3"], ["--tbstart", "1"],
                 ["--dropin", "--precision", "fp32"], ["--dropin", "--init", "1"], ["--step0", "-2"], ["--dropin", "--step0", "4"], ["--dropin", "--amp", "1e-3"],
                 ["--grid", "8,8"], ["--precision", "fp8"], ["--bogus", "1"], ["--steps"]):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=30)
        assert r.returncode == 2, (args, r.returncode, r.stderr)
