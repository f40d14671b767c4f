"""The C time-loop driver (lib/islbm_run, built from csrc/islbm_run.c) on the GPU.

It runs the reference's main-program loop in C over the C ABI, either on the library's own
solver (native) or through the reference-named calls on caller-owned padded arrays
(drop-in), and writes the reference's result files (fileIO.h:52-67, 262-292) and
Monitor.dat (monitor.h:4-26).  The two modes agree with each other, with the FP64 oracle,
and a run restarted from its own result files continues as if it had not stopped.
"""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRID = (16, 32, 32)
STEPS = 20
U0 = (0.0, 0.05, 0.0)
NAMES = ["u", "v", "w", "rho"] + [f"f{q}" for q in range(19)]


def driver(pkg):
    exe = os.path.join(os.path.dirname(pkg.islbm.LIB_PATH), "islbm_run")
    if not os.path.exists(exe):
        raise RuntimeError(f"driver missing: {exe} (run build.py)")
    return exe


def run(pkg, out, *args):
    cmd = [driver(pkg), "--grid", ",".join(map(str, GRID)), "--u0", ",".join(map(str, U0)), "--dir", str(out), *args]
    print("islbm_run", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def result(out, name):
    nx, ny, nz = GRID
    a = np.fromfile(os.path.join(out, "result", f"{name}_0.bin"), dtype=np.float64)
    assert a.size == (nx + 7) * (ny + 7) * (nz + 6)
    return a.reshape(ny + 7, nz + 6, nx + 7)


def unique(a):
    """padded [J][K][I] -> unique [k][j][i]"""
    nx, ny, nz = GRID
    return np.ascontiguousarray(a[3:3 + ny, 3:3 + nz, 3:3 + nx].transpose(1, 0, 2))


def monitor(out):
    return np.loadtxt(os.path.join(out, "Monitor.dat"), ndmin=2)


def force(out):
    with open(os.path.join(out, "result", "0_force.dat")) as fp:
        return float(fp.read())


def oracle(po, steps, ndtfrc):
    nx, ny, nz = GRID
    o = po.Oracle(po.default_params(nx=nx, ny=ny, nz=nz, tau=0.6833, re=50.0, mesh_mode=0, u_ref=0.0, ndtfrc=ndtfrc))
    o.init_equilibrium(u0=U0, amp=0.0, seed=1)
    o.step(steps)
    return o


def test_native_and_dropin_write_the_same_files(pkg, tmp_path):
    """FP64 native stepping and the reference's calls on padded arrays: every result file
    (populations, fields, ghost layers included), the body force after two controller
    updates and every Monitor.dat line agree."""
    a = run(pkg, tmp_path / "native", "--steps", str(STEPS), "--ndtfrc", "10", "--ndtmit", "4")
    b = run(pkg, tmp_path / "dropin", "--steps", str(STEPS), "--ndtfrc", "10", "--ndtmit", "4", "--dropin")
    assert a["mode"] == "native" and b["mode"] == "dropin" and a["steps"] == b["steps"] == STEPS
    for name in NAMES:
        x, y = result(tmp_path / "native", name), result(tmp_path / "dropin", name)
        assert np.max(np.abs(x - y)) < 1e-13, name
    fa, fb = force(tmp_path / "native"), force(tmp_path / "dropin")
    assert abs(fa - fb) <= 1e-12 * abs(fa), (fa, fb)
    ma, mb = monitor(tmp_path / "native"), monitor(tmp_path / "dropin")
    assert ma.shape == mb.shape == (STEPS // 4, 5)
    assert np.array_equal(ma[:, 0], np.arange(4, STEPS + 1, 4))
    assert np.max(np.abs(ma - mb)) < 1e-13
    assert abs(a["force"] - fa) < 1e-14 and abs(b["force"] - fb) < 1e-14


def test_native_fp64_matches_the_oracle(pkg, po, tmp_path):
    """The driver's FP64 result files hold the oracle's populations and fields after 20 steps
    with two forcing updates (relative L2 <= 1e-10), and its force is the oracle's."""
    run(pkg, tmp_path, "--steps", str(STEPS), "--ndtfrc", "10")
    o = oracle(po, STEPS, 10)
    want = o.get_f()
    got = np.stack([unique(result(tmp_path, f"f{q}")) for q in range(19)])
    assert pkg.rel_l2(got, want) <= 1e-10
    r, u, v, w = o.macros()
    assert pkg.rel_l2_vec([unique(result(tmp_path, n)) for n in ("u", "v", "w")], (u, v, w)) <= 1e-10
    assert pkg.rel_l2(unique(result(tmp_path, "rho")), r) <= 1e-10
    assert abs(force(tmp_path) - o.force) <= 1e-12 * abs(o.force)
    # ghost layers hold the periodic images; z ghost layers the rest state
    f3 = result(tmp_path, "f3")
    nx, ny, nz = GRID
    assert np.array_equal(f3[0:3, 3:3 + nz], f3[ny:ny + 3, 3:3 + nz])
    assert np.array_equal(f3[:, 3:3 + nz, 0:3], f3[:, 3:3 + nz, nx:nx + 3])
    assert np.all(f3[:, :3] == 1.0 / 18.0)


def test_native_fp32_matches_the_oracle(pkg, po, tmp_path):
    """The FP32 tiled kernel through the driver: velocity within 2e-5 of the oracle."""
    s = run(pkg, tmp_path, "--steps", str(STEPS), "--ndtfrc", "10", "--precision", "fp32")
    assert s["precision"] == pkg.FP32
    o = oracle(po, STEPS, 10)
    r, u, v, w = o.macros()
    assert pkg.rel_l2_vec([unique(result(tmp_path, n)) for n in ("u", "v", "w")], (u, v, w)) <= 2e-5


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
def test_restart_continues_the_run(pkg, tmp_path, prec):
    """INIT = 1 (InitialUsingBkpData) from the driver's own result files after 10 steps, then
    10 more steps from step count 10, equals 20 uninterrupted steps (BF16 included: its
    stores are dithered by the step count); statistics from step 10 on count 10 samples."""
    full, part = tmp_path / "full", tmp_path / "part"
    run(pkg, full, "--steps", "20", "--ndtfrc", "10", "--tbstart", "10", "--precision", prec)
    run(pkg, part, "--steps", "10", "--ndtfrc", "10", "--precision", prec)
    run(pkg, part, "--steps", "10", "--ndtfrc", "10", "--init", "1", "--step0", "10", "--tbstart", "0", "--precision", prec)
    for name in NAMES:
        assert np.max(np.abs(result(full, name) - result(part, name))) < 1e-13, name
    # 0_force.dat holds 15 decimals ("%.15lf"): the restart starts from the force rounded there
    assert abs(force(full) - force(part)) <= 2e-15
    for d in (full, part):
        with open(os.path.join(d, "statistics", "accu.dat")) as fp:
            assert int(fp.read()) == 10
    u_full = np.fromfile(os.path.join(full, "statistics", "U", "U_0.bin"))
    u_part = np.fromfile(os.path.join(part, "statistics", "U", "U_0.bin"))
    assert u_full.size == u_part.size > 0
    assert np.max(np.abs(u_full - u_part)) < 1e-12
