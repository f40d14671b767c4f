"""The pinning checks catch plausible mistakes: each case plants one mistake in a copy of
the oracle (a swapped direction, weight family, bounce-back pair, BFL operand or weight,
stencil start, Smagorinsky factor, relaxation rate, transposed inverse or force sign),
builds it, and runs the check that is meant to catch it against that build.  The check
must fail (pytest exit status 1: a failed assertion, not an error or a crash)."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle")
PIN = "tests/test_oracle_pinning.py::"

# (name, original text, planted text, check that must fail)
MUTANTS = [
    ("xi family of 15", "s->yp0, NY6, J, s->xiw[XI15], nyz, ixi);", "s->yp0, NY6, J, s->xiw[XI16], nyz, ixi);",
     "test_each_direction_streams_from_its_departure_point"),
    ("x weights of 8", "F[8] = ip_x_y_xi(fo[8], b3, nline, nface, s->xp2,", "F[8] = ip_x_y_xi(fo[8], b3, nline, nface, s->xp0,",
     "test_each_direction_streams_from_its_departure_point"),
    ("y weights of 9", "s->xp0, NX6, I, s->yp2, NY6, J, s->xiw[XI4], nyz, ixi);\n    F[10]",
     "s->xp0, NX6, I, s->yp0, NY6, J, s->xiw[XI4], nyz, ixi);\n    F[10]",
     "test_each_direction_streams_from_its_departure_point"),
    ("xi family of 12", "F[12] = ip_x_xi(fo[12], bxz, nline, s->xp2, NX6, I, s->xiw[XI5]",
     "F[12] = ip_x_xi(fo[12], bxz, nline, s->xp2, NX6, I, s->xiw[XI6]", "test_each_direction_streams_from_its_departure_point"),
    ("xi stencil start", "if (K <= 6) cz = 3;", "if (K <= 6) cz = 4;", "test_each_direction_streams_from_its_departure_point"),
    ("bounce-back pair", "F[11] = fo[14][index]; F[12] = fo[13][index];", "F[11] = fo[13][index]; F[12] = fo[14][index];",
     "test_flat_walls_bounce_back_the_opposite_population"),
    ("top bounce-back", "F[13] = fo[12][index]; F[14] = fo[11][index];", "F[13] = fo[11][index]; F[14] = fo[12][index];",
     "test_flat_walls_bounce_back_the_opposite_population"),
    ("BFL linear operands", "F[3] = (1 / (2 * q)) * fo[4][index] + ((2 * q - 1) / (2 * q)) * fo[3][index];",
     "F[3] = (1 / (2 * q)) * fo[3][index] + ((2 * q - 1) / (2 * q)) * fo[4][index];", "test_bfl_linear_branch_is_bouzidi"),
    ("BFL linear 16", "F[16] = (1 / (2 * q)) * fo[17][index] + ((2 * q - 1) / (2 * q)) * fo[16][index];",
     "F[16] = (1 / (2 * q)) * fo[18][index] + ((2 * q - 1) / (2 * q)) * fo[16][index];", "test_bfl_linear_branch_is_bouzidi"),
    ("BFL x weight of 7", "F[7] = ip_x_y_xi(fo[10], bXY, nline, nface, s->bw[BX37]",
     "F[7] = ip_x_y_xi(fo[10], bXY, nline, nface, s->bw[BX38]", "test_bfl_interpolated_branch_reaches_the_reflected_start_point"),
    ("BFL source of 15", "F[15] = ip_y_xi(fo[18], bY,", "F[15] = ip_y_xi(fo[17], bY,",
     "test_bfl_interpolated_branch_reaches_the_reflected_start_point"),
    ("BFL source of 9", "F[9] = ip_x_y_xi(fo[8], bXY,", "F[9] = ip_x_y_xi(fo[7], bXY,",
     "test_bfl_interpolated_branch_reaches_the_reflected_start_point"),
    ("LES off-diagonal factor", "+ 2.0 * (P[3] * P[3] + P[4] * P[4] + P[5] * P[5]);",
     "+ 1.0 * (P[3] * P[3] + P[4] * P[4] + P[5] * P[5]);", "test_smagorinsky_relaxation_time_satisfies_hou_relation"),
    ("LES filter width", "const double dzl = (s->z[(long)J * NZ6 + K + 1] - s->z[(long)J * NZ6 + K - 1]) / 2.0;",
     "const double dzl = (s->z[(long)J * NZ6 + K + 1] - s->z[(long)J * NZ6 + K - 1]);",
     "test_smagorinsky_relaxation_time_satisfies_hou_relation"),
    ("stress rate on m10", "S[9] = S[11] = S[13] = S[14] = S[15] = 1.0 / tau_eff;",
     "S[10] = S[11] = S[13] = S[14] = S[15] = 1.0 / tau_eff;", "test_collision_relaxes_each_moment_at_its_reference_rate"),
    ("transposed inverse", "r = r - s->Mi[q][a] * S[a]", "r = r - s->Mi[a][q] * S[a]",
     "test_collision_relaxes_each_moment_at_its_reference_rate"),
    ("force sign", "r = (E[q][1] > 0) ? r + c * fy * s->dt : r - c * fy * s->dt;",
     "r = (E[q][1] > 0) ? r - c * fy * s->dt : r + c * fy * s->dt;", "test_collision_relaxes_each_moment_at_its_reference_rate"),
]


def _build(tmp, name, old, new):
    d = os.path.join(tmp, name.replace(" ", "_"))
    os.makedirs(d, exist_ok=True)
    for f in ("islbm_oracle.c", "islbm_oracle.h"):
        shutil.copy(os.path.join(SRC, f), d)
    path = os.path.join(d, "islbm_oracle.c")
    text = open(path).read()
    assert text.count(old) == 1, f"{name}: the text to mutate is not unique in the oracle"
    open(path, "w").write(text.replace(old, new))
    so = os.path.join(d, "libmutant.so")
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-o", so, path, "-lm"],
                   check=True, capture_output=True)
    return so


def _run(so, check):
    env = dict(os.environ, ISLBM_ORACLE_LIB=so, OMP_NUM_THREADS="2")
    target = PIN + check if check else PIN.rstrip(":")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", target],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-2000:]


def test_every_planted_mistake_is_caught(tmp_path):
    def one(m):
        name, old, new, check = m
        so = _build(str(tmp_path), name, old, new)
        return name, _run(so, check)

    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(one, MUTANTS))
    missed = [(name, rc, out) for name, (rc, out) in results if rc != 1]
    assert not missed, "\n".join(f"{n}: exit {rc}\n{out}" for n, rc, out in missed)


def test_unmutated_build_passes_the_same_checks(tmp_path):
    so = _build(str(tmp_path), "none", "if (K <= 6) cz = 3;", "if (K <= 6) cz = 3;")
    rc, out = _run(so, "")
    assert rc == 0, out
