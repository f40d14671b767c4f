"""The oracle's step pinned by mathematics, term by term.

Each check derives its expected value independently of the oracle's tables: departure
points from the lattice velocities and the mesh map, wall distances from a root finder on
the hill profile (itself pinned bit for bit against the reference's model.h), the moment
basis and relaxation rates from the reference's own MRT_Matrix.h (tests/golden), and the
Smagorinsky relaxation time from Hou et al.'s implicit relation.  A swapped direction,
weight family, bounce-back pair, BFL operand or relaxation rate in the oracle fails one
of them (tests/test_oracle_mutations.py plants such mistakes and checks that).

Reference axes: x spanwise, y streamwise, z wall-normal.  Direction q moves by
e_q * minSize per step (evolution.h:451-475 with the weights of initialization.h:185-258).
"""
import json
import os

import numpy as np
import pytest
from scipy.optimize import brentq

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ref_geometry.json")))

# D3Q19 velocities in the reference's order (initialization.h:7-12)
E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
              [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = np.array([int(np.where((E == -E[q]).all(1))[0][0]) for q in range(19)])

ALPHA = 0.8
LY = 5.142 + 3.858 * ALPHA
# every BFL family has flagged nodes of both branches (q > 1/2 and q < 1/2) on this grid
HILL = dict(nx=16, ny=48, nz=32, hill_alpha=ALPHA, ly=LY, mesh_mode=0, tau=0.8, re=50.0, u_ref=0.0)


def make(po, **kw):
    p = dict(HILL)
    p.update(kw)
    return po.Oracle(po.default_params(**p))


class Geo:
    """Coordinates of the unique nodes and the computational map, from first principles."""

    def __init__(self, po, o):
        p = o.p
        self.po, self.p = po, p
        sc = o.scalars()
        self.ms = sc["dt"]
        self.x = sc["dx"] * np.arange(p.nx)
        self.y = sc["dy"] * np.arange(p.ny)
        self.xi = o.table("xi")                      # pinned by test_oracle_golden / the mesh map
        self.z = o.table("z").reshape(p.ny, p.nz)    # z[j][k]

    def H(self, y):
        return self.po.hill(float(y), self.p.ly, self.p.hill_alpha, self.p.flat)

    def xi_of(self, z, y):
        """Computational coordinate of the physical point (y, z): each mesh column is the
        affine image of the xi grid between H(y) + minSize/2 and LZ - minSize/2."""
        H = self.H(y)
        return self.p.lxi * (z - (H + self.ms / 2)) / (self.p.lz - H - self.ms)


def poly_fields(g, seed):
    """f_q = P_q(x, y, xi): a different polynomial of degree <= 3 in each computational
    coordinate per direction (exactly reproduced by the 7-point Lagrange weights)."""
    rng = np.random.default_rng(seed)
    C = rng.uniform(-1, 1, size=(19, 4, 4, 4))
    C[:, 0, 0, 0] += 3.0
    sx, sy, sxi = g.p.lx, g.p.ly, g.p.lxi

    def P(q, x, y, xi):
        X = np.asarray(x) / sx
        Y = np.asarray(y) / sy
        Z = np.asarray(xi) / sxi
        px = [np.ones_like(X), X, X * X, X * X * X]
        py = [np.ones_like(Y), Y, Y * Y, Y * Y * Y]
        pz = [np.ones_like(Z), Z, Z * Z, Z * Z * Z]
        out = 0.0
        for a in range(4):
            for b in range(4):
                for c in range(4):
                    out = out + C[q, a, b, c] * px[a] * py[b] * pz[c]
        return out

    nz, ny, nx = g.p.nz, g.p.ny, g.p.nx
    Z, Y, X = np.meshgrid(g.xi, g.y, g.x, indexing="ij")  # [k][j][i]
    f = np.stack([P(q, X, Y, Z) for q in range(19)])
    return f, P


def interior_nodes(g):
    """Nodes whose x and y stencils (i-3..i+3, j-3..j+3) do not cross the periodic wrap,
    so that a polynomial (non-periodic) field is seen unchanged by every tap."""
    return [(i, j) for i in range(3, g.p.nx - 3) for j in range(3, g.p.ny - 3)]


def flagged(o, fam):
    fl = o.table(f"bfl_flag{fam}").reshape(2, -1)
    return fl == 1


def wall_t(g, j, k, dy, dz):
    """Distance t in [0, minSize] (per axis) along -(0, dy, dz) from node (j, k) to the hill
    surface: H(y - dy t) = z - dz t, by a bracketing root finder."""
    y, z = g.y[j], g.z[j, k]
    f = lambda t: g.H(y - dy * t) - (z - dz * t)
    return brentq(f, 0.0, g.ms, xtol=1e-15, rtol=1e-15, maxiter=500)


# directions whose departure point (e_q) is tested against each BFL family, and the
# (dy, dz) along which that family meets the hill
BFL_FAMILIES = {"3": ((3, 7, 8), (1, 0)), "4": ((4, 9, 10), (-1, 0)), "15": ((15,), (1, 1)), "16": ((16,), (-1, 1))}


def wall_affected(o, g):
    """(q, k, j) entries that a wall rule replaces: bounce-back at the bottom and top
    layers, and BFL-flagged families on the first two layers."""
    out = set()
    nz = g.p.nz
    for q in range(19):
        if E[q][2] > 0:
            out |= {(q, 0, j) for j in range(g.p.ny)}
        if E[q][2] < 0:
            out |= {(q, nz - 1, j) for j in range(g.p.ny)}
    for fam, (qs, _) in BFL_FAMILIES.items():
        fl = flagged(o, fam)
        for layer in (0, 1):
            for j in np.where(fl[layer])[0]:
                out |= {(q, layer, int(j)) for q in qs}
    return out


# ------------------------------------------------------------------------------------- (a)
@pytest.mark.parametrize("mesh", [dict(), dict(mesh_mode=1, stretch=2.0)])
def test_each_direction_streams_from_its_departure_point(po, mesh):
    """Interpolated streaming moves every direction's polynomial field exactly to its own
    departure point (x, y, z) - e_q minSize, the xi coordinate of that point taken on the
    departure column (evolution.h:451-475, initialization.h:185-258).  Checked on the hill
    flanks, where H(y - minSize) != H(y + minSize), at every level including the clamped
    stencils of the first and last four levels."""
    o = make(po, **mesh)
    g = Geo(po, o)
    f, P = poly_fields(g, seed=11)
    o.set_f(f)
    o.stream_only()
    F = o.get_f()
    skip = wall_affected(o, g)
    ms = g.ms
    nx, ny, nz = g.p.nx, g.p.ny, g.p.nz
    ii = np.arange(3, nx - 3)
    worst = 0.0
    checked = 0
    for q in range(19):
        cx, cy, cz = E[q]
        for j in range(3, ny - 3):
            yd = g.y[j] - cy * ms
            ks = np.array([k for k in range(nz) if (q, k, j) not in skip], dtype=int)
            xis = np.array([g.xi_of(g.z[j, k] - cz * ms, yd) for k in ks])
            want = P(q, g.x[ii][None, :] - cx * ms, yd, xis[:, None])
            worst = max(worst, float(np.max(np.abs(F[q, ks][:, j, ii] - want))))
            checked += want.size
    assert checked > 19 * 300
    assert worst < 1e-11, worst


def test_departure_columns_differ_on_the_flanks(po):
    """The nodes above carry information: on the hill flanks the xi coordinate of a
    departure point depends on which column it is taken on (a swapped family moves it)."""
    o = make(po)
    g = Geo(po, o)
    diffs = [abs(g.xi_of(g.z[j, 5], g.y[j] - g.ms) - g.xi_of(g.z[j, 5], g.y[j] + g.ms)) for j in range(3, 10)]
    assert max(diffs) > 1e-2


# ------------------------------------------------------------------------------------- (b)
def test_flat_walls_bounce_back_the_opposite_population(po):
    """Half-way bounce-back (evolution.h:482-496): on the bottom layer every direction with
    e_z > 0 receives the node's stored opposite population, on the top layer every direction
    with e_z < 0; the other directions keep their interpolated values.  Every f_q differs."""
    o = make(po, flat=1)
    g = Geo(po, o)
    rng = np.random.default_rng(5)
    f = rng.uniform(0.5, 1.5, size=(19, g.p.nz, g.p.ny, g.p.nx)) * W[:, None, None, None]
    o.set_f(f)
    o.stream_only()
    F = o.get_f()
    nz = g.p.nz
    for q in range(19):
        if E[q][2] > 0:
            assert np.array_equal(F[q, 0], f[OPP[q], 0]), q
        if E[q][2] < 0:
            assert np.array_equal(F[q, nz - 1], f[OPP[q], nz - 1]), q
    # the in-plane directions of the wall layers stream normally (not bounced): with a
    # field that is constant per direction they keep their value
    o2 = make(po, flat=1)
    const = np.broadcast_to(W[:, None, None, None] * (1 + np.arange(19)[:, None, None, None] / 100),
                            (19, nz, g.p.ny, g.p.nx)).copy()
    o2.set_f(const)
    o2.stream_only()
    F2 = o2.get_f()
    for q in range(19):
        for k in (0, nz - 1):
            if (k == 0 and E[q][2] > 0) or (k == nz - 1 and E[q][2] < 0):
                continue
            assert np.max(np.abs(F2[q, k] - const[q, k])) < 1e-14, (q, k)


# ------------------------------------------------------------------------------ (c), (d)
def test_bfl_wall_distances_match_a_root_finder(po):
    """q of every flagged node is the fraction of minSize between the node and the hill
    surface along the family's direction (initialization.h:260-333), found here by brentq."""
    o = make(po)
    g = Geo(po, o)
    for fam, (_, (dy, dz)) in BFL_FAMILIES.items():
        fl = flagged(o, fam)
        q = o.table(f"bfl_q{fam}").reshape(2, -1)
        n = 0
        for layer in (0, 1):
            for j in np.where(fl[layer])[0]:
                t = wall_t(g, int(j), layer, dy, dz)
                assert q[layer, j] == pytest.approx(t / g.ms, abs=1e-10), (fam, layer, j)
                n += 1
        assert n > 0, fam
    # and every flag is right: the departure point is inside the hill exactly where flagged
    for fam, (_, (dy, dz)) in BFL_FAMILIES.items():
        fl = flagged(o, fam)
        for layer in (0, 1):
            for j in range(g.p.ny):
                inside = g.H(g.y[j] - dy * g.ms) > g.z[j, layer] - dz * g.ms
                assert bool(fl[layer, j]) == inside, (fam, layer, j)


def test_bfl_linear_branch_is_bouzidi(po):
    """q > 1/2: F_q = 1/(2q) f_qbar + (2q - 1)/(2q) f_q with the node's stored populations
    (evolution.h:516-518, 536-538, 553, 564), on values that differ per direction."""
    o = make(po)
    g = Geo(po, o)
    rng = np.random.default_rng(8)
    f = rng.uniform(0.5, 1.5, size=(19, g.p.nz, g.p.ny, g.p.nx)) * W[:, None, None, None]
    o.set_f(f)
    o.stream_only()
    F = o.get_f()
    n = 0
    for fam, (qs, (dy, dz)) in BFL_FAMILIES.items():
        fl = flagged(o, fam)
        for layer in (0, 1):
            for j in np.where(fl[layer])[0]:
                q = wall_t(g, int(j), layer, dy, dz) / g.ms
                if q <= 0.5 + 1e-9:
                    continue
                for d in qs:
                    want = f[OPP[d], layer, j] / (2 * q) + (2 * q - 1) / (2 * q) * f[d, layer, j]
                    assert np.max(np.abs(F[d, layer, j] - want)) < 1e-12, (fam, d, layer, j)
                    n += 1
    assert n > 0


def test_bfl_interpolated_branch_reaches_the_reflected_start_point(po):
    """q < 1/2: the arriving population is the opposite one interpolated at the point it
    left from, x_f + delta e_q with delta = minSize (1 - 2q), so that it travels one step
    through the wall reflection (evolution.h:520-570; weights initialization.h:286-328 read
    at layer * NY6 + j).  On polynomial fields the interpolation is exact."""
    o = make(po)
    g = Geo(po, o)
    f, P = poly_fields(g, seed=21)
    o.set_f(f)
    o.stream_only()
    F = o.get_f()
    n = 0
    worst = 0.0
    for fam, (qs, (dy, dz)) in BFL_FAMILIES.items():
        fl = flagged(o, fam)
        for layer in (0, 1):
            for j in np.where(fl[layer])[0]:
                j = int(j)
                if j < 3 or j > g.p.ny - 4:
                    continue
                q = wall_t(g, j, layer, dy, dz) / g.ms
                if q >= 0.5 - 1e-9:
                    continue
                delta = g.ms * (1 - 2 * q)
                for d in qs:
                    cx, cy, cz = E[d]
                    ys, zs = g.y[j] + cy * delta, g.z[j, layer] + cz * delta
                    for i in range(3, g.p.nx - 3):
                        want = P(OPP[d], g.x[i] + cx * delta, ys, g.xi_of(zs, ys))
                        worst = max(worst, abs(F[d, layer, j, i] - want))
                        n += 1
    assert n > 0
    assert worst < 1e-11, worst


# ------------------------------------------------------------------------------- (e), (f)
def equilibrium(rho, u):
    eu = E @ u
    return W * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * (u @ u))


def neq_from_stress(Pi):
    """Non-equilibrium populations of zero mass and momentum whose second moment is Pi
    (second-order Hermite term; D3Q19 is isotropic to 4th order)."""
    return 4.5 * W * np.array([E[q] @ Pi @ E[q] - np.trace(Pi) / 3 for q in range(19)])


def test_smagorinsky_relaxation_time_satisfies_hou_relation(po):
    """LES (LES/2.Hou et al. (1996).tex:489-518): for a prescribed Pi_neq the effective
    relaxation time tau* satisfies sqrt(Q) = (sqrt 2 / 3) rho tau* dt |S| together with
    tau* = tau0 + 3 (Cs Delta)^2 |S| / dt, Q = Pi_neq : Pi_neq, Delta = (dx dy dz_local)^(1/3),
    and the collision relaxes the traceless part of Pi_neq by exactly (1 - 1/tau*)."""
    cs = 0.17
    o = make(po, cs_smag=cs)
    o.force = 0.0
    g = Geo(po, o)
    sc = o.scalars()
    tau0, dt = sc["tau"], sc["dt"]
    rng = np.random.default_rng(3)
    for trial in range(6):
        j, k = int(rng.integers(0, g.p.ny)), int(rng.integers(1, g.p.nz - 1))
        A = rng.uniform(-1, 1, (3, 3)) * 10 ** rng.uniform(-6, -3)
        Pi = A + A.T
        rho = 1 + 0.01 * rng.uniform(-1, 1)
        u = 0.05 * rng.uniform(-1, 1, 3)
        fin = equilibrium(rho, u) + neq_from_stress(Pi)
        fout, mac = o.collide_one(fin, j, k)
        ts = mac[4]
        Q = float(np.sum(Pi * Pi))
        dz = (g.z[j, k + 1] - g.z[j, k - 1]) / 2
        C = (cs * np.cbrt(sc["dx"] * sc["dy"] * dz)) ** 2
        S = (ts - tau0) * dt / (3 * C)          # |S| from the Smagorinsky closure
        assert np.sqrt(Q) == pytest.approx(np.sqrt(2) / 3 * rho * ts * dt * S, rel=1e-10)
        assert ts > tau0
        Pi_out = np.einsum("q,qa,qb->ab", fout - equilibrium(rho, u), E, E)
        tl = lambda T: T - np.trace(T) / 3 * np.eye(3)
        assert np.max(np.abs(tl(Pi_out) - (1 - 1 / ts) * tl(Pi))) < 1e-12 * max(1.0, np.max(np.abs(Pi)) * 1e6)


@pytest.mark.parametrize("les", [0.0, 0.1])
def test_collision_relaxes_each_moment_at_its_reference_rate(po, les):
    """MRT with body force (MRT_Matrix.h:4-68, MRT_Process.h:4-179): with the reference's
    own moment matrix M and rates s (tests/golden), a collision of any state gives
    M f' = M f - S (M f - M f_eq) + M g, where g adds +-F dt/6 to 3, 4 and +-F dt/12 to the
    other y-moving directions; mass is kept and the y momentum gains F dt."""
    M = np.array(GOLD["M"]).reshape(19, 19)
    s = np.array(GOLD["relaxation"], dtype=float)
    o = make(po, tau=GOLD["tau"], cs_smag=les)
    Fb = 3.7e-4
    o.force = Fb
    sc = o.scalars()
    dt = sc["dt"]
    rng = np.random.default_rng(17)
    for trial in range(5):
        j, k = int(rng.integers(0, 48)), int(rng.integers(1, 31))
        rho = 1 + 0.02 * rng.uniform(-1, 1)
        u = 0.05 * rng.uniform(-1, 1, 3)
        fin = equilibrium(rho, u) * (1 + 0.02 * rng.uniform(-1, 1, 19))
        fout, mac = o.collide_one(fin, j, k)
        r = fin.sum()
        uu = (E.T @ fin) / r
        assert mac[0] == pytest.approx(r, rel=1e-14)
        assert np.allclose(mac[1:4], uu, atol=1e-15)
        rates = s.copy()
        rates[[9, 11, 13, 14, 15]] = 1.0 / mac[4]
        if les == 0.0:
            assert mac[4] == GOLD["tau"]
        g = np.array([(1 / 6 if abs(E[q]).sum() == 1 else 1 / 12) * E[q][1] * Fb * dt for q in range(19)])
        m, meq = M @ fin, M @ equilibrium(r, uu)
        want = m - rates * (m - meq) + M @ g
        assert np.max(np.abs(M @ fout - want)) < 1e-13
        assert fout.sum() == pytest.approx(r, abs=1e-14)
        assert (E.T @ fout)[1] == pytest.approx(r * uu[1] + Fb * dt, abs=1e-15)
