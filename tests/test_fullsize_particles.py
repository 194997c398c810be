"""The particle workloads at their full benchmark sizes, through properties
that do not depend on size and through the oracle where it stays affordable.

Lennard-Jones, 262,144 particles at density 0.85 in a periodic cube with the
truncated-shifted cutoff 2.5 and a cell/Verlet list (the benchmark's
configuration): the total force vanishes (Newton's third law), energy and
gradient are invariant under a rigid shift of every particle (periodic box),
FP32 agrees with FP64 to 1e-5, and the forces on a random sample of particles
equal a brute-force minimum-image sum over all particles (numpy, FP64).

Bar-and-hinge sheet, 1024 x 1024 nodes (2.09 M triangles): energy and
gradient against the oracle, whose cost is linear in the number of elements,
with the bar and hinge lists derived independently of the library (the
Python restatement of the reference's list construction).
"""
import numpy as np
import pytest

from oracle_bind import OracleBaH

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def lj_system(n_side=64, density=0.85, seed=2025):
    a = (1.0 / density) ** (1.0 / 3.0)
    L = n_side * a
    rng = np.random.default_rng(seed)
    grid = np.stack(np.meshgrid(*[np.arange(n_side) * a] * 3, indexing="ij"), -1).reshape(-1, 3)
    x0 = (grid + rng.uniform(-0.1, 0.1, grid.shape)).reshape(-1)
    return x0, L


def brute_force(xyz, idx, L, rc=2.5):
    """Gradient on particles idx of the truncated-shifted LJ (sigma = epsilon =
    1) in a periodic cube of side L: minimum image over every other particle."""
    out = np.zeros((len(idx), 3))
    for k, i in enumerate(idx):
        d = xyz[i] - xyz
        d -= L * np.round(d / L)
        r2 = (d * d).sum(axis=1)
        r2[i] = np.inf
        m = r2 < rc * rc
        inv6 = 1.0 / r2[m] ** 3
        f = -24.0 * (2.0 * inv6 * inv6 - inv6) / r2[m]
        out[k] = (d[m] * f[:, None]).sum(axis=0)
    return out


def test_lj_262144_properties(gpu):
    m = gpu
    x0, L = lj_system()
    n = x0.size // 3
    assert n == 262144
    pot = m.Lj3d().setCutoff(2.5, [L, L, L], 0.3)
    s = m.State(pot, x0)
    print("LJ 262144: FP64 energy+gradient", flush=True)
    e, g = s.energyGradient()
    gx = g.reshape(-1, 3)
    # Newton's third law: the forces of the whole system sum to zero
    assert np.abs(gx.sum(axis=0)).max() <= 1e-10 * np.abs(gx).sum()
    # a rigid shift of every particle (wrapped into the box) changes nothing
    shift = np.array([0.37, -1.21, 5.13])
    y = ((x0.reshape(-1, 3) + shift) % L).reshape(-1)
    e2, g2 = s.energyGradient(y)
    assert abs(e2 - e) <= 1e-12 * abs(e)
    assert rel_err(g2, g) <= 1e-12
    # the forces on sampled particles against a brute-force minimum-image sum
    rng = np.random.default_rng(9)
    idx = rng.choice(n, 64, replace=False)
    bf = brute_force(x0.reshape(-1, 3), idx, L)
    assert rel_err(gx[idx], bf) <= 1e-12
    # FP32 storage and arithmetic (FP64 energy accumulation), against FP64 at
    # the same (FP32-rounded) coordinates: the repulsive wall turns the
    # rounding of the positions alone into force differences of 1e-4
    xr = x0.astype(np.float32).astype(np.float64)
    er, gr = s.energyGradient(xr)
    s32 = m.State(pot, xr, "fp32")
    e32, g32 = s32.energyGradient()
    assert abs(e32 - er) <= 1e-5 * abs(er)
    assert rel_err(g32, gr) <= 1e-5


def test_sheet_1024_against_the_oracle(gpu):
    from bench import sheet_triangles

    m = gpu
    nside = 1024
    tri = sheet_triangles(nside, nside)
    i, j = np.meshgrid(np.arange(nside), np.arange(nside), indexing="ij")
    rng = np.random.default_rng(77)
    x0 = np.stack([0.95 * (i - 0.5 * j), (np.sqrt(3) / 2) * j, rng.normal(0, 0.01, i.shape)], -1).reshape(-1)
    edge = np.where((i == 0) | (i == nside - 1), 1, 0).reshape(-1).nonzero()[0]
    fixed = np.stack([3 * edge, 3 * edge + 1, 3 * edge + 2], 1).reshape(-1)
    pot = m.BarAndHinge().setTriangulation(tri).setRigidity(1.0, 0.01).setLength0(1.0).setTheta0(np.pi)
    pot.setConstraints(fixed.tolist())
    print("sheet 1024^2: lists", flush=True)
    bonds, hinges = m.triangulation_lists(tri)
    assert len(bonds) == 3141633 and len(hinges) == 3137541
    nnode = nside * nside
    fx = np.zeros(3 * nnode, np.uint8)
    fx[fixed] = 1
    o = OracleBaH(nnode, bonds, np.ones(len(bonds)), np.ones(len(bonds)), hinges, np.full(len(hinges), 0.01),
                  np.full(len(hinges), np.pi), fixed=fx)
    x1 = x0 + rng.normal(0, 0.02, x0.size)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        print("sheet 1024^2:", prec, flush=True)
        xs = x1.astype(np.float32).astype(np.float64) if prec == "fp32" else x1
        s = m.State(pot, x0, prec)
        e, g = s.energyGradient(xs)
        eo, go = o.energy(xs), o.gradient(xs)
        assert abs(e - eo) <= tol * abs(eo)
        assert rel_err(g, go) <= tol
        assert np.all(g[fixed] == 0)
        # the element types at the State's own coordinates: bars, hinges (no
        # force), their sum
        if prec == "fp64":
            eb, eh = s.componentEnergy(0), s.componentEnergy(1)
            e_own = s.energy()
            assert abs(eb + eh - e_own) <= 1e-12 * abs(e_own)
            assert s.componentEnergy(2) == 0.0
