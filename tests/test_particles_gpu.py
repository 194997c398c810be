"""Lennard-Jones and bar-and-hinge potentials on the GPU vs the reference's
golden vectors, its unit-test known answers and the oracle."""
import math

import numpy as np
import pytest

from conftest import golden
from oracle_bind import OracleBaH, OracleLJ, OracleLJPBC, oracle_fire, oracle_lbfgs

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5}


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def r32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


# ---------------------------------------------------------------- Lennard-Jones
def test_lj_dimer_known_answer(gpu):
    # State_test.cpp:33-51 through every State entry point
    m = gpu
    s = m.State(m.Lj3d(), [0, 0, 0, 2, 0, 0])
    gex = [-93.0 / 512, 0, 0, 93.0 / 512, 0, 0]
    assert s.energy() == pytest.approx(-63.0 / 1024, rel=1e-15)
    np.testing.assert_allclose(s.gradient(), gex, atol=1e-15)
    e, g = s.energyGradient()
    assert e == pytest.approx(-63.0 / 1024, rel=1e-15)
    np.testing.assert_allclose(g, gex, atol=1e-15)
    assert s.energy([0, 0, 0, 2, 0, 0]) == pytest.approx(-63.0 / 1024, rel=1e-15)
    assert s.blockEnergy() == s.procEnergy() == s.allEnergy()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", range(3))
def test_lj_golden_vectors(gpu, case, prec):
    m = gpu
    c = golden("lj")[case]
    x = np.array(c["x"]) if prec == "fp64" else r32(c["x"])
    s = m.State(m.LjNd(c["ndim"]), x, prec)
    e, g = s.energyGradient()
    o = OracleLJ(c["ndim"], c["n"])
    eo, go = o.energy(x), o.gradient(x)
    assert abs(e - eo) <= TOL[prec] * abs(eo)
    assert rel_err(g, go) <= TOL[prec]
    if prec == "fp64":
        assert e == pytest.approx(c["energy"], rel=1e-12)
        assert rel_err(g, c["gradient"]) <= 1e-12


def lattice(n_side, a, jitter, rng):
    g = np.arange(n_side) * a
    pts = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    return (pts + rng.uniform(-jitter, jitter, pts.shape)).reshape(-1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("n_side", [5, 9])
def test_lj_cutoff_vs_oracle(gpu, n_side, prec):
    """Truncated-shifted LJ in a periodic box (cell list for 9^3, brute-force
    list for the small 5^3 box) vs the oracle's minimum-image all-pairs sum."""
    m = gpu
    rng = np.random.default_rng(n_side)
    a = 1.1
    L = n_side * a
    x = lattice(n_side, a, 0.15, rng)
    x = x if prec == "fp64" else r32(x)
    n = n_side ** 3
    pot = m.Lj3d().setCutoff(2.5, [L, L, L], 0.3)
    s = m.State(pot, x, prec)
    e, g = s.energyGradient()
    assert pot.listBuilds(s) == 1
    # the longest list: the most minimum-image neighbours within rc + skin
    p = x.reshape(-1, 3).astype(np.float64)
    dv = p[:, None, :] - p[None, :, :]
    dv -= L * np.rint(dv / L)
    r2 = (dv * dv).sum(-1)
    np.fill_diagonal(r2, np.inf)
    assert pot.maxNeighbours(s) == int((r2 < 2.8 ** 2).sum(1).max())
    o = OracleLJPBC(n, L, 2.5)
    assert abs(e - o.energy(x)) <= TOL[prec] * abs(o.energy(x))
    assert rel_err(g, o.gradient(x)) <= TOL[prec]
    # move particles beyond skin/2: the list must be rebuilt, results stay exact
    y = x + rng.uniform(-0.3, 0.3, x.shape)
    y = y if prec == "fp64" else r32(y)
    e2, g2 = s.energyGradient(y)
    assert pot.listBuilds(s) == 2
    assert abs(e2 - o.energy(y)) <= TOL[prec] * abs(o.energy(y))
    assert rel_err(g2, o.gradient(y)) <= TOL[prec]
    # particles far outside the box (whole periods added, different per
    # particle): the same energy and forces as their images inside it
    k = rng.integers(-4, 5, (n, 1)) * L
    z = (y.reshape(-1, 3) + k).reshape(-1)
    z = z if prec == "fp64" else r32(z)
    e3, g3 = s.energyGradient(z)
    assert abs(e3 - o.energy(z)) <= 10 * TOL[prec] * abs(o.energy(z))
    assert rel_err(g3, o.gradient(z)) <= 10 * TOL[prec]


def test_lj_fire_cluster(gpu):
    """Fire relaxes a 27-particle cluster; same minimum as the oracle's Fire."""
    m = gpu
    rng = np.random.default_rng(11)
    x0 = lattice(3, 1.12, 0.05, rng)
    s = m.State(m.Lj3d(), x0)
    f = m.Fire().setMaxIter(20000)
    f.minimise(s)
    o = OracleLJ(3, 27)
    _, so = oracle_fire(o, x0, max_iter=20000)
    assert f.stats["converged"] and so["converged"]
    assert s.energy() == pytest.approx(so["energy"], rel=1e-8)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fire_dt0_first_step(gpu, prec):
    """Fire with an initial time step: iteration 0 has v = 0, so p = 0, v = -dt0 g
    and dt halves before the step (Fire.cpp:50-67): x1 = x0 - 0.5 dt0^2 g(x0)."""
    m = gpu
    rng = np.random.default_rng(5)
    n_side, a = 5, 1.1
    L = n_side * a
    x0 = lattice(n_side, a, 0.1, rng)
    x0 = x0 if prec == "fp64" else r32(x0)
    pot = m.Lj3d().setCutoff(2.5, [L, L, L], 0.3)
    s = m.State(pot, x0, prec)
    g0 = s.gradient()
    f = m.Fire(0.02).setDt0(0.002).setMaxIter(0)
    x1 = f.minimise(s)
    np.testing.assert_allclose(x1, x0 - 0.5 * 0.002 ** 2 * g0, rtol=TOL[prec], atol=TOL[prec] * np.abs(x0).max())


def test_fire_max_force_lj_cutoff_vs_oracle(gpu):
    """The LJ benchmark's minimiser settings on a small periodic box: Fire with
    dt0 0.002, dtMax 0.02, no line search, converged when the largest
    per-particle force is below 1e-6; same minimum as the oracle's Fire on the
    minimum-image all-pairs potential."""
    m = gpu
    rng = np.random.default_rng(31)
    n_side, a = 6, (1 / 0.85) ** (1 / 3)  # density 0.85
    L = n_side * a
    x0 = lattice(n_side, a, 0.08, rng)
    n = n_side ** 3
    pot = m.Lj3d().setCutoff(2.5, [L, L, L], 0.3)
    s = m.State(pot, x0)
    s.convergence = 1e-6
    f = m.Fire(0.02).setDt0(0.002).setMaxForceConvergence(3).setMaxIter(50000)
    x = f.minimise(s)
    o = OracleLJPBC(n, L, 2.5)
    xo, so = oracle_fire(o, x0, dt_max=0.02, dt0=0.002, max_force_dim=3, max_iter=50000, linesearch="none",
                         convergence=1e-6)
    print(f.stats, so, flush=True)
    assert f.stats["converged"] and so["converged"]
    assert f.stats["grad_max"] < 1e-6
    gmax = np.sqrt((o.gradient(x).reshape(-1, 3) ** 2).sum(axis=1)).max()
    assert gmax < 2e-6
    assert s.energy() == pytest.approx(so["energy"], rel=1e-8)
    assert abs(f.stats["iterations"] - so["iterations"]) <= max(20, 0.2 * so["iterations"])


# ---------------------------------------------------------------- BarAndHinge
def bah_from_case(m, c, use_tri=False):
    p = m.BarAndHinge()
    if use_tri:
        p.setTriangulation(c["tri"]).setRigidity(c["kb"], c["kh"])
        if c["l0"] is not None:
            p.setLength0(c["l0"])
        if c["t0"] is not None:
            p.setTheta0(c["t0"])
    else:
        p.setBondList(c["bars"]).setHingeList(c["hinges"]).setRigidity(c["bar_k"], c["hinge_k"])
        p.setLength0(c["bar_l0"]).setTheta0(c["hinge_t0"])
    if c["force"] is not None:
        p.setForce(c["force"])
    if c["fixed"]:
        p.setConstraints(c["fixed"])
    return p


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", range(2))
def test_bah_golden_vectors(gpu, case, prec):
    m = gpu
    c = golden("bah")[case]
    x = np.array(c["x"]) if prec == "fp64" else r32(c["x"])
    s = m.State(bah_from_case(m, c), c["x0"], prec)
    e, g = s.energyGradient(x)
    nnode = len(x) // 3
    fixed = None
    if c["fixed"]:
        fixed = np.zeros(3 * nnode, np.uint8)
        fixed[c["fixed"]] = 1
    o = OracleBaH(nnode, c["bars"], c["bar_k"], c["bar_l0"], c["hinges"], c["hinge_k"], c["hinge_t0"],
                  force=None if c["force"] is None else np.array(c["force"]), fixed=fixed)
    eo, go = o.energy(x), o.gradient(x)
    assert abs(e - eo) <= TOL[prec] * abs(eo)
    assert rel_err(g, go) <= TOL[prec]
    if prec == "fp64":
        assert e == pytest.approx(c["energy"], rel=1e-12)
        assert rel_err(g, c["gradient"]) <= 1e-12
        if c["fixed"]:
            assert np.all(g[c["fixed"]] == 0)
        # the same sheet built from its triangulation (O(T) list construction)
        s2 = m.State(bah_from_case(m, c, use_tri=True), c["x0"], prec)
        e2, g2 = s2.energyGradient(x)
        assert e2 == pytest.approx(e, rel=1e-14)
        assert rel_err(g2, g) <= 1e-14


def test_bah_two_triangle_minima(gpu):
    # BarAndHinge_test.cpp:13-56: L-BFGS minima for five rest angles, within 1e-4
    m = gpu
    c = golden("bah")[2]
    dz = math.sqrt(2) / 2
    s2 = math.sqrt(2)
    expected = [[0, 0, 0.5 + dz, 0, -s2, 0.5 - dz, 0, s2, 0.5 - dz, 0, 0, 0.5 + dz],
                [-1, 0, 1, 0, -s2, 0, 0, s2, 0, 1, 0, 1],
                [-s2, 0, 0.5, 0, -s2, 0.5, 0, s2, 0.5, s2, 0, 0.5],
                [-1, 0, 0, 0, -s2, 1, 0, s2, 1, 1, 0, 0],
                [0, 0, 0.5 - dz, 0, -s2, 0.5 + dz, 0, s2, 0.5 + dz, 0, 0, 0.5 - dz]]
    for mn_ref, ex in zip(c["minima"], expected):
        p = m.BarAndHinge().setTriangulation(c["tri"]).setRigidity(c["kb"], c["kh"]).setTheta0(mn_ref["theta0"])
        s = m.State(p, c["x0"])
        x = m.Lbfgs().minimise(s)
        np.testing.assert_allclose(x, ex, atol=1e-4)
        np.testing.assert_allclose(x, mn_ref["x"], atol=1e-6)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_bah_sheet_vs_oracle(gpu, prec):
    """The benchmark sheet's construction at 256^2 nodes (equilateral lattice
    compressed by 0.95, clamped x edges, z noise): energy and gradient at the
    start and at a perturbed point against the oracle, for the energy pass
    that stores the hinge angles, the gradient-only pass and a trial point."""
    import torch
    from bench import sheet_triangles

    m = gpu
    nside = 256
    tri = sheet_triangles(nside, nside)
    i, j = np.meshgrid(np.arange(nside), np.arange(nside), indexing="ij")
    rng = np.random.default_rng(78)
    x0 = np.stack([0.95 * (i - 0.5 * j), (np.sqrt(3) / 2) * j, rng.normal(0, 0.01, i.shape)], -1).reshape(-1)
    edge = np.where((i == 0) | (i == nside - 1), 1, 0).reshape(-1).nonzero()[0]
    fixed = np.stack([3 * edge, 3 * edge + 1, 3 * edge + 2], 1).reshape(-1)
    pot = m.BarAndHinge().setTriangulation(tri).setRigidity(1.0, 0.01).setLength0(1.0).setTheta0(np.pi)
    pot.setConstraints(fixed.tolist())
    bonds, hinges = m.triangulation_lists(tri)
    nnode = nside * nside
    fx = np.zeros(3 * nnode, np.uint8)
    fx[fixed] = 1
    o = OracleBaH(nnode, bonds, np.ones(len(bonds)), np.ones(len(bonds)), hinges, np.full(len(hinges), 0.01),
                  np.full(len(hinges), np.pi), fixed=fx)
    x1 = x0 + rng.normal(0, 0.02, x0.size)
    xs = r32(x1) if prec == "fp32" else x1
    s = m.State(pot, x0, prec)
    e, g = s.energyGradient(xs)
    eo, go = o.energy(xs), o.gradient(xs)
    assert abs(e - eo) <= TOL[prec] * abs(eo)
    assert rel_err(g, go) <= TOL[prec]
    assert np.all(g[fixed] == 0)
    assert rel_err(s.gradient(xs), go) <= TOL[prec]
    # energy at a trial point x + t*d (materialised before the energy pass)
    h = pot.handle(x0)
    dt = torch.float64 if prec == "fp64" else torch.float32
    xd = torch.tensor(xs, dtype=dt, device="cuda")
    d = torch.tensor(rng.normal(0, 0.01, x0.size), dtype=dt, device="cuda")
    t = 0.625
    et = m.potential_eval(h, prec, xd.data_ptr(), x0.size, d_ptr=d.data_ptr(), t=t)
    xt = (xd + torch.tensor(t, dtype=dt, device="cuda") * d).double().cpu().numpy()
    assert abs(et - o.energy(xt)) <= TOL[prec] * abs(o.energy(xt))


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_bah_fused_step_matches_separate_passes(gpu, prec):
    """The L-BFGS line search's fused accepted step for the bar-and-hinge
    sheet (x + t*d written out, then energy and gradient there in one energy
    pass and the gather) gives bit for bit the trial energy of the separate
    trial pass, and the gradient pass at the point it wrote; fixed DOFs stay
    zero."""
    import torch
    from bench import sheet_triangles

    m = gpu
    nside = 40
    print("case bah fused step", nside, prec, flush=True)
    tri = sheet_triangles(nside, nside)
    i, j = np.meshgrid(np.arange(nside), np.arange(nside), indexing="ij")
    rng = np.random.default_rng(81)
    x0 = np.stack([0.95 * (i - 0.5 * j), (np.sqrt(3) / 2) * j, rng.normal(0, 0.01, i.shape)], -1).reshape(-1)
    edge = np.where((i == 0) | (i == nside - 1), 1, 0).reshape(-1).nonzero()[0]
    fixed = np.stack([3 * edge, 3 * edge + 1, 3 * edge + 2], 1).reshape(-1)
    pot = m.BarAndHinge().setTriangulation(tri).setRigidity(1.0, 0.01).setLength0(1.0).setTheta0(np.pi)
    pot.setConstraints(fixed.tolist())
    n = x0.size
    h = pot.handle(x0)
    px = "fp32" if prec == "fp32" else "fp64"
    pv = "fp64" if prec == "fp64" else "fp32"
    tx = torch.float32 if px == "fp32" else torch.float64
    tv = torch.float32 if pv == "fp32" else torch.float64
    x = torch.tensor(x0 + rng.normal(0, 0.02, n), dtype=tx, device="cuda")
    d = torch.tensor(rng.normal(0, 0.01, n), dtype=tv, device="cuda")
    t = 0.625
    xo = torch.empty_like(x)
    g = torch.empty(n, dtype=tv, device="cuda")
    e = m.potential_eval_step(h, px, pv, x.data_ptr(), n, d.data_ptr(), t, xo.data_ptr(), g.data_ptr())
    gref = torch.empty(n, dtype=tv, device="cuda")
    if prec == "mixed":
        eref = m.potential_eval_mixed(h, x.data_ptr(), n, d_ptr=d.data_ptr(), t=t)
        m.potential_eval_mixed(h, xo.data_ptr(), n, g_ptr=gref.data_ptr(), want_energy=False)
    else:
        eref = m.potential_eval(h, px, x.data_ptr(), n, d_ptr=d.data_ptr(), t=t)
        m.potential_eval(h, px, xo.data_ptr(), n, g_ptr=gref.data_ptr(), want_energy=False)
    torch.cuda.synchronize()
    assert e == eref
    assert torch.equal(g, gref)
    assert torch.all(g[torch.tensor(fixed, device="cuda")] == 0)
    xt = x.double() + t * d.double()
    ulp = 2.0 ** -22 if px == "fp32" else 2.0 ** -50
    assert torch.allclose(xo.double(), xt, rtol=ulp, atol=ulp)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_bah_lbfgs_fused_step_keeps_the_trajectory(gpu, prec):
    """L-BFGS on the bar-and-hinge sheet takes its accepted first trials
    through the fused step, which keeps the hinge store for the gradient that
    follows.  A maxStep too large to bind runs the same algorithm through the
    separate trial energy, axpy and gradient passes: after 300 iterations the
    coordinates, energy and evaluation counts agree bit for bit."""
    from bench import sheet_triangles

    m = gpu
    nside = 24
    print("case bah lbfgs fused trajectory", nside, prec, flush=True)
    tri = sheet_triangles(nside, nside)
    i, j = np.meshgrid(np.arange(nside), np.arange(nside), indexing="ij")
    rng = np.random.default_rng(82)
    x0 = np.stack([0.95 * (i - 0.5 * j), (np.sqrt(3) / 2) * j, rng.normal(0, 0.01, i.shape)], -1).reshape(-1)
    edge = np.where((i == 0) | (i == nside - 1), 1, 0).reshape(-1).nonzero()[0]
    fixed = np.stack([3 * edge, 3 * edge + 1, 3 * edge + 2], 1).reshape(-1)
    out = []
    for cap in (0.0, 1e30):
        p = m.BarAndHinge().setTriangulation(tri).setRigidity(1.0, 0.01).setLength0(1.0).setTheta0(np.pi)
        p.setConstraints(fixed.tolist())
        s = m.State(p, x0, "fp64")
        lb = m.Lbfgs().setM(10).setMaxIter(300).setPrecision(prec)
        if cap:
            lb.setMaxStep(cap)
        lb.minimise(s)
        st = lb.stats
        out.append((s.coords(), s.energy(), st["iterations"], st["n_energy"], st["n_gradient"]))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


def test_bah_mixed_evaluation_and_minimisation(gpu):
    """The mixed schedule's evaluation of the bar-and-hinge sheet: FP64
    coordinates and arithmetic with FP32 direction and gradient.  The energy
    is the FP64 evaluation's bit for bit, the gradient its FP32 rounding
    (fixed DOFs zero), and a trial point x + t d with d in FP32 gives the FP64
    energy at d widened to FP64.  L-BFGS m=10 with either schedule then
    reaches the tolerance (the oracle's gradient at the coordinates left),
    with the energy the oracle's at those coordinates (1e-12)."""
    import torch
    from bench import sheet_triangles

    m = gpu
    nside = 24
    tri = sheet_triangles(nside, nside)
    i, j = np.meshgrid(np.arange(nside), np.arange(nside), indexing="ij")
    rng = np.random.default_rng(79)
    x0 = np.stack([0.95 * (i - 0.5 * j), (np.sqrt(3) / 2) * j, rng.normal(0, 0.01, i.shape)], -1).reshape(-1)
    edge = np.where((i == 0) | (i == nside - 1), 1, 0).reshape(-1).nonzero()[0]
    fixed = np.stack([3 * edge, 3 * edge + 1, 3 * edge + 2], 1).reshape(-1)

    def make():
        p = m.BarAndHinge().setTriangulation(tri).setRigidity(1.0, 0.01).setLength0(1.0).setTheta0(np.pi)
        p.setConstraints(fixed.tolist())
        return p

    pot = make()
    n = x0.size
    h = pot.handle(x0)
    x = torch.tensor(x0 + rng.normal(0, 0.02, n), dtype=torch.float64, device="cuda")
    d32 = torch.tensor(rng.normal(0, 0.01, n), dtype=torch.float32, device="cuda")
    g64 = torch.empty(n, dtype=torch.float64, device="cuda")
    g32 = torch.empty(n, dtype=torch.float32, device="cuda")
    e64 = m.potential_eval(h, "fp64", x.data_ptr(), n, g_ptr=g64.data_ptr())
    emx = m.potential_eval_mixed(h, x.data_ptr(), n, g_ptr=g32.data_ptr())
    torch.cuda.synchronize()
    assert emx == e64
    assert torch.equal(g32, g64.float())
    assert torch.all(g32[torch.tensor(fixed, device="cuda")] == 0)
    t = 0.625
    assert m.potential_eval_mixed(h, x.data_ptr(), n, d_ptr=d32.data_ptr(), t=t) == m.potential_eval(
        h, "fp64", x.data_ptr(), n, d_ptr=d32.double().data_ptr(), t=t)

    bonds, hinges = m.triangulation_lists(tri)
    fx = np.zeros(n, np.uint8)
    fx[fixed] = 1
    o = OracleBaH(nside * nside, bonds, np.ones(len(bonds)), np.ones(len(bonds)), hinges,
                  np.full(len(hinges), 0.01), np.full(len(hinges), np.pi), fixed=fx)
    out = {}
    for prec in ("fp64", "mixed"):
        s = m.State(make(), x0, "fp64")
        lb = m.Lbfgs().setM(10).setMaxIter(200000).setPrecision(prec)
        lb.minimise(s)
        st = lb.stats
        print(prec, st["iterations"], st["converged"], st["energy"], flush=True)
        assert st["converged"], (prec, st)
        xs = s.coords()
        assert abs(s.energy() - o.energy(xs)) <= 1e-12 * abs(o.energy(x0))
        g = o.gradient(xs)
        assert np.sqrt(np.mean(g * g)) < 1.0001 * s.convergence
        out[prec] = s.energy()
    # The compressed sheet buckles into a very flat minimum: at a gradient RMS
    # of 1e-6 its energy is pinned only to about 1e-5 of the initial energy,
    # and the two schedules stop at different points of that valley.
    assert abs(out["mixed"] - out["fp64"]) <= 1e-5 * abs(o.energy(x0))
