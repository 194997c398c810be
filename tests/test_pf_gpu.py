"""PhaseField stencil on the GPU vs the oracle and the reference's golden vectors.

Tolerances (from the task's correctness bar): FP64 path energy and gradient
within 1e-12 relative (gradient norm-wise), FP32 path within 1e-5 relative,
evaluated on identical coordinates (for FP32: the FP32-rounded ones).
"""
import numpy as np
import pytest

from conftest import golden
from oracle_bind import OraclePF

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5}


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make_pf(m, grid, nf, st, iw, dens="hard", res=1.0):
    pf = m.PhaseField().setNFluid(nf).setGridSize(list(grid)).setResolution(res).setSurfaceTension(st)
    if iw:
        pf.setInterfaceSize(iw if len(iw) > 1 else iw[0])
    pf.setDensityConstraint(dens)
    return pf


def check(m, grid, nf, st, iw, dens, x, prec, res=1.0):
    print("case", grid, nf, dens, prec, flush=True)
    pf = make_pf(m, grid, nf, st, iw, dens, res)
    xs = x.astype(np.float32).astype(np.float64) if prec == "fp32" else x
    s = m.State(pf, xs, prec)
    o = OraclePF(tuple(grid), nf, res, st, iw, dens)
    e, g = s.energyGradient()
    eo, go = o.energy(xs), o.gradient(xs)
    assert abs(e - eo) <= TOL[prec] * max(abs(eo), 1e-300), (e, eo)
    assert rel_err(g, go) <= TOL[prec]
    assert abs(s.energy() - eo) <= TOL[prec] * abs(eo)
    assert rel_err(s.gradient(), go) <= TOL[prec]
    return s


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", range(9))
def test_golden_vectors(gpu, case, prec):
    c = golden("pf")[case]
    x = np.array(c["x"])
    s = check(gpu, c["grid"], c["nfluid"], c["surface_tension"], c["interface_size"], c["density"], x, prec,
              c["resolution"])
    if prec == "fp64":
        assert s.energy() == pytest.approx(c["energy"], rel=1e-12)
        assert rel_err(s.gradient(), c["gradient"]) <= 1e-12


def test_known_answers(gpu):
    m = gpu
    # PhaseField_test.cpp:35-42
    s = m.State(make_pf(m, (1, 1, 1), 1, [1.0], []), [1.0])
    assert s.energy() == 0 and s.gradient()[0] == 0
    s.coords([0.1])
    assert s.energy() == pytest.approx(3.0 / 16 * 1.21 * 0.81, rel=1e-15)
    assert s.gradient()[0] == pytest.approx(-3.0 / 4 * 0.099, rel=1e-15)
    # 3-site periodic chain, hand-derived (see test_oracle)
    for grid in [(3, 1, 1), (1, 3, 1), (1, 1, 3)]:
        s = m.State(make_pf(m, grid, 1, [1.0], []), [-1, 0, 1])
        assert s.energy() == pytest.approx(3.0 / 16 + 4.5, rel=1e-15)
        np.testing.assert_allclose(s.gradient(), [-4.5, 0, 4.5], atol=1e-15)
    # PhaseField_test.cpp:182-200: fixed fluid -> zero gradient
    pf = make_pf(m, (2, 2, 1), 3, [1.0], [], "none").setFixFluid(0)
    g = m.State(pf, np.ones(12)).gradient()
    assert np.all(g[0::3] == 0)


SHAPES = [
    ((17, 13, 11), 1, [1.0], [1.0], "hard"),
    ((64, 64, 1), 1, [1.0], [1.0], "hard"),
    ((1, 1, 300), 1, [0.8], [2.0], "hard"),
    ((2, 9, 2), 1, [1.0], [1.5], "hard"),
    ((32, 32, 32), 3, [1.0, 0.8, 1.2], [2.0], "hard"),
    ((9, 300, 2), 3, [1.0, 0.8, 1.2], [2.0], "hard"),
    ((12, 10, 8), 2, [1.0], [1.0], "hard"),
    ((8, 8, 8), 4, [1.0, 0.9, 1.1, 0.8, 1.2, 0.7], [1.3], "soft"),
    ((6, 7, 5), 5, [1.0] * 5, [1.0], "none"),
    ((5, 6, 7), 3, [1.0, 0.8, 1.2], [2.0], "fixed"),
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_random_shapes(gpu, shape, prec):
    grid, nf, st, iw, dens = SHAPES[shape]
    rng = np.random.default_rng(100 + shape)
    n = int(np.prod(grid)) * nf
    x = rng.uniform(-1.1, 1.1, n) if nf == 1 else rng.uniform(-0.1, 1.1, n)
    if nf == 4:
        st = [1.0, 0.9, 1.1, 0.8]
    check(gpu, grid, nf, st, iw, dens, x, prec)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_trial_point_and_constraints(gpu, prec):
    """Energy at x + t*d (fused) equals the energy at the materialised point,
    and the projection kernel equals the oracle's specialisedConstraints."""
    import torch

    m = gpu
    grid, nf = (20, 18, 16), 3
    rng = np.random.default_rng(7)
    n = int(np.prod(grid)) * nf
    x0, d0, v0 = rng.uniform(0, 1, n), rng.normal(0, 0.1, n), rng.normal(0, 1, n)
    pf = make_pf(m, grid, nf, [1.0, 0.8, 1.2], [2.0])
    h = pf.handle(x0)
    dt = torch.float64 if prec == "fp64" else torch.float32
    x, d, v = (torch.tensor(a, dtype=dt, device="cuda") for a in (x0, d0, v0))
    t = -0.375
    e1 = m.potential_eval(h, prec, x.data_ptr(), n, d_ptr=d.data_ptr(), t=t)
    xt = x + torch.tensor(t, dtype=dt, device="cuda") * d
    torch.cuda.synchronize()
    e2 = m.potential_eval(h, prec, xt.data_ptr(), n)
    assert e1 == e2
    m.potential_constrain(h, prec, v.data_ptr(), n)
    torch.cuda.synchronize()
    o = OraclePF(grid, nf, 1.0, [1.0, 0.8, 1.2], [2.0])
    vin = v0.astype(np.float32).astype(np.float64) if prec == "fp32" else v0
    assert rel_err(v.double().cpu().numpy(), o.constrain(vin)) <= TOL[prec]
    # deterministic: repeated evaluation is bitwise identical
    e3 = m.potential_eval(h, prec, x.data_ptr(), n, d_ptr=d.data_ptr(), t=t)
    assert e3 == e1


MODE_SHAPES = [
    ((1, 70, 130), 1, [1.0], [1.5], "none"),
    ((3, 17, 65), 2, [1.0], [2.0], "hard"),
    ((5, 33, 64), 3, [1.0, 0.8, 1.2], [2.0], "hard"),
    ((2, 9, 200), 3, [1.0, 0.8, 1.2], [1.5], "soft"),
    ((4, 20, 40), 4, [1.0, 0.9, 1.1, 0.8], [1.3], "fixed"),
    ((3, 5, 7), 6, [1.0] * 6, [1.0], "hard"),
]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", range(len(MODE_SHAPES)))
def test_evaluation_modes_agree(gpu, shape, prec):
    """Every launch mode of the stencil on shapes whose tiles overhang the grid:
    gradient-only equals the energy+gradient launch's gradient and energy-only
    its energy, bit for bit; the energy at a trial point x + t*d equals the
    energy at the materialised point, bit for bit."""
    import torch

    m = gpu
    grid, nf, st, iw, dens = MODE_SHAPES[shape]
    print("case modes", grid, nf, dens, prec, flush=True)
    rng = np.random.default_rng(500 + shape)
    n = int(np.prod(grid)) * nf
    x0 = rng.uniform(-1.1, 1.1, n) if nf == 1 else rng.uniform(-0.1, 1.1, n)
    h = make_pf(m, grid, nf, st, iw, dens).handle(x0)
    dt = torch.float64 if prec == "fp64" else torch.float32
    x = torch.tensor(x0, dtype=dt, device="cuda")
    d = torch.tensor(rng.normal(0, 0.2, n), dtype=dt, device="cuda")
    g1 = torch.empty_like(x)
    g2 = torch.empty_like(x)
    e1 = m.potential_eval(h, prec, x.data_ptr(), n, g_ptr=g1.data_ptr())
    m.potential_eval(h, prec, x.data_ptr(), n, g_ptr=g2.data_ptr(), want_energy=False)
    e3 = m.potential_eval(h, prec, x.data_ptr(), n)
    torch.cuda.synchronize()
    assert torch.equal(g1, g2)
    assert e1 == e3
    t = 0.8125
    et = m.potential_eval(h, prec, x.data_ptr(), n, d_ptr=d.data_ptr(), t=t)
    xt = x + torch.tensor(t, dtype=dt, device="cuda") * d
    torch.cuda.synchronize()
    assert et == m.potential_eval(h, prec, xt.data_ptr(), n)


MIXED_SHAPES = [
    ((12, 10, 9), 1, [1.0], [1.5], "none"),
    ((11, 1, 13), 2, [1.0], [2.0], "hard"),
    ((20, 18, 16), 3, [1.0, 0.8, 1.2], [2.0], "hard"),
    ((8, 8, 8), 4, [1.0, 0.9, 1.1, 0.8], [1.3], "soft"),
    ((5, 6, 7), 3, [1.0, 0.8, 1.2], [2.0], "fixed"),
]


@pytest.mark.parametrize("shape", range(len(MIXED_SHAPES)))
def test_mixed_storage_matches_fp64(gpu, shape):
    """FP64 coordinates with FP32 direction/gradient storage: the arithmetic is
    the FP64 kernel's, so the gradient is exactly the FP64 gradient rounded to
    FP32, and energies (at x and at x + t*d) equal the FP64 evaluation with the
    FP32 direction widened."""
    import torch

    m = gpu
    grid, nf, st, iw, dens = MIXED_SHAPES[shape]
    print("case mixed", grid, nf, dens, flush=True)
    rng = np.random.default_rng(300 + shape)
    n = int(np.prod(grid)) * nf
    x0 = rng.uniform(-1.1, 1.1, n) if nf == 1 else rng.uniform(-0.1, 1.1, n)
    pf = make_pf(m, grid, nf, st, iw, dens)
    h = pf.handle(x0)
    x = torch.tensor(x0, dtype=torch.float64, device="cuda")
    d32 = torch.tensor(rng.normal(0, 0.1, n), dtype=torch.float32, device="cuda")
    d64 = d32.double()
    g64 = torch.empty(n, dtype=torch.float64, device="cuda")
    g32 = torch.empty(n, dtype=torch.float32, device="cuda")
    e64 = m.potential_eval(h, "fp64", x.data_ptr(), n, g_ptr=g64.data_ptr())
    emx = m.potential_eval_mixed(h, x.data_ptr(), n, g_ptr=g32.data_ptr())
    torch.cuda.synchronize()
    assert emx == e64
    assert torch.equal(g32, g64.float())
    t = -0.625
    assert m.potential_eval_mixed(h, x.data_ptr(), n, d_ptr=d32.data_ptr(), t=t) == m.potential_eval(
        h, "fp64", x.data_ptr(), n, d_ptr=d64.data_ptr(), t=t)
    # against the oracle too (FP64 bar)
    o = OraclePF(grid, nf, 1.0, st, iw, dens)
    assert abs(emx - o.energy(x0)) <= TOL["fp64"] * abs(o.energy(x0))
    assert rel_err(g32.double().cpu().numpy(), o.gradient(x0)) <= 1e-7


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
@pytest.mark.parametrize("shape", range(len(MIXED_SHAPES)))
def test_fused_step_equals_separate_passes(gpu, shape, prec):
    """The L-BFGS line search's fused accepted step (one pass: energy and
    gradient at x + t*d, and x + t*d written out) equals, bit for bit, the
    separate trial energy, the axpy and the gradient pass at the materialised
    point it replaces."""
    import torch

    m = gpu
    grid, nf, st, iw, dens = MIXED_SHAPES[shape]
    print("case fused", grid, nf, dens, prec, flush=True)
    rng = np.random.default_rng(700 + shape)
    n = int(np.prod(grid)) * nf
    x0 = rng.uniform(-1.1, 1.1, n) if nf == 1 else rng.uniform(-0.1, 1.1, n)
    h = make_pf(m, grid, nf, st, iw, dens).handle(x0)
    px = "fp32" if prec == "fp32" else "fp64"
    pv = "fp64" if prec == "fp64" else "fp32"
    tx = torch.float32 if px == "fp32" else torch.float64
    tv = torch.float32 if pv == "fp32" else torch.float64
    x = torch.tensor(x0, dtype=tx, device="cuda")
    d = torch.tensor(rng.normal(0, 0.1, n), dtype=tv, device="cuda")
    t = 0.75
    xo = torch.empty_like(x)
    g = torch.empty(n, dtype=tv, device="cuda")
    e = m.potential_eval_step(h, px, pv, x.data_ptr(), n, d.data_ptr(), t, xo.data_ptr(), g.data_ptr())
    xt = x + torch.tensor(t, dtype=tx, device="cuda") * d.to(tx)
    gref = torch.empty(n, dtype=tv, device="cuda")
    if prec == "mixed":
        eref = m.potential_eval_mixed(h, x.data_ptr(), n, d_ptr=d.data_ptr(), t=t)
        m.potential_eval_mixed(h, xt.data_ptr(), n, g_ptr=gref.data_ptr(), want_energy=False)
    else:
        eref = m.potential_eval(h, px, x.data_ptr(), n, d_ptr=d.data_ptr(), t=t)
        m.potential_eval(h, px, xt.data_ptr(), n, g_ptr=gref.data_ptr(), want_energy=False)
    torch.cuda.synchronize()
    assert torch.equal(xo, xt)
    assert torch.equal(g, gref)
    assert e == eref


def test_fused_step_rejects_general_kernel(gpu):
    import torch

    m = gpu
    grid, nf = (6, 5, 4), 2
    n = int(np.prod(grid)) * nf
    pf = make_pf(m, grid, nf, [1.0], [1.0], "hard")
    pf.stencilKernel = 1
    h = pf.handle(np.full(n, 0.5))
    x = torch.full((n,), 0.5, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        m.potential_eval_step(h, "fp64", "fp64", x.data_ptr(), n, x.data_ptr(), 1.0, x.data_ptr(), x.data_ptr())


def test_mixed_storage_rejects_five_fluids(gpu):
    import torch

    m = gpu
    grid, nf = (4, 5, 6), 5
    n = int(np.prod(grid)) * nf
    pf = make_pf(m, grid, nf, [1.0] * 5, [1.0], "none")
    h = pf.handle(np.full(n, 0.2))
    x = torch.full((n,), 0.2, dtype=torch.float64, device="cuda")
    g = torch.empty(n, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        m.potential_eval_mixed(h, x.data_ptr(), n, g_ptr=g.data_ptr())


def test_full_size_256cubed_ternary(gpu):
    """The bench workload (256^3 sites x 3 fluids, 50.3M DOF) at full size:
    FP64 energy and gradient vs the oracle, FP32 vs FP64, and the hard
    density invariant."""
    import torch

    m = gpu
    grid, nf = (256, 256, 256), 3
    n = 256 ** 3 * 3
    rng = np.random.default_rng(2024)
    u = rng.uniform(0, 1, (n // 3, 3))
    x = (u / u.sum(axis=1, keepdims=True)).reshape(-1)
    o = OraclePF(grid, nf, 1.0, [1.0, 0.8, 1.2], [2.0])
    print("oracle 256^3 ...", flush=True)
    eo = o.energy(x)
    go = o.gradient(x)
    pf = make_pf(m, grid, nf, [1.0, 0.8, 1.2], [2.0])
    h = pf.handle(x)
    xd = torch.from_numpy(x).cuda()
    g = torch.empty_like(xd)
    e = m.potential_eval(h, "fp64", xd.data_ptr(), n, g_ptr=g.data_ptr())
    torch.cuda.synchronize()
    assert abs(e - eo) <= 1e-12 * abs(eo)
    gd = g.cpu().numpy()
    assert rel_err(gd, go) <= 1e-12
    assert np.abs(gd.reshape(-1, 3).sum(axis=1)).max() < 1e-12
    x32 = xd.float()
    g32 = torch.empty_like(x32)
    e32 = m.potential_eval(h, "fp32", x32.data_ptr(), n, g_ptr=g32.data_ptr())
    torch.cuda.synchronize()
    assert abs(e32 - eo) <= 1e-5 * abs(eo)
    assert rel_err(g32.double().cpu().numpy(), go) <= 1e-5


def test_autotune_keeps_the_fastest_shape_and_the_same_gradient(gpu):
    """The library's launch-shape autotuner (minim_pf_autotune): it keeps the
    fastest shape per precision, a new potential of the same shape picks it up
    from the process-wide cache at init, and every shape gives the same
    gradient bit for bit and the same energy to the order of the block sums."""
    m = gpu
    m.autotune_clear()
    rng = np.random.default_rng(21)
    grid = (40, 36, 44)
    u = rng.uniform(0, 1, (int(np.prod(grid)), 3))
    x = (u / u.sum(axis=1, keepdims=True)).reshape(-1)
    pf = m.PhaseField().setNFluid(3).setGridSize(list(grid)).setSurfaceTension([1.0, 0.8, 1.2]).setInterfaceSize(2.0)
    assert pf.tunedShape("fp64", x) == (0, 0)
    times = pf.autotune(x, "fp64", blocks=(64, 256, 1024), marches=(0, 4, 8))
    assert len(times) == 9 and all(t > 0 for t in times.values())
    best = min(times, key=times.get)
    assert pf.tunedShape("fp64", x) == best  # a fresh handle: from the cache
    assert pf.tunedShape("fp32", x) == (0, 0)  # per precision
    ref = None
    for shape in [(64, 4), (1024, 0), None]:
        if shape is None:  # the tuned shape through the cache
            pf.stencilTargetBlocks, pf.stencilXChunk = 0, 0
        else:
            pf.stencilTargetBlocks, pf.stencilXChunk = shape
        e, g = m.State(pf, x).energyGradient()
        if ref is None:
            ref = (e, g)
        else:
            assert np.array_equal(g, ref[1]), shape
            assert abs(e - ref[0]) <= 1e-12 * abs(ref[0])
    m.autotune_clear()


def test_autotune_general_kernel_and_vector_passes(gpu):
    """The autotuner on the general (solid-aware) kernel, whose gradient is the
    same for every shape, and on the L-BFGS direction and Gram passes, after
    which a minimisation converges to the same energy."""
    m = gpu
    m.autotune_clear()
    rng = np.random.default_rng(22)
    grid = (24, 20, 28)
    ng = int(np.prod(grid))
    u = rng.uniform(0, 1, (ng, 3))
    x = (u / u.sum(axis=1, keepdims=True)).reshape(-1)
    solid = np.zeros(grid, np.uint8)
    solid[:, :, 0] = 1
    pf = (m.PhaseField().setNFluid(3).setGridSize(list(grid)).setSurfaceTension([1.0, 0.8, 1.2])
          .setInterfaceSize(2.0).setSolid(solid.reshape(-1).tolist()).setPressure([0.0, 0.1, 0.0]))
    g_before = m.State(pf, x, "fp32").energyGradient()[1]
    times = pf.autotune(x, "fp32", blocks=(128, 512), marches=(0, 8))
    assert len(times) == 4
    assert pf.tunedShape("fp32", x) == min(times, key=times.get)
    g_after = m.State(pf, x, "fp32").energyGradient()[1]
    assert np.array_equal(g_before, g_after)
    # the vector passes of L-BFGS
    plain = m.PhaseField().setNFluid(3).setGridSize(list(grid)).setSurfaceTension([1.0, 0.8, 1.2]).setInterfaceSize(2.0)
    s = m.State(plain, x)
    m.Lbfgs().setM(5).setMaxIter(3000).minimise(s)
    e_ref = s.energy()
    lb = m.Lbfgs().setM(5).setMaxIter(3000)
    s2 = m.State(plain, x)
    r = lb.autotune(s2, blocks=(128, 1024), reps=3)
    assert set(r["direction"]) == {128, 1024} and set(r["gram"]) == {128, 1024}
    assert r["best"][0] == min(r["direction"], key=r["direction"].get)
    assert r["best"][1] == min(r["gram"], key=r["gram"].get)
    lb.minimise(s2)
    assert lb.stats["converged"]
    e0 = m.State(plain, x).energy()
    assert abs(s2.energy() - e_ref) <= 1e-8 * max(abs(e_ref), abs(e0))
    m.autotune_clear()
