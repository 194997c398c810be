"""The element form of the phase field (PhaseFieldUnstructured) on the GPU
against the oracle, the reference's unit-test known answers
(test/unit/PhaseFieldUnstructured_test.cpp) and the gather-form PhaseField.

Within a chunk of elements the gradient is accumulated with LDS atomics, so
its summation order (and, in the last bits, its value) varies from launch to
launch.  Tolerances: FP64 within
1e-12 relative (gradients norm-wise; energies relative to the sum of the
magnitudes of the energy terms), FP32 within 1e-5, on identical (for FP32:
FP32-rounded) coordinates.
"""
import math

import numpy as np
import pytest

from oracle_bind import OraclePFU, OraclePFX, oracle_lbfgs
from test_pf_ext_gpu import device_pf, rel_err
from test_pf_ext_oracle import structured_solid
from test_pfu_oracle import unstructured_case

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5}


def oracle_of(c, x0=None):
    return OraclePFU(c["grid"], c["x0"] if x0 is None else x0, nfluid=c["nfluid"], resolution=c["resolution"],
                     surface_tension=c["surface_tension"], interface_size=c["interface_size"],
                     density=c["density"], **c["opts"])


def valid_case(seed):
    """unstructured_case for the first seed from `seed` the reference accepts."""
    while True:
        c = unstructured_case(seed)
        try:
            return c, oracle_of(c)
        except ValueError:
            seed += 1000


def energy_scale(c, x):
    """Sum of the magnitudes of the energy terms, from the gather form of the
    same physics (the element form's total can cancel)."""
    try:
        o = OraclePFX(c["grid"], c["x0"], nfluid=c["nfluid"], resolution=c["resolution"],
                      surface_tension=c["surface_tension"], interface_size=c["interface_size"],
                      density=c["density"], **c["opts"])
        return max(1.0, sum(abs(v[0]) for v in o.components(x).values()))
    except ValueError:
        return 1.0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("seed", range(24))
def test_element_form_vs_oracle(gpu, seed, prec):
    m = gpu
    c, o = valid_case(seed)
    print("case pfu", seed, c["grid"], c["nfluid"], c["density"], sorted(c["opts"]), prec, flush=True)
    x = c["x0"] + c["rng"].normal(0, 0.05, c["x0"].size)
    xs = x.astype(np.float32).astype(np.float64) if prec == "fp32" else x
    s = m.State(device_pf(m, c, "PhaseFieldUnstructured"), c["x0"], prec)
    s.coords(xs)
    eo, go = o.energy(xs), o.gradient(xs)
    es = max(energy_scale(c, xs), abs(eo))
    assert abs(s.energy() - eo) <= TOL[prec] * es, (s.energy(), eo)
    assert rel_err(s.gradient(), go) <= TOL[prec]
    e, g = s.energyGradient()
    assert abs(e - eo) <= TOL[prec] * es
    assert rel_err(g, go) <= TOL[prec]


@pytest.mark.parametrize("seed", range(8))
def test_element_form_constraints_vs_oracle(gpu, seed):
    import torch

    m = gpu
    c, o = valid_case(100 + seed)
    print("case pfu-constrain", seed, c["grid"], c["nfluid"], c["density"], flush=True)
    h = device_pf(m, c, "PhaseFieldUnstructured").handle(c["x0"])
    v0 = c["rng"].normal(0, 1, c["x0"].size)
    v = torch.tensor(v0, dtype=torch.float64, device="cuda")
    m.potential_constrain(h, "fp64", v.data_ptr(), v0.size)
    torch.cuda.synchronize()
    assert rel_err(v.cpu().numpy(), o.constrain(v0)) <= 1e-15


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_element_form_modes_agree(gpu, prec):
    """Energy-only, gradient-only and combined launches give the same energy
    (bit for bit: the energy reduction has a fixed order) and the same
    gradient (to the atomics' rounding); the energy at x + t*d equals the
    energy at the materialised point."""
    import torch

    m = gpu
    c, _ = valid_case(7)
    c["opts"]["pressure"] = [0.5, -0.25, 1.0, 0.75][: c["nfluid"]]
    pf = device_pf(m, c, "PhaseFieldUnstructured")
    pf.setSolid(structured_solid(c["grid"], None, "wall"))
    h = pf.handle(c["x0"])
    n = c["x0"].size
    dt = torch.float64 if prec == "fp64" else torch.float32
    x = torch.tensor(c["x0"], dtype=dt, device="cuda")
    d = torch.tensor(c["rng"].normal(0, 0.2, n), dtype=dt, device="cuda")
    g1, g2 = torch.empty_like(x), torch.empty_like(x)
    e1 = m.potential_eval(h, prec, x.data_ptr(), n, g_ptr=g1.data_ptr())
    m.potential_eval(h, prec, x.data_ptr(), n, g_ptr=g2.data_ptr(), want_energy=False)
    e3 = m.potential_eval(h, prec, x.data_ptr(), n)
    torch.cuda.synchronize()
    assert e1 == e3
    assert rel_err(g2.cpu().numpy(), g1.cpu().numpy()) <= (1e-15 if prec == "fp64" else 1e-6)
    t = 0.8125
    et = m.potential_eval(h, prec, x.data_ptr(), n, d_ptr=d.data_ptr(), t=t)
    xt = x + torch.tensor(t, dtype=dt, device="cuda") * d
    torch.cuda.synchronize()
    assert et == m.potential_eval(h, prec, xt.data_ptr(), n)


def test_element_form_unit_test_answers(gpu):
    """PhaseFieldUnstructured_test.cpp:27-65, 68-80, 104-129, 132-180 (the
    cases that do not erase elements) and 190-208 through the library."""
    m = gpu

    def pfu(grid, **kw):
        p = m.PhaseFieldUnstructured().setGridSize(list(grid))
        if "res" in kw:
            p.setResolution(kw["res"])
        if "solid" in kw:
            p.setSolid(kw["solid"])
        if "ca" in kw:
            p.setContactAngle(kw["ca"])
        if "pressure" in kw:
            p.setPressure(kw["pressure"])
        if "force" in kw:
            p.setForce(kw["force"])
        if "volume" in kw:
            p.setVolume(kw["volume"], kw["vol_const"])
        elif "vol_const" in kw:
            p.setVolumeFixed(True, kw["vol_const"])
        return p

    def near(a, b):
        np.testing.assert_allclose(np.asarray(a, float), np.asarray(b, float), rtol=1e-6, atol=1e-6)

    s = m.State(pfu((1, 1, 1)), [0.1])
    near(s.energy(), 3.0 / 16 * 1.21 * 0.81)
    near(s.gradient(), [-3.0 / 4 * 0.099])
    for axis in range(3):
        g = [1, 1, 1]
        g[axis] = 5
        s = m.State(pfu(g, solid=[1, 0, 0, 0, 1]), [0, -1, 0, 1, 0])
        near(s.energy(), 3.0 / 16 + 2 * (3.0 / 4))
        near(s.gradient(), [0, -1.5, 0, 1.5, 0])
        g = [1, 1, 1]
        g[axis] = 3
        s = m.State(pfu(g, solid=[0, 1, 0]), [-1, 0, 1])
        near(s.energy(), 3.0)
        near(s.gradient(), [-3, 0, 3])
    for f, gw in (([-2, 0, 0], [-.5] * 4 + [.5] * 4), ([0, -2, 0], [-.5, -.5, .5, .5, -.5, -.5, .5, .5]),
                  ([0, 0, -2], [-.5, .5, -.5, .5, -.5, .5, -.5, .5])):
        s = m.State(pfu((2, 2, 2), force=f), [1.0] * 8)
        near(s.energy(), 0)
        near(s.gradient(), gw)
    ones = [1.0] * 6
    wall = [1, 0, 0, 0, 0, 1]
    s = m.State(pfu((6, 1, 1), solid=wall, pressure=[10]), ones)
    near(s.energy(), -30)
    near(s.gradient(), [0, -2.5, -5, -5, -2.5, 0])
    s = m.State(pfu((6, 1, 1), solid=wall, vol_const=1), ones)
    near(s.energy(), 0)
    s = m.State(pfu((6, 1, 1), solid=wall, volume=[1], vol_const=100), ones)
    near(s.energy(), 100 * 2 ** 2)
    near(s.gradient(), [0] * 6)
    s = m.State(pfu((5, 1, 1), res=2.0, solid=[1, 0, 0, 0, 1]), [0, -1, 0, 1, 0])
    near(s.energy(), 1.5 / 16 * 8 + 2 * (6.0 / 4) * 2)
    near(s.gradient(), [0, -6, 0, 6, 0])
    s = m.State(pfu((6, 1, 1), res=2.0, solid=wall, ca=[90] * 6, pressure=[10]), ones)
    near(s.energy(), -240)
    near(s.gradient(), [0, -20, -40, -40, -20, 0])
    s = m.State(pfu((6, 1, 1), res=2.0, solid=wall, ca=[90] * 6, volume=[8], vol_const=100), ones)
    near(s.energy(), 100.0 / 16 * (2 * 8) ** 2)
    near(s.gradient(), [0] * 6)
    # a fixed fluid's DOFs are constrained (:190-208)
    p = m.PhaseFieldUnstructured().setNFluid(3).setGridSize([2, 2, 1]).setDensityConstraint("none").setFixFluid(0)
    s = m.State(p, [1.0] * 12)
    v = np.arange(1.0, 13.0)
    want = v.copy()
    want[0::3] = 0
    near(s.applyConstraints(v), want)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_element_form_equals_gather_form(gpu, prec):
    """Without solid nodes the two formulations are the same sum arranged
    differently (each bond's square split between its two sites here, counted
    once at its lower site in PhaseField): the library's two kernels agree to
    rounding on bulk, gradient, density and pressure terms."""
    m = gpu
    for grid, nf, st, iw, dens, pres in [((12, 10, 9), 1, [1.0], 1.0, "hard", None),
                                         ((9, 8, 20), 3, [1.0, 0.8, 1.2], 2.0, "hard", [0.3, -0.2, 0.1]),
                                         ((1, 40, 33), 2, [1.0], 1.5, "none", [0.5, 0.0]),
                                         ((6, 5, 4), 3, [1.0, 0.9, 1.1], 1.3, "soft", None)]:
        rng = np.random.default_rng(3)
        x = rng.uniform(-0.1, 1.1, int(np.prod(grid)) * nf)
        out = []
        for cls in ("PhaseField", "PhaseFieldUnstructured"):
            pf = getattr(m, cls)().setNFluid(nf).setGridSize(list(grid)).setSurfaceTension(st).setInterfaceSize(iw)
            pf.setDensityConstraint(dens)
            if pres is not None:
                pf.setPressure(pres)
            out.append(m.State(pf, x, prec).energyGradient())
        tol = 1e-12 if prec == "fp64" else 1e-5
        assert abs(out[1][0] - out[0][0]) <= tol * max(1.0, abs(out[0][0])), (grid, nf, out[1][0], out[0][0])
        assert rel_err(out[1][1], out[0][1]) <= tol, (grid, nf)


@pytest.mark.parametrize("seed", [1, 2, 7])
def test_lbfgs_element_form_vs_oracle(gpu, seed):
    """L-BFGS without line search through the element form follows the
    oracle's trajectory (the atomics' rounding differences stay far below the
    bound over 60 iterations)."""
    m = gpu
    c, o = valid_case(300 + seed)
    print("case pfu-lbfgs", seed, c["grid"], c["nfluid"], c["density"], sorted(c["opts"]), flush=True)
    xo, so = oracle_lbfgs(o, c["x0"], m=5, max_iter=60, linesearch="none")
    s = m.State(device_pf(m, c, "PhaseFieldUnstructured"), c["x0"])
    lb = m.Lbfgs().setM(5).setMaxIter(60).setLinesearch("none")
    x = lb.minimise(s)
    assert lb.iter == so["iterations"]
    assert np.max(np.abs(np.asarray(x) - xo)) <= 1e-8 * max(1.0, np.max(np.abs(xo)))


@pytest.mark.parametrize("seed", [44, 48, 49])
def test_lbfgs_backtracking_element_form_converges_like_oracle(gpu, seed):
    """L-BFGS with backtracking (energy-only evaluations at x + t*d) on
    layouts with solid nodes, forces, confinement and fixed fluids converges
    where the oracle does, to the same minimum."""
    m = gpu
    c, _ = valid_case(seed)
    for k in ("volume_fixed", "volume", "pressure", "force"):
        c["opts"].pop(k, None)
    o = oracle_of(c)
    print("case pfu-lbfgs-bt", seed, c["grid"], c["nfluid"], c["density"], sorted(c["opts"]), flush=True)
    xo, so = oracle_lbfgs(o, c["x0"], m=5, max_iter=2000)
    assert so["converged"]
    s = m.State(device_pf(m, c, "PhaseFieldUnstructured"), c["x0"])
    lb = m.Lbfgs().setM(5).setMaxIter(2000)
    x = np.asarray(lb.minimise(s))
    print("iterations", lb.iter, so["iterations"], flush=True)
    assert lb.iter < 2000
    assert abs(o.energy(x) - so["energy"]) <= 1e-8 * max(1.0, abs(so["energy"]))


@pytest.mark.parametrize("two_step", [False, True])
def test_element_form_diffuse_solid_vs_oracle(gpu, two_step):
    """PhaseFieldUnstructured::diffuseSolid (PhaseFieldUnstructured.cpp:968-994):
    100 L-BFGS steps without line search under confinement; bound 1e-6 per
    entry, as for PhaseField's (rounding differences grow along the
    trajectory)."""
    m = gpu
    grid = (8, 7, 1)
    solid = structured_solid(grid, None, "channel")
    sub = np.zeros(56, np.uint8)
    sub[(np.arange(56) // 7) < 3] = 1
    pf = (m.PhaseFieldUnstructured().setNFluid(3).setGridSize(list(grid)).setSurfaceTension([1.0, 0.8, 1.2])
          .setInterfaceSize(1.0).setSolid(solid))
    got = pf.diffuseSolid(sub, 0, two_step)
    init = np.zeros(56 * 3)
    init[0::3] = sub
    kw = dict(nfluid=3, surface_tension=[1.0, 0.8, 1.2], interface_size=[1.0], density="none", solid=solid,
              confinement=[100.0, 0.0, 0.0])
    x, _ = oracle_lbfgs(OraclePFU(grid, init, **kw), init, m=5, max_iter=100, linesearch="none", convergence=1e-8)
    if two_step:
        o2 = OraclePFU(grid, x, **kw)
        x, _ = oracle_lbfgs(o2, x, m=5, max_iter=100, linesearch="none", convergence=o2.convergence)
    assert np.max(np.abs(got - x)) <= 1e-6


def test_element_form_full_size_256cubed(gpu):
    """The headline input (256^3 sites x 3 fluids, 50.3M elements) through the
    element form: energy and constrained gradient equal the gather form's (the
    same sum arranged differently) in FP64, and the gradient satisfies the
    hard density constraint."""
    import torch

    m = gpu
    grid, nf = (256, 256, 256), 3
    n = 256 ** 3 * 3
    rng = np.random.default_rng(2024)
    u = rng.uniform(0, 1, (n // 3, 3))
    x = (u / u.sum(axis=1, keepdims=True)).reshape(-1)
    print("element form 256^3 ...", flush=True)
    out = {}
    xd = torch.from_numpy(x).cuda()
    for cls in ("PhaseField", "PhaseFieldUnstructured"):
        pf = getattr(m, cls)().setNFluid(nf).setGridSize(list(grid)).setSurfaceTension([1.0, 0.8, 1.2])
        pf.setInterfaceSize(2.0)
        h = pf.handle(x)
        g = torch.empty_like(xd)
        e = m.potential_eval(h, "fp64", xd.data_ptr(), n, g_ptr=g.data_ptr())
        torch.cuda.synchronize()
        out[cls] = (e, g)
        del h
    (e0, g0), (e1, g1) = out["PhaseField"], out["PhaseFieldUnstructured"]
    assert abs(e1 - e0) <= 1e-12 * abs(e0)
    assert (torch.linalg.norm(g1 - g0) / torch.linalg.norm(g0)).item() <= 1e-12
    assert g1.view(-1, 3).sum(dim=1).abs().max().item() < 1e-12


def test_element_form_mixed_schedule_widens(gpu):
    """The element form has no mixed FP64/FP32 evaluation: the mixed L-BFGS
    schedule runs its FP32 phase and then widens every vector, and converges
    to the FP64 run's minimum."""
    m = gpu
    c, o = valid_case(44)
    for k in ("volume_fixed", "volume", "pressure", "force"):
        c["opts"].pop(k, None)
    o = oracle_of(c)
    res = {}
    for mode in ("fp64", "mixed"):
        s = m.State(device_pf(m, c, "PhaseFieldUnstructured"), c["x0"])
        lb = m.Lbfgs().setM(5).setMaxIter(3000).setPrecision(mode)
        x = np.asarray(lb.minimise(s))
        assert lb.stats["converged"], mode
        res[mode] = o.energy(x)
    assert abs(res["mixed"] - res["fp64"]) <= 1e-6 * max(1.0, abs(res["fp64"]))
