"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against the reference's own unit-test known
answers, against the golden vectors that the reference library wrote
(tests/golden/, made by make_golden.py from oracle/_ref), against the live
reference build when it is present, and against what the mathematics fixes
(hand-derived energies, finite differences, symmetries, invariants).
"""
import math
import os

import numpy as np
import pytest

from conftest import golden
from oracle_bind import (OracleBaH, OracleLJ, OracleLJPBC, OraclePF, RefState, oracle_dot, oracle_fire,
                         oracle_graddescent, oracle_lbfgs, ref_available)

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref (reference build) not present")


# ---------------------------------------------------------------- PhaseField
def test_pf_coefficients_known_answers():
    # PhaseField_test.cpp:36-37: sigma = w = 1 -> kappa = kappaP = 3
    o = OraclePF((1, 1, 1), 1, 1.0, [1.0], [])
    assert o.kappa[0] == 3 and o.kappaP[0] == 3
    assert o.convergence == 1e-8  # PhaseField.cpp:146
    # PhaseField_test.cpp:139-140: resolution 2 -> kappa 1.5, kappaP 6
    o = OraclePF((5, 1, 1), 1, 2.0, [1.0], [])
    assert o.kappa[0] == 1.5 and o.kappaP[0] == 6
    assert o.convergence == 4e-8
    # ternary combination of pairwise tensions (PhaseField.cpp:113-132)
    o = OraclePF((2, 2, 2), 3, 1.0, [1.0, 0.8, 1.2], [2.0, 2.0, 2.0])
    np.testing.assert_allclose(o.kappa, [0.9, 2.1, 1.5], rtol=1e-15)
    np.testing.assert_allclose(o.kappaP, [3.6, 8.4, 6.0], rtol=1e-15)


def test_pf_bulk_known_answer():
    # PhaseField_test.cpp:35-42: single site
    o = OraclePF((1, 1, 1), 1, 1.0, [1.0], [])
    assert o.energy([1.0]) == 0 and o.gradient([1.0])[0] == 0
    assert o.energy([0.1]) == pytest.approx(3.0 / 16 * 1.21 * 0.81, rel=1e-15)
    assert o.gradient([0.1])[0] == pytest.approx(-3.0 / 4 * 0.099, rel=1e-15)


def test_pf_periodic_hand_derived():
    # 3-site periodic chain phi = (-1, 0, 1), kappa = kappaP = 3, no solid:
    # bulk 3/16 at the middle site; gradient energy 0.75 * (2.5 + 1 + 2.5);
    # gradient 2f(2c - c- - c+) with 2f = 1.5 -> (-4.5, 0, 4.5)
    for grid in [(3, 1, 1), (1, 3, 1), (1, 1, 3)]:
        o = OraclePF(grid, 1, 1.0, [1.0], [])
        assert o.energy([-1, 0, 1]) == pytest.approx(3.0 / 16 + 4.5, rel=1e-15)
        np.testing.assert_allclose(o.gradient([-1, 0, 1]), [-4.5, 0, 4.5], rtol=0, atol=1e-15)


def test_pf_fix_fluid_known_answer():
    # PhaseField_test.cpp:182-200: fixed fluid 0 -> zero gradient
    o = OraclePF((2, 2, 1), 3, 1.0, [1.0], [], density="none", fix_fluid=[True, False, False])
    g = o.gradient(np.ones(12))
    assert np.all(g[0::3] == 0)


def test_pf_invalid_sizes():
    with pytest.raises(ValueError):
        OraclePF((2, 2, 2), 3, 1.0, [1.0, 2.0], [])  # PhaseField.cpp:86
    with pytest.raises(ValueError):
        OraclePF((2, 2, 2), 3, 1.0, [1.0], [1.0, 2.0])  # PhaseField.cpp:95 / :972


@pytest.mark.parametrize("case", range(9))
def test_pf_golden_vectors(case):
    c = golden("pf")[case]
    o = OraclePF(tuple(c["grid"]), c["nfluid"], c["resolution"], c["surface_tension"], c["interface_size"],
                 c["density"])
    x = np.array(c["x"])
    assert o.energy(x) == c["energy"]  # bit-exact with the single-thread reference
    np.testing.assert_array_equal(o.gradient(x), np.array(c["gradient"]))


def test_pf_finite_difference_and_symmetry():
    rng = np.random.default_rng(1)
    for grid, nf, st, dens in [((5, 4, 3), 1, [1.0], "none"), ((4, 3, 3), 3, [1.0, 0.8, 1.2], "none"),
                               ((3, 4, 2), 2, [0.9], "soft")]:
        o = OraclePF(grid, nf, 1.3, st, [1.7], dens)
        x = rng.uniform(-0.8, 1.2, o.ndof)
        g = o.gradient(x, constrained=False)
        h = 1e-6
        for i in rng.choice(o.ndof, 8, replace=False):
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            fd = (o.energy(xp) - o.energy(xm)) / (2 * h)
            assert fd == pytest.approx(g[i], rel=1e-6, abs=1e-7)
    # binary double well is even: E(-phi) = E(phi), g(-phi) = -g(phi)
    o = OraclePF((6, 5, 1), 1, 1.0, [1.0], [1.0])
    x = rng.uniform(-1, 1, o.ndof)
    assert o.energy(-x) == pytest.approx(o.energy(x), rel=1e-14)
    np.testing.assert_allclose(o.gradient(-x), -o.gradient(x), rtol=1e-13, atol=1e-14)


def test_pf_translation_invariance_and_projection():
    rng = np.random.default_rng(2)
    o = OraclePF((5, 4, 3), 3, 1.0, [1.0, 0.8, 1.2], [2.0])
    x = rng.uniform(0, 1, o.ndof)
    xs = np.roll(x.reshape(5, 4, 3, 3), shift=(2, -1, 1), axis=(0, 1, 2)).reshape(-1)
    assert o.energy(xs) == pytest.approx(o.energy(x), rel=1e-13)
    gs = np.roll(o.gradient(x).reshape(5, 4, 3, 3), shift=(2, -1, 1), axis=(0, 1, 2)).reshape(-1)
    np.testing.assert_allclose(o.gradient(xs), gs, rtol=1e-12, atol=1e-13)
    # hard density constraint: components of every site's gradient sum to zero
    assert np.abs(o.gradient(x).reshape(-1, 3).sum(axis=1)).max() < 1e-13


@needs_ref
def test_pf_oracle_vs_live_reference():
    rng = np.random.default_rng(3)
    for grid, nf, st, iw, dens in [((7, 5, 3), 1, [1.0], [1.2], "hard"), ((3, 6, 4), 3, [1.0, 0.8, 1.2], [2.0], "hard"),
                                   ((2, 2, 9), 4, [1.0, 0.9, 1.1, 0.8], [1.3], "soft")]:
        x = rng.uniform(-0.2, 1.1, int(np.prod(grid)) * nf)
        o = OraclePF(grid, nf, 1.0, st, iw, dens)
        r = RefState.phasefield(grid, x, nf, 1.0, st, iw, dens)
        if os.environ.get("OMP_NUM_THREADS") == "1":  # the reference's only deterministic mode
            assert o.energy(x) == r.energy(x)
            np.testing.assert_array_equal(o.gradient(x), r.gradient(x))
        else:
            assert o.energy(x) == pytest.approx(r.energy(x), rel=1e-13)
            np.testing.assert_allclose(o.gradient(x), r.gradient(x), rtol=1e-12, atol=1e-13)


# ---------------------------------------------------------------- minimisers
@pytest.mark.parametrize("case", range(7))
def test_lbfgs_golden_trajectory(case):
    c = golden("lbfgs")[case]
    o = OraclePF(tuple(c["grid"]), c["nfluid"], 1.0, c["surface_tension"], c["interface_size"])
    x, st = oracle_lbfgs(o, np.array(c["x0"]), m=c["m"], max_iter=c["max_iter"], max_step=c["max_step"],
                         init_step=c["init_step"], linesearch=c["linesearch"])
    assert st["iterations"] == c["iterations"]
    assert st["n_energy"] == c["n_energy"] and st["n_gradient"] == c["n_gradient"]
    np.testing.assert_array_equal(x, np.array(c["x"]))
    assert st["energy"] == c["energy"]


def test_mt19937_64_known_answers():
    """The input generator reproduces C++ std::mt19937_64 (the standard's
    required 10000th output for seed 5489, and libstdc++'s first outputs for
    the benchmark seed)."""
    from mt64 import MT19937_64, cfg1_input

    g = MT19937_64(5489)
    v = [g.next() for _ in range(10000)]
    assert v[-1] == 9981545732273789042
    h = MT19937_64(20240)
    assert [h.next() for _ in range(3)] == [2090563417169947710, 6791348416396944384, 8999006810243764886]
    x = cfg1_input()
    assert x.shape == (16384,) and x.min() >= -0.1 and x.max() < 0.1


def test_cfg1_lbfgs_golden():
    """128x128x1 binary field, L-BFGS m=5 to the reference tolerance: the
    oracle reproduces the reference's run bit for bit."""
    from mt64 import uniform

    c = golden("cfg1")[0]
    x0 = uniform(c["seed"], int(np.prod(c["grid"])), c["lo"], c["hi"])
    o = OraclePF(tuple(c["grid"]), 1, 1.0, c["surface_tension"], c["interface_size"])
    assert o.energy(x0) == c["initial_energy"]
    x, st = oracle_lbfgs(o, x0, m=c["m"], max_iter=c["max_iter"])
    assert st["iterations"] == c["iterations"] and st["converged"]
    assert st["n_energy"] == c["n_energy"] and st["n_gradient"] == c["n_gradient"]
    assert st["energy"] == c["energy"]
    assert float(x.sum()) == c["x_sum"]


@pytest.mark.parametrize("case", range(4))
def test_fire_graddescent_golden_trajectory(case):
    c = golden("fire_gd")[case]
    o = OraclePF(tuple(c["grid"]), 1, 1.0, [1.0], c["interface_size"])
    if c["kind"] == "fire":
        x, st = oracle_fire(o, np.array(c["x0"]), dt_max=c["param"], max_iter=c["max_iter"],
                            linesearch=c["linesearch"])
    else:
        x, st = oracle_graddescent(o, np.array(c["x0"]), alpha=c["param"], max_iter=c["max_iter"],
                                   linesearch=c["linesearch"])
    assert st["iterations"] == c["iterations"]
    np.testing.assert_array_equal(x, np.array(c["x"]))


def test_fire_dt0_and_max_force_options():
    """orc_fire_minimise_ex: with both options off it is orc_fire_minimise bit for
    bit; with dt0 the first step is -0.5 dt0^2 g (v = -dt0 g, then dt halves,
    Fire.cpp:50-67); with max-force convergence the largest per-particle force
    of the result is below the tolerance."""
    import ctypes as C

    from oracle_bind import OrcStats, _d, lib

    rng = np.random.default_rng(8)
    L = 5 * 1.1
    base = np.stack(np.meshgrid(*[np.arange(5) * 1.1] * 3, indexing="ij"), -1).reshape(-1, 3)
    x0 = (base + rng.uniform(-0.08, 0.08, base.shape)).reshape(-1)
    o = OracleLJPBC(125, L, 2.5)
    xa, sa = oracle_fire(o, x0, dt_max=0.01, max_iter=300, linesearch="none", convergence=1e-6)
    xb = np.array(x0)
    st = OrcStats()
    dto = C.c_double(0)
    lib().orc_fire_minimise_ex(C.byref(o.pot), _d(xb), 0.01, 0.0, 0, 300, 1, 1e-6, C.byref(st), C.byref(dto))
    np.testing.assert_array_equal(xa, xb)
    x1, _ = oracle_fire(o, x0, dt_max=0.02, dt0=0.002, max_force_dim=3, max_iter=0, linesearch="none")
    np.testing.assert_allclose(x1, x0 - 0.5 * 0.002 ** 2 * o.gradient(x0), rtol=1e-15, atol=1e-15)
    xc, sc = oracle_fire(o, x0, dt_max=0.02, dt0=0.002, max_force_dim=3, max_iter=20000, linesearch="none",
                         convergence=1e-6)
    assert sc["converged"]
    assert np.sqrt((o.gradient(xc).reshape(-1, 3) ** 2).sum(axis=1)).max() < 1e-6


def test_fire_first_step_known_answer():
    # Fire_test.cpp:25-40 on a PhaseField state: with dtMax estimated as
    # 0.1/sqrt(|g|), the first step is exactly (0.5 dtMax) * (-dtMax * g).
    rng = np.random.default_rng(4)
    o = OraclePF((6, 6, 1), 1, 1.0, [1.0], [1.0])
    x0 = rng.uniform(-0.5, 0.5, o.ndof)
    g = o.gradient(x0)
    x, st = oracle_fire(o, x0, dt_max=0.0, max_iter=0)
    dt = 0.1 / math.sqrt(math.sqrt(oracle_dot(g, g)))
    assert st["dt_max"] == dt
    np.testing.assert_array_equal(x, x0 + (0.5 * dt) * (-dt * g))


def test_lbfgs_first_step_and_quadratic():
    # Lbfgs.cpp:91-96: iteration 0 moves by -g * initStep / |g| (no line search)
    rng = np.random.default_rng(5)
    o = OraclePF((5, 5, 1), 1, 1.0, [1.0], [1.0])
    x0 = rng.uniform(-0.3, 0.3, o.ndof)
    g = o.gradient(x0)
    x, _ = oracle_lbfgs(o, x0, max_iter=0, init_step=0.01, linesearch="none")
    np.testing.assert_array_equal(x, x0 + (-g) * (0.01 / math.sqrt(oracle_dot(g, g))))


# ---------------------------------------------------------------- Lennard-Jones
def test_lj_dimer_known_answer():
    # State_test.cpp:33-51: dimer at r = 2
    o = OracleLJ(3, 2)
    x = [0, 0, 0, 2, 0, 0]
    assert o.energy(x) == -63.0 / 1024
    np.testing.assert_allclose(o.gradient(x), [-93.0 / 512, 0, 0, 93.0 / 512, 0, 0], rtol=1e-15)


@pytest.mark.parametrize("case", range(3))
def test_lj_golden_vectors(case):
    c = golden("lj")[case]
    o = OracleLJ(c["ndim"], c["n"])
    x = np.array(c["x"])
    assert o.energy(x) == pytest.approx(c["energy"], rel=1e-14)
    np.testing.assert_allclose(o.gradient(x), c["gradient"], rtol=1e-13, atol=1e-14)


def test_ljpbc_pins():
    rng = np.random.default_rng(6)
    n = 20
    x = rng.uniform(0, 3.0, 3 * n)
    # rc beyond every distance and a box so large that no image is nearer:
    # truncated-shifted = plain all-pairs minus one shift per pair
    rc, box = 50.0, 1000.0
    p = OracleLJPBC(n, box, rc)
    a = OracleLJ(3, n)
    src6 = (1 / rc) ** 6
    shift = 4 * (src6 * src6 - src6)
    assert p.energy(x) == pytest.approx(a.energy(x) - shift * n * (n - 1) / 2, rel=1e-12)
    np.testing.assert_allclose(p.gradient(x), a.gradient(x), rtol=1e-12, atol=1e-13)
    # periodic: energy and forces invariant under whole-box shifts of particles
    L = 6.0
    q = OracleLJPBC(n, L, 2.5)
    y = rng.uniform(0, L, 3 * n)
    y2 = y + L * rng.integers(-2, 3, y.shape)
    assert q.energy(y2) == pytest.approx(q.energy(y), rel=1e-12)
    # momentum conservation and finite differences
    g = q.gradient(y)
    assert np.abs(g.reshape(-1, 3).sum(axis=0)).max() < 1e-10 * max(1.0, np.abs(g).max())
    h = 1e-6
    for i in rng.choice(3 * n, 6, replace=False):
        yp, ym = y.copy(), y.copy()
        yp[i] += h
        ym[i] -= h
        assert (q.energy(yp) - q.energy(ym)) / (2 * h) == pytest.approx(g[i], rel=1e-5, abs=1e-6)
    # shifted pair potential is zero at the cutoff and continuous
    two = OracleLJPBC(2, 100.0, 2.5)
    e_in = two.energy([0, 0, 0, 2.5 - 1e-9, 0, 0])
    e_out = two.energy([0, 0, 0, 2.5 + 1e-9, 0, 0])
    assert abs(e_in) < 1e-9 and e_out == 0


# ---------------------------------------------------------------- BarAndHinge
def _bah_from_case(c):
    nnode = len(c["x"]) // 3
    fixed = None
    if c["fixed"]:
        fixed = np.zeros(3 * nnode, dtype=np.uint8)
        fixed[c["fixed"]] = 1
    force = None if c["force"] is None else np.array(c["force"])
    return OracleBaH(nnode, c["bars"], c["bar_k"], c["bar_l0"], c["hinges"], c["hinge_k"], c["hinge_t0"],
                     force=force, fixed=fixed)


@pytest.mark.parametrize("case", range(2))
def test_bah_golden_vectors(case):
    c = golden("bah")[case]
    o = _bah_from_case(c)
    x = np.array(c["x"])
    assert o.energy(x) == pytest.approx(c["energy"], rel=1e-13)
    np.testing.assert_allclose(o.gradient(x), c["gradient"], rtol=1e-12, atol=1e-13)


def test_bah_gradient_quirk_and_fd():
    # bending gradient = derivative of its energy; stretching gradient = 2x it
    # (BarAndHinge.cpp:279 vs :282)
    c = golden("bah")[0]
    nnode = len(c["x"]) // 3
    x = np.array(c["x"])
    bend = OracleBaH(nnode, np.zeros((0, 2), int), [], [], c["hinges"], c["hinge_k"], c["hinge_t0"])
    bar = OracleBaH(nnode, c["bars"], c["bar_k"], c["bar_l0"], np.zeros((0, 4), int), [], [])
    h = 1e-6
    rng = np.random.default_rng(7)
    for i in rng.choice(3 * nnode, 6, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd_b = (bend.energy(xp) - bend.energy(xm)) / (2 * h)
        fd_s = (bar.energy(xp) - bar.energy(xm)) / (2 * h)
        assert bend.gradient(x)[i] == pytest.approx(fd_b, rel=1e-5, abs=1e-8)
        assert bar.gradient(x)[i] == pytest.approx(2 * fd_s, rel=1e-5, abs=1e-8)


def test_bah_minima_known_answers():
    # BarAndHinge_test.cpp:13-56 expected minima, within 1e-4
    c = golden("bah")[2]
    dz = math.sqrt(2) / 2
    s2 = math.sqrt(2)
    expected = [[0, 0, 0.5 + dz, 0, -s2, 0.5 - dz, 0, s2, 0.5 - dz, 0, 0, 0.5 + dz],
                [-1, 0, 1, 0, -s2, 0, 0, s2, 0, 1, 0, 1],
                [-s2, 0, 0.5, 0, -s2, 0.5, 0, s2, 0.5, s2, 0, 0.5],
                [-1, 0, 0, 0, -s2, 1, 0, s2, 1, 1, 0, 0],
                [0, 0, 0.5 - dz, 0, -s2, 0.5 + dz, 0, s2, 0.5 + dz, 0, 0, 0.5 - dz]]
    for m, ex in zip(c["minima"], expected):
        np.testing.assert_allclose(m["x"], ex, atol=1e-4)


@pytest.mark.parametrize("case", range(2))
def test_interface_minima_lbfgs_golden(case):
    """L-BFGS runs whose minimum has interfaces (E far from 0): the oracle
    reproduces the reference's run bit for bit, and the converged state is a
    stationary point of the oracle's energy."""
    from mt64 import minima_input

    c = golden("minima")[case]
    x0 = minima_input(c)
    o = OraclePF(tuple(c["grid"]), c["nfluid"], 1.0, c["surface_tension"], c["interface_size"])
    assert o.energy(x0) == c["initial_energy"]
    x, st = oracle_lbfgs(o, x0, m=c["m"], max_iter=c["max_iter"])
    assert st["iterations"] == c["iterations"] and st["converged"]
    assert st["energy"] == c["energy"]
    prof = x.reshape(c["grid"][0], -1, c["nfluid"]).mean(axis=1)
    np.testing.assert_array_equal(prof, np.array(c["profile"]))
    g = o.gradient(x)
    assert math.sqrt(np.dot(g, g) / g.size) == c["grad_rms_at_x"] < o.convergence
    # the interfaces carry the energy: E ~ sum of tension x area
    area = c["grid"][1] * c["grid"][2]
    tension = 2 * c["surface_tension"][0] if c["nfluid"] == 1 else sum(c["surface_tension"])
    assert abs(c["energy"] - tension * area) < 0.02 * tension * area
