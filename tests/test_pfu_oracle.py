"""Pins the oracle's PhaseFieldUnstructured (the element form of the phase
field, src/potentials/PhaseFieldUnstructured.cpp) before the GPU tests trust it:
the reference's unit-test known answers (test/unit/PhaseFieldUnstructured_test.cpp
:27-80, 104-129, 132-180 where no elements are erased), the reference build on random layouts (energies,
gradients, projections, element counts, L-BFGS trajectories), and finite
differences of the energy.
"""
import numpy as np
import pytest

from oracle_bind import OraclePFU, RefState, oracle_lbfgs, ref_available
from test_pf_ext_oracle import close, random_case

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref (reference build) not present")


def near(a, b, tol=1e-6):
    np.testing.assert_allclose(np.asarray(a, float), np.asarray(b, float), rtol=tol, atol=tol)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_bulk_known_answers(axis):
    # PhaseFieldUnstructured_test.cpp:27-65
    o = OraclePFU((1, 1, 1), [0.1])
    assert o.kappa[0] == 3
    near(o.energy([0.1]), 3.0 / 16 * 1.21 * 0.81)
    near(o.gradient([0.1]), [-3.0 / 4 * 0.099])
    grid = [1, 1, 1]
    grid[axis] = 5
    x = [0, -1, 0, 1, 0]
    o = OraclePFU(grid, x, solid=[1, 0, 0, 0, 1])
    near(o.energy(x), 3.0 / 16 + 2 * (3.0 / 4))
    near(o.gradient(x), [0, -1.5, 0, 1.5, 0])
    grid = [1, 1, 1]
    grid[axis] = 3
    x = [-1, 0, 1]
    o = OraclePFU(grid, x, solid=[0, 1, 0])
    near(o.energy(x), 3.0)
    near(o.gradient(x), [-3, 0, 3])


def test_force_known_answers():
    # :68-80
    x = [1.0] * 8
    for f, g in (([-2, 0, 0], [-.5] * 4 + [.5] * 4), ([0, -2, 0], [-.5, -.5, .5, .5, -.5, -.5, .5, .5]),
                 ([0, 0, -2], [-.5, .5, -.5, .5, -.5, .5, -.5, .5])):
        o = OraclePFU((2, 2, 2), x, force=f)
        near(o.energy(x), 0)
        near(o.gradient(x), g)


def test_pressure_and_volume_known_answers():
    # :104-129
    x = [1.0] * 6
    o = OraclePFU((6, 1, 1), x, solid=[1, 0, 0, 0, 0, 1], pressure=[10])
    near(o.energy(x), -30)
    near(o.gradient(x), [0, -2.5, -5, -5, -2.5, 0])
    o = OraclePFU((6, 1, 1), x, solid=[1, 0, 0, 0, 0, 1], volume_fixed=True, vol_const=1)
    near(o.volume, [3])
    near(o.energy(x), 0)
    o = OraclePFU((6, 1, 1), x, solid=[1, 0, 0, 0, 0, 1], volume=[1], vol_const=100)
    near(o.energy(x), 100 * 2 ** 2)
    near(o.gradient(x), [0] * 6)


def test_resolution_known_answers():
    # :132-180 (the cases that do not erase elements)
    x = [0, -1, 0, 1, 0]
    o = OraclePFU((5, 1, 1), x, resolution=2.0, solid=[1, 0, 0, 0, 1])
    assert o.kappa[0] == 1.5
    near(o.energy(x), 1.5 / 16 * 8 + 2 * (6.0 / 4) * 2)
    near(o.gradient(x), [0, -6, 0, 6, 0])
    x = [1.0] * 6
    o = OraclePFU((6, 1, 1), x, resolution=2.0, solid=[1, 0, 0, 0, 0, 1], contact_angle=[90] * 6, pressure=[10])
    near(o.energy(x), -240)
    near(o.gradient(x), [0, -20, -40, -40, -20, 0])
    o = OraclePFU((6, 1, 1), x, resolution=2.0, solid=[1, 0, 0, 0, 0, 1], contact_angle=[90] * 6,
                  volume=[8], vol_const=100)
    near(o.energy(x), 100.0 / 16 * (2 * 8) ** 2)
    near(o.gradient(x), [0] * 6)


def unstructured_case(seed):
    """A random_case the element form accepts: uniform tensions, and no soft
    constraint with two fluids (whose element the reference reads past)."""
    c = random_case(seed)
    if len(c["surface_tension"]) not in (1, c["nfluid"]):
        c["surface_tension"] = c["surface_tension"][: c["nfluid"]]
    if c["nfluid"] == 2 and c["density"] == "soft":
        c["density"] = "none"
    if seed % 4 == 3 and c["nfluid"] > 1:
        c["density"] = "fixed"  # the N-1 fluid form (FLUID_ENERGY_ALL)
    return c


def make_pfu_pair(c):
    try:
        o = OraclePFU(c["grid"], c["x0"], nfluid=c["nfluid"], resolution=c["resolution"],
                      surface_tension=c["surface_tension"], interface_size=c["interface_size"],
                      density=c["density"], **c["opts"])
    except ValueError:
        with pytest.raises(ValueError):
            RefState.phasefield_unstructured(c["grid"], c["x0"], c["nfluid"], c["resolution"],
                                             c["surface_tension"], c["interface_size"], c["density"], **c["opts"])
        return None, None
    r = RefState.phasefield_unstructured(c["grid"], c["x0"], c["nfluid"], c["resolution"], c["surface_tension"],
                                         c["interface_size"], c["density"], **c["opts"])
    return o, r


@needs_ref
@pytest.mark.parametrize("seed", range(40))
def test_element_form_vs_reference(seed):
    c = unstructured_case(seed)
    o, r = make_pfu_pair(c)
    if o is None:
        return
    assert o.nelements == r.nelements()
    rng = c["rng"]
    for x in (c["x0"], c["x0"] + rng.normal(0, 0.05, c["x0"].size)):
        close(o.energy(x), r.energy(x))
        close(o.gradient(x), r.gradient(x))
        e, g = r.energy_gradient(x)
        eo, go = o.energy_gradient(x)
        close(eo, e)
        close(o.constrain(go), g)
    v = rng.normal(0, 1, c["x0"].size)
    close(o.constrain(v), r.constrain(v))


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 7])
def test_lbfgs_element_form_vs_reference(seed):
    o = None
    while o is None:
        c = unstructured_case(300 + seed)
        o, r = make_pfu_pair(c)
        seed += 1000
    xo, so = oracle_lbfgs(o, c["x0"], m=5, max_iter=60, linesearch="none")
    xr, sr = r.minimise("lbfgs", m=5, max_iter=60, linesearch="none")
    assert so["iterations"] == sr["iterations"]
    close(xo, xr, 1e-9)


@pytest.mark.parametrize("seed", range(10))
def test_element_form_gradient_is_energy_derivative(seed):
    c = unstructured_case(400 + seed)
    c["opts"].pop("confinement", None)
    # the volume term's gradient is weighted towards interfaces, not the
    # derivative of its energy (PhaseFieldUnstructured.cpp:301-306)
    c["opts"].pop("volume_fixed", None)
    try:
        o = OraclePFU(c["grid"], c["x0"], nfluid=c["nfluid"], resolution=c["resolution"],
                      surface_tension=c["surface_tension"], interface_size=c["interface_size"],
                      density=c["density"], **c["opts"])
    except ValueError:
        return
    x = c["x0"]
    x = np.where(np.abs(x - 0.01) < 0.02, 0.2, x) if c["nfluid"] > 1 else np.where(np.abs(x + 0.98) < 0.04, 0.3, x)
    _, g = o.energy_gradient(x)
    rng = np.random.default_rng(seed)
    for i in rng.choice(x.size, size=min(8, x.size), replace=False):
        h = 1e-6
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (o.energy_gradient(xp)[0] - o.energy_gradient(xm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
