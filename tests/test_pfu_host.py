"""Host side of the library's PhaseFieldUnstructured (no GPU: init builds the
element list on the host and uploads it at the first evaluation): element
counts, coefficients, node volumes and convergence against the oracle, and the
reference's input errors."""
import numpy as np
import pytest

import minim_amd as m
from oracle_bind import OraclePFU
from test_pf_ext_gpu import device_pf
from test_pfu_oracle import unstructured_case


def oracle_or_none(c):
    try:
        return OraclePFU(c["grid"], c["x0"], nfluid=c["nfluid"], resolution=c["resolution"],
                         surface_tension=c["surface_tension"], interface_size=c["interface_size"],
                         density=c["density"], **c["opts"])
    except ValueError:
        return None


@pytest.mark.parametrize("seed", range(40))
def test_elements_and_coefficients_match_oracle(seed):
    c = unstructured_case(seed)
    o = oracle_or_none(c)
    pf = device_pf(m, c, "PhaseFieldUnstructured")
    if o is None:
        with pytest.raises(Exception):
            pf.nElements(c["x0"])
        return
    assert pf.nElements(c["x0"]) == o.nelements
    k, kp, conv = pf.coefficients(c["x0"])
    np.testing.assert_array_equal(k, o.kappa)
    assert conv == o.convergence
    np.testing.assert_array_equal(pf.nodeVolumes(c["x0"]), o.node_vol)


def test_element_form_rejects_what_it_cannot_hold():
    x = np.zeros(4 * 3 * 2 * 2)
    base = lambda: m.PhaseFieldUnstructured().setNFluid(2).setGridSize([4, 3, 2])  # noqa: E731
    with pytest.raises(ValueError, match="three fluids"):
        base().setDensityConstraint("soft").nElements(x)
    with pytest.raises(ValueError, match="surfaceTension"):
        base().setSurfaceTension([1.0] * 48).nElements(x)
    with pytest.raises(ValueError):
        base().setSurfaceTension(lambda i, j, k: [1.0, 1.0])
    with pytest.raises(ValueError, match="contactAngle"):
        base().setContactAngle([90.0] * 24).nElements(x)
    with pytest.raises(NotImplementedError):
        base().energyComponents(x)
