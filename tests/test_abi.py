"""CPU checks of the C ABI boundary: the library loads, exports every symbol
include/minim_hip.h declares, and host-only logic (parameter set-up, error
mapping) behaves like the reference without touching a GPU."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "minim_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void\*|const char\*)\s*\*?\s*(minim_[a-z0-9_]+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from minim_amd import _capi

    L = _capi.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_capi.SIGNATURES), "ctypes table and header disagree"
    assert L.minim_abi_version() == 10


def _pf_handle(grid, nf, st, iw, density=0, res=1.0, fix=None):
    from minim_amd import _capi as A

    L = A.load()
    p = A.PFParams()
    for i in range(3):
        p.grid[i] = grid[i]
    p.n_fluid = nf
    p.resolution = res
    sta = np.ascontiguousarray(st, dtype=np.float64)
    iwa = np.ascontiguousarray(iw if iw else [0.0], dtype=np.float64)
    p.surface_tension = sta.ctypes.data_as(A.dp)
    p.n_surface_tension = sta.size
    p.interface_size = iwa.ctypes.data_as(A.dp)
    p.n_interface_size = len(iw)
    p.density_constraint = density
    p.density_const = 1.0
    if fix is not None:
        fx = (C.c_ubyte * nf)(*fix)
        p.fix_fluid = fx
    h = C.c_void_p()
    rc = L.minim_pf_create(C.byref(p), C.byref(h))
    return rc, h


def test_pf_setup_matches_reference_coefficients():
    """PhaseField::init runs on the host (no GPU needed): kappa, kappaP and the
    convergence criterion must match the oracle/reference values."""
    from minim_amd import _capi as A
    from oracle_bind import OraclePF

    L = A.load()
    for grid, nf, st, iw, res in [((1, 1, 1), 1, [1.0], [], 1.0), ((5, 1, 1), 1, [1.0], [], 2.0),
                                  ((4, 4, 4), 3, [1.0, 0.8, 1.2], [2.0], 1.0),
                                  ((3, 3, 3), 4, [1.0, 0.9, 1.1, 0.8], [1.3], 0.7),
                                  ((2, 2, 2), 2, [0.6], [1.4], 1.0)]:
        rc, h = _pf_handle(grid, nf, st, iw, res=res)
        assert rc == 0
        n = int(np.prod(grid)) * nf
        x = np.zeros(n)
        assert L.minim_potential_init(h, x.ctypes.data_as(A.dp), n) == 0
        k, kp, conv = np.zeros(nf), np.zeros(nf), C.c_double()
        assert L.minim_pf_coefficients(h, k.ctypes.data_as(A.dp), kp.ctypes.data_as(A.dp)) == 0
        assert L.minim_potential_convergence(h, C.byref(conv)) == 0
        o = OraclePF(grid, nf, res, st, iw)
        np.testing.assert_array_equal(k, o.kappa)
        np.testing.assert_array_equal(kp, o.kappaP)
        assert conv.value == o.convergence
        L.minim_potential_destroy(h)


def test_pf_errors_map_to_invalid_argument():
    from minim_amd import _capi as A

    L = A.load()
    # wrong number of surface tensions (PhaseField.cpp:86)
    rc, h = _pf_handle((2, 2, 2), 3, [1.0, 2.0], [])
    assert rc == 0
    x = np.zeros(24)
    assert L.minim_potential_init(h, x.ctypes.data_as(A.dp), 24) == A.MINIM_EINVAL
    assert b"surfaceTension" in L.minim_last_error()
    # coordinates that do not match the grid (PhaseField.cpp:138-140)
    rc, h = _pf_handle((2, 2, 2), 1, [1.0], [])
    assert L.minim_potential_init(h, x.ctypes.data_as(A.dp), 7) == A.MINIM_EINVAL
    assert b"does not match" in L.minim_last_error()
    # bad density constraint name / code (PhaseField.cpp:802-815)
    rc, h = _pf_handle((2, 2, 2), 1, [1.0], [], density=7)
    assert rc == A.MINIM_EINVAL
    # evaluating before init is a runtime error
    rc, h = _pf_handle((2, 2, 2), 1, [1.0], [])
    assert L.minim_potential_eval(h, 0, C.c_void_p(1), 8, None, 0.0, None, None, None) == A.MINIM_ERUNTIME


def test_python_mirror_setters():
    import minim_amd as m

    pf = m.PhaseField().setNFluid(3).setGridSize([2, 3, 4]).setSurfaceTension([1, 0.8, 1.2]).setInterfaceSize(2.0)
    assert pf.interfaceSize == [2.0, 2.0, 2.0]  # scalar setter sized by the current nFluid
    k, kp, conv = pf.coefficients()
    np.testing.assert_allclose(k, [0.9, 2.1, 1.5], rtol=1e-15)
    np.testing.assert_allclose(kp, [3.6, 8.4, 6.0], rtol=1e-15)
    assert conv == pytest.approx(1e-8)
    with pytest.raises(ValueError):
        m.PhaseField().setDensityConstraint("bogus")
    with pytest.raises(ValueError):
        m.Lbfgs().setLinesearch("bogus")
    lb = m.Lbfgs().setMaxIter(10)
    assert lb.maxIter == 10  # Lbfgs_test.cpp:7-10
    with pytest.raises(ValueError):
        m.Lbfgs().setM(99)._handle()


def test_triangulation_lists_match_reference_order():
    from minim_amd import triangulation_lists
    from conftest import golden

    c = golden("bah")[0]
    bonds, hinges = triangulation_lists(c["tri"])
    np.testing.assert_array_equal(bonds, c["bars"])
    np.testing.assert_array_equal(hinges, c["hinges"])


def test_site_types_match_oracle():
    """The library's host classification of solid layouts (PhaseField::getType)
    equals the oracle's, site for site, and rejects the same layouts."""
    import numpy as np
    from minim_amd import site_types
    from oracle_bind import oracle_site_types
    from test_pf_ext_oracle import structured_solid

    rng = np.random.default_rng(4)
    for grid in [(6, 7, 5), (8, 6, 1), (1, 9, 7), (5, 2, 4), (4, 4, 4), (3, 1, 1)]:
        for kind in ["wall", "channel", "pillar", "box"]:
            s = structured_solid(grid, rng, kind)
            assert np.array_equal(site_types(grid, s), oracle_site_types(grid, s))
    # random sparse solids: both accept or both reject
    for seed in range(60):
        r = np.random.default_rng(seed)
        grid = (5, 6, 4)
        s = (r.random(120) < 0.03).astype(np.uint8)
        try:
            want = oracle_site_types(grid, s)
        except RuntimeError:
            with pytest.raises(Exception, match="Undefined surface type"):
                site_types(grid, s)
            continue
        assert np.array_equal(site_types(grid, s), want)


def test_element_form_and_fused_step_errors():
    """Host-side checks of the element-form and fused-step entry points (no
    kernel is launched): the wrong potential kind, use before init and NULL
    pointers map to the reference's exception kinds."""
    from minim_amd import _capi as A

    L = A.load()
    rc, h = _pf_handle((2, 2, 2), 1, [1.0], [])
    assert rc == 0
    n, b = C.c_size_t(), C.c_size_t()
    # not a PhaseFieldUnstructured
    assert L.minim_pfu_element_list(h, C.byref(n), C.byref(b)) == A.MINIM_EINVAL
    # the fused step before init, and with a NULL direction
    assert L.minim_potential_eval_step(h, 0, 0, C.c_void_p(1), 8, C.c_void_p(1), 1.0, C.c_void_p(1), None,
                                       C.c_void_p(1), None) == A.MINIM_ERUNTIME
    assert L.minim_potential_eval_step(h, 0, 0, C.c_void_p(1), 8, None, 1.0, C.c_void_p(1), None,
                                       C.c_void_p(1), None) == A.MINIM_EINVAL
    L.minim_potential_destroy(h)
    p = A.PFParams()
    for i in range(3):
        p.grid[i] = 2
    p.n_fluid = 1
    p.resolution = 1.0
    st = np.ones(1)
    p.surface_tension = st.ctypes.data_as(A.dp)
    p.n_surface_tension = 1
    p.density_const = 1.0
    hu = C.c_void_p()
    assert L.minim_pfu_create(C.byref(p), C.byref(hu)) == 0
    # element list before init
    assert L.minim_pfu_element_list(hu, C.byref(n), C.byref(b)) == A.MINIM_ERUNTIME
    x = np.zeros(8)
    assert L.minim_potential_init(hu, x.ctypes.data_as(A.dp), 8) == 0
    assert L.minim_pfu_element_list(hu, C.byref(n), C.byref(b)) == 0
    assert n.value == 8 and b.value > 0
    # the element form has no fused step
    assert L.minim_potential_eval_step(hu, 0, 0, C.c_void_p(1), 8, C.c_void_p(1), 1.0, C.c_void_p(1), None,
                                       C.c_void_p(1), None) == A.MINIM_EINVAL
    # node volumes: no surface areas in the element form
    nv = np.zeros(8)
    assert L.minim_pf_node_volumes(hu, nv.ctypes.data_as(A.dp), None) == 0
    np.testing.assert_array_equal(nv, np.ones(8))
    assert L.minim_pf_node_volumes(hu, nv.ctypes.data_as(A.dp), nv.ctypes.data_as(A.dp)) == A.MINIM_EINVAL
    L.minim_potential_destroy(hu)


def test_batch_entry_point_errors():
    """minim_minimise_batch without a GPU: an empty batch is a no-op, a negative
    count, NULL states and a minimiser other than L-BFGS are rejected."""
    from minim_amd import _capi as A

    L = A.load()
    p = A.MinimiserParams()
    p.kind = A.MINIM_LBFGS
    p.max_iter = 10
    p.m = 5
    p.init_step = 1.0
    h = C.c_void_p()
    assert L.minim_minimiser_create(C.byref(p), C.byref(h)) == 0
    assert L.minim_minimise_batch(h, None, 0, None) == 0
    assert L.minim_minimise_batch(h, None, -1, None) == A.MINIM_EINVAL
    assert L.minim_minimise_batch(h, None, 2, None) == A.MINIM_EINVAL
    arr = (C.c_void_p * 1)(None)
    assert L.minim_minimise_batch(h, arr, 1, None) == A.MINIM_EINVAL
    L.minim_minimiser_destroy(h)
    p.kind = A.MINIM_FIRE
    assert L.minim_minimiser_create(C.byref(p), C.byref(h)) == 0
    assert L.minim_minimise_batch(h, None, 0, None) == A.MINIM_EINVAL
    assert b"Lbfgs" in L.minim_last_error()
    L.minim_minimiser_destroy(h)
    # Lbfgs::setBatchBudget rejects a negative budget
    p.kind = A.MINIM_LBFGS
    p.batch_budget = -1
    assert L.minim_minimiser_create(C.byref(p), C.byref(h)) == A.MINIM_EINVAL
    assert b"budget" in L.minim_last_error()
    with pytest.raises(ValueError):
        __import__("minim_amd").Lbfgs().setBatchBudget(-1)


def test_batch_control_values():
    """batch_control 0..3 (device, graph, host, resident) are accepted and
    anything else is rejected, in the C ABI and the Python mirror."""
    from minim_amd import _capi as A

    import minim_amd as m

    L = A.load()
    p = A.MinimiserParams()
    p.kind = A.MINIM_LBFGS
    p.max_iter = 10
    p.m = 5
    p.init_step = 1.0
    h = C.c_void_p()
    for v in range(4):
        p.batch_control = v
        assert L.minim_minimiser_create(C.byref(p), C.byref(h)) == 0, v
        L.minim_minimiser_destroy(h)
    for v in (-1, 4):
        p.batch_control = v
        assert L.minim_minimiser_create(C.byref(p), C.byref(h)) == A.MINIM_EINVAL, v
    for mode in ("graph", "device", "host", "resident"):
        m.Lbfgs().setBatchControl(mode)._handle()
    with pytest.raises(ValueError):
        m.Lbfgs().setBatchControl("bogus")
