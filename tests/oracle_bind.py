"""ctypes bindings to the checker: oracle/liboracle.so (C restatement) and
oracle/_ref/libminim_ref.so (the reference compiled from its own sources).

Test infrastructure only.  The product package (minim_amd) never imports this.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
# the reference build (bit-exact single-threaded); MINIM_REF_SO may name
# another build of the same sources (bench.py's CPU baseline uses the native-ISA one)
REF_SO = os.environ.get("MINIM_REF_SO") or os.path.join(ROOT, "oracle", "_ref", "libminim_ref.so")

DENSITY = {"hard": 0, "gradient": 0, "soft": 1, "energy penalty": 1, "none": 2, "fixed": 3}
ORC_MAXF = 16

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


class OrcPot(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("ndof", C.c_size_t), ("eg", C.c_void_p),
                ("constrain", C.c_void_p), ("convergence", C.c_double)]


class OrcPF(C.Structure):
    _fields_ = [("grid", C.c_int * 3), ("nfluid", C.c_int), ("resolution", C.c_double),
                ("kappa", C.c_double * ORC_MAXF), ("kappaP", C.c_double * ORC_MAXF),
                ("st_mean", C.c_double), ("density", C.c_int), ("density_const", C.c_double),
                ("fix", C.c_ubyte * ORC_MAXF), ("convergence", C.c_double)]


class OrcPFX(C.Structure):
    _fields_ = [("pf", OrcPF), ("ngrid", C.c_size_t), ("nkgrid", C.c_int), ("kappa", _dp), ("kappaP", _dp),
                ("solid", C.POINTER(C.c_ubyte)), ("type", _ip), ("node_vol", _dp), ("surface_area", _dp),
                ("contact_angle", _dp), ("pressure", C.c_double * ORC_MAXF), ("any_force", C.c_int),
                ("fmag", C.c_double * ORC_MAXF), ("fnorm", (C.c_double * 3) * ORC_MAXF),
                ("conf", C.c_double * ORC_MAXF), ("ff_init", _dp), ("volume_fixed", C.c_int),
                ("vol_const", C.c_double), ("volume", C.c_double * ORC_MAXF), ("has_volume", C.c_int),
                ("total_volume2", C.c_double)]


ORC_PFU_NIDOF = 16


class OrcPFU(C.Structure):
    _fields_ = [("grid", C.c_int * 3), ("nfluid", C.c_int), ("resolution", C.c_double), ("density", C.c_int),
                ("density_const", C.c_double), ("kappa", C.c_double * ORC_MAXF), ("kappaP", C.c_double * ORC_MAXF),
                ("st_mean", C.c_double), ("convergence", C.c_double), ("fix", C.c_ubyte * ORC_MAXF),
                ("ngrid", C.c_size_t), ("node_vol", _dp), ("volume_fixed", C.c_int), ("vol_const", C.c_double),
                ("volume", C.c_double * ORC_MAXF), ("nel", C.c_size_t), ("el_type", _ip), ("el_idof", _ip),
                ("el_par", _dp)]


class OrcStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("n_energy", C.c_long),
                ("n_gradient", C.c_long), ("energy", C.c_double), ("grad_rms", C.c_double)]


class OrcLJ(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nparticle", C.c_int), ("sigma", C.c_double), ("epsilon", C.c_double)]


class OrcLJPBC(C.Structure):
    _fields_ = [("nparticle", C.c_int), ("sigma", C.c_double), ("epsilon", C.c_double),
                ("rc", C.c_double), ("box", C.c_double)]


class OrcBaH(C.Structure):
    _fields_ = [("nnode", C.c_int),
                ("nbar", C.c_int), ("bar", _ip), ("bar_k", _dp), ("bar_l0", _dp),
                ("nhinge", C.c_int), ("hinge", _ip), ("hinge_k", _dp), ("hinge_t0", _dp),
                ("force", _dp), ("force_per_node", C.c_int), ("fixed", C.POINTER(C.c_ubyte))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        L.orc_pf_init.argtypes = [C.POINTER(OrcPF), _ip, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int,
                                  C.c_int, C.c_double, C.POINTER(C.c_ubyte)]
        L.orc_pf_energy_gradient.argtypes = [C.POINTER(OrcPF), _dp, _dp, _dp]
        L.orc_pf_constraints.argtypes = [C.POINTER(OrcPF), _dp]
        L.orc_pf_potential.argtypes = [C.POINTER(OrcPF), C.POINTER(OrcPot)]
        _ub = C.POINTER(C.c_ubyte)
        L.orc_pfx_init.argtypes = [C.POINTER(OrcPFX), _ip, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int,
                                   C.c_int, C.c_double, _ub, _ub, _dp, C.c_int, _dp, _dp, _dp, C.c_int,
                                   C.c_double, _dp, _dp]
        L.orc_pfx_free.argtypes = [C.POINTER(OrcPFX)]
        L.orc_pfx_energy_gradient.argtypes = [C.POINTER(OrcPFX), _dp, _dp, _dp]
        L.orc_pfx_constraints.argtypes = [C.POINTER(OrcPFX), _dp]
        L.orc_pfx_components.argtypes = [C.POINTER(OrcPFX), _dp, _dp]
        L.orc_pfx_potential.argtypes = [C.POINTER(OrcPFX), C.POINTER(OrcPot)]
        L.orc_pf_site_types.argtypes = [_ip, _ub, _ip]
        L.orc_pfu_init.argtypes = [C.POINTER(OrcPFU), _ip, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int,
                                   C.c_int, C.c_double, _ub, _ub, _dp, C.c_int, _dp, _dp, _dp, C.c_int,
                                   C.c_double, _dp, _dp]
        L.orc_pfu_free.argtypes = [C.POINTER(OrcPFU)]
        L.orc_pfu_energy_gradient.argtypes = [C.POINTER(OrcPFU), _dp, _dp, _dp]
        L.orc_pfu_constraints.argtypes = [C.POINTER(OrcPFU), _dp]
        L.orc_pfu_potential.argtypes = [C.POINTER(OrcPFU), C.POINTER(OrcPot)]
        L.orc_dot.argtypes = [_dp, _dp, C.c_size_t]
        L.orc_dot.restype = C.c_double
        L.orc_lbfgs_minimise.argtypes = [C.POINTER(OrcPot), _dp, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_int, C.c_double, C.POINTER(OrcStats)]
        L.orc_fire_minimise.argtypes = [C.POINTER(OrcPot), _dp, C.c_double, C.c_int, C.c_int, C.c_double,
                                        C.POINTER(OrcStats), _dp]
        L.orc_graddescent_minimise.argtypes = [C.POINTER(OrcPot), _dp, C.c_double, C.c_int, C.c_int,
                                               C.c_double, C.POINTER(OrcStats)]
        L.orc_lj_energy_gradient.argtypes = [C.POINTER(OrcLJ), _dp, _dp, _dp]
        L.orc_lj_potential.argtypes = [C.POINTER(OrcLJ), C.c_size_t, C.POINTER(OrcPot)]
        L.orc_ljpbc_energy_gradient.argtypes = [C.POINTER(OrcLJPBC), _dp, _dp, _dp]
        L.orc_ljpbc_potential.argtypes = [C.POINTER(OrcLJPBC), C.POINTER(OrcPot)]
        L.orc_fire_minimise_ex.argtypes = [C.POINTER(OrcPot), _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                           C.c_int, C.c_double, C.POINTER(OrcStats), _dp]
        L.orc_bah_energy_gradient.argtypes = [C.POINTER(OrcBaH), _dp, _dp, _dp]
        L.orc_bah_potential.argtypes = [C.POINTER(OrcBaH), C.POINTER(OrcPot)]
        _lib = L
    return _lib


# ---------------------------------------------------------------- PhaseField
class OraclePF:
    """CPU restatement of PhaseField (src/potentials/PhaseField.cpp)."""

    def __init__(self, grid, nfluid=1, resolution=1.0, surface_tension=(1.0,), interface_size=(),
                 density="hard", density_const=1.0, fix_fluid=None):
        L = lib()
        self.pf = OrcPF()
        g = (C.c_int * 3)(*grid)
        st = np.ascontiguousarray(surface_tension, dtype=np.float64)
        iw = np.ascontiguousarray(interface_size if len(interface_size) else [0.0], dtype=np.float64)
        fix = None
        if fix_fluid is not None:
            fix = (C.c_ubyte * nfluid)(*[1 if f else 0 for f in fix_fluid])
        rc = L.orc_pf_init(C.byref(self.pf), g, nfluid, resolution, _d(st), len(st), _d(iw),
                           len(interface_size), DENSITY[density], density_const, fix)
        if rc != 0:
            raise ValueError("PhaseField: invalid parameters (code %d)" % rc)
        self.ndof = int(np.prod(grid)) * nfluid
        self.pot = OrcPot()
        L.orc_pf_potential(C.byref(self.pf), C.byref(self.pot))

    @property
    def kappa(self):
        return np.array(self.pf.kappa[: self.pf.nfluid])

    @property
    def kappaP(self):
        return np.array(self.pf.kappaP[: self.pf.nfluid])

    @property
    def convergence(self):
        return self.pf.convergence

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        e = C.c_double(0)
        lib().orc_pf_energy_gradient(C.byref(self.pf), _d(x), C.byref(e), None)
        return e.value

    def gradient(self, x, constrained=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        lib().orc_pf_energy_gradient(C.byref(self.pf), _d(x), None, _d(g))
        if constrained:
            lib().orc_pf_constraints(C.byref(self.pf), _d(g))
        return g

    def constrain(self, data):
        data = np.array(data, dtype=np.float64)
        lib().orc_pf_constraints(C.byref(self.pf), _d(data))
        return data


COMPONENTS = ("fluid", "surface", "pressure", "density constraint", "force", "confinement", "volume constraint")


def expand_force(nfluid, force3, fluids=()):
    """PhaseField::setForce (PhaseField.cpp:876-887): one force for every fluid
    (one fluid, or no list), else only for the listed fluids."""
    f = np.zeros((nfluid, 3))
    if nfluid == 1 or len(fluids) == 0:
        f[:] = force3
    else:
        for i in fluids:
            f[i] = force3
    return f


def _opt(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


class OraclePFX:
    """CPU restatement of PhaseField with solid nodes, wetting, pressure, force,
    confinement, per-site surface tensions and the volume projection."""

    def __init__(self, grid, x0, nfluid=1, resolution=1.0, surface_tension=(1.0,), interface_size=(),
                 density="hard", density_const=1.0, fix_fluid=None, solid=None, contact_angle=None,
                 pressure=None, force=None, force_fluids=(), confinement=None, volume_fixed=False,
                 vol_const=0.01, volume=None):
        L = lib()
        self.p = OrcPFX()
        self.nfluid = nfluid
        if volume is not None:
            volume_fixed = True  # setVolume fixes the volumes (PhaseField.cpp:823-828)
        self._keep = dict(st=np.ascontiguousarray(surface_tension, dtype=np.float64),
                          iw=np.ascontiguousarray(interface_size if len(interface_size) else [0.0], dtype=np.float64),
                          x0=np.ascontiguousarray(x0, dtype=np.float64),
                          solid=_opt(solid, np.uint8), ca=_opt(contact_angle, np.float64),
                          pres=_opt(pressure, np.float64),
                          force=None if force is None else np.ascontiguousarray(
                              expand_force(nfluid, force, force_fluids)).reshape(-1),
                          conf=_opt(confinement, np.float64), vol=_opt(volume, np.float64))
        k = self._keep
        fix = None
        if fix_fluid is not None:
            fix = (C.c_ubyte * nfluid)(*[1 if f else 0 for f in fix_fluid])

        def dp(a):
            return None if a is None else _d(a)

        ub = None if k["solid"] is None else k["solid"].ctypes.data_as(C.POINTER(C.c_ubyte))
        rc = L.orc_pfx_init(C.byref(self.p), (C.c_int * 3)(*grid), nfluid, resolution, _d(k["st"]), len(k["st"]),
                            _d(k["iw"]), len(interface_size), DENSITY[density], density_const, fix, ub,
                            dp(k["ca"]), 0 if k["ca"] is None else k["ca"].size, dp(k["pres"]), dp(k["force"]),
                            dp(k["conf"]), 1 if volume_fixed else 0, vol_const, dp(k["vol"]), _d(k["x0"]))
        if rc != 0:
            raise ValueError("PhaseField: invalid parameters (code %d)" % rc)
        self.ndof = int(np.prod(grid)) * nfluid
        self.ngrid = int(np.prod(grid))
        self.pot = OrcPot()
        L.orc_pfx_potential(C.byref(self.p), C.byref(self.pot))

    def __del__(self):
        try:
            lib().orc_pfx_free(C.byref(self.p))
        except Exception:
            pass

    @property
    def convergence(self):
        return self.p.pf.convergence

    @property
    def node_vol(self):
        return np.ctypeslib.as_array(self.p.node_vol, shape=(self.ngrid,)).copy()

    @property
    def surface_area(self):
        return np.ctypeslib.as_array(self.p.surface_area, shape=(self.ngrid,)).copy()

    @property
    def site_type(self):
        return np.ctypeslib.as_array(self.p.type, shape=(self.ngrid,)).copy()

    @property
    def volume(self):
        return np.array(self.p.volume[: self.nfluid])

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        e = C.c_double(0)
        lib().orc_pfx_energy_gradient(C.byref(self.p), _d(x), C.byref(e), None)
        return e.value

    def energy_gradient(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        e = C.c_double(0)
        lib().orc_pfx_energy_gradient(C.byref(self.p), _d(x), C.byref(e), _d(g))
        return e.value, g

    def gradient(self, x, constrained=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        lib().orc_pfx_energy_gradient(C.byref(self.p), _d(x), None, _d(g))
        if constrained:
            lib().orc_pfx_constraints(C.byref(self.p), _d(g))
        return g

    def constrain(self, data):
        data = np.array(data, dtype=np.float64)
        lib().orc_pfx_constraints(C.byref(self.p), _d(data))
        return data

    def components(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(14)
        lib().orc_pfx_components(C.byref(self.p), _d(x), _d(out))
        return {n: [out[2 * i], out[2 * i + 1]] for i, n in enumerate(COMPONENTS)}


class OraclePFU:
    """CPU restatement of PhaseFieldUnstructured (src/potentials/PhaseFieldUnstructured.cpp)."""

    def __init__(self, grid, x0, nfluid=1, resolution=1.0, surface_tension=(1.0,), interface_size=(),
                 density="hard", density_const=1.0, fix_fluid=None, solid=None, contact_angle=None,
                 pressure=None, force=None, force_fluids=(), confinement=None, volume_fixed=False,
                 vol_const=0.01, volume=None):
        L = lib()
        self.p = OrcPFU()
        self.nfluid = nfluid
        self._keep = dict(st=np.ascontiguousarray(surface_tension, dtype=np.float64),
                          iw=np.ascontiguousarray(interface_size if len(interface_size) else [0.0], dtype=np.float64),
                          x0=np.ascontiguousarray(x0, dtype=np.float64),
                          solid=_opt(solid, np.uint8), ca=_opt(contact_angle, np.float64),
                          pres=_opt(pressure, np.float64),
                          force=None if force is None else np.ascontiguousarray(
                              expand_force(nfluid, force, force_fluids)).reshape(-1),
                          conf=_opt(confinement, np.float64), vol=_opt(volume, np.float64))
        k = self._keep
        fix = None
        if fix_fluid is not None:
            fix = (C.c_ubyte * nfluid)(*[1 if f else 0 for f in fix_fluid])

        def dp(a):
            return None if a is None else _d(a)

        ub = None if k["solid"] is None else k["solid"].ctypes.data_as(C.POINTER(C.c_ubyte))
        rc = L.orc_pfu_init(C.byref(self.p), (C.c_int * 3)(*grid), nfluid, resolution, _d(k["st"]), len(k["st"]),
                            _d(k["iw"]), len(interface_size), DENSITY[density], density_const, fix, ub,
                            dp(k["ca"]), 0 if k["ca"] is None else k["ca"].size, dp(k["pres"]), dp(k["force"]),
                            dp(k["conf"]), 1 if volume_fixed else 0, vol_const, dp(k["vol"]), _d(k["x0"]))
        if rc != 0:
            raise ValueError("PhaseFieldUnstructured: invalid parameters (code %d)" % rc)
        self.ndof = int(np.prod(grid)) * nfluid
        self.pot = OrcPot()
        L.orc_pfu_potential(C.byref(self.p), C.byref(self.pot))

    def __del__(self):
        try:
            lib().orc_pfu_free(C.byref(self.p))
        except Exception:
            pass

    @property
    def convergence(self):
        return self.p.convergence

    @property
    def nelements(self):
        return self.p.nel

    @property
    def kappa(self):
        return np.array(self.p.kappa[: self.nfluid])

    @property
    def node_vol(self):
        return np.ctypeslib.as_array(self.p.node_vol, shape=(self.p.ngrid,)).copy()

    @property
    def volume(self):
        return np.array(self.p.volume[: self.nfluid])

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        e = C.c_double(0)
        lib().orc_pfu_energy_gradient(C.byref(self.p), _d(x), C.byref(e), None)
        return e.value

    def energy_gradient(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        e = C.c_double(0)
        lib().orc_pfu_energy_gradient(C.byref(self.p), _d(x), C.byref(e), _d(g))
        return e.value, g

    def gradient(self, x, constrained=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        lib().orc_pfu_energy_gradient(C.byref(self.p), _d(x), None, _d(g))
        if constrained:
            lib().orc_pfu_constraints(C.byref(self.p), _d(g))
        return g

    def constrain(self, data):
        data = np.array(data, dtype=np.float64)
        lib().orc_pfu_constraints(C.byref(self.p), _d(data))
        return data


def oracle_site_types(grid, solid):
    s = np.ascontiguousarray(solid, dtype=np.uint8)
    t = np.zeros(s.size, np.int32)
    rc = lib().orc_pf_site_types((C.c_int * 3)(*grid), s.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                 t.ctypes.data_as(_ip))
    if rc != 0:
        raise RuntimeError("PhaseField: Undefined surface type")
    return t


def _stats(st):
    return dict(iterations=st.iterations, converged=bool(st.converged), n_energy=st.n_energy,
                n_gradient=st.n_gradient, energy=st.energy, grad_rms=st.grad_rms)


def oracle_lbfgs(pot_holder, x0, m=5, max_iter=100000, max_step=0.0, init_step=1.0,
                 linesearch="backtracking", convergence=None):
    x = np.array(x0, dtype=np.float64)
    st = OrcStats()
    conv = pot_holder.pot.convergence if convergence is None else convergence
    lib().orc_lbfgs_minimise(C.byref(pot_holder.pot), _d(x), m, max_iter, max_step, init_step,
                             0 if linesearch == "backtracking" else 1, conv, C.byref(st))
    return x, _stats(st)


def oracle_fire(pot_holder, x0, dt_max=0.0, max_iter=100000, linesearch="backtracking", convergence=None,
                dt0=0.0, max_force_dim=0):
    x = np.array(x0, dtype=np.float64)
    st = OrcStats()
    dto = C.c_double(0)
    conv = pot_holder.pot.convergence if convergence is None else convergence
    if dt0 == 0.0 and max_force_dim == 0:
        lib().orc_fire_minimise(C.byref(pot_holder.pot), _d(x), dt_max, max_iter,
                                0 if linesearch == "backtracking" else 1, conv, C.byref(st), C.byref(dto))
    else:
        lib().orc_fire_minimise_ex(C.byref(pot_holder.pot), _d(x), dt_max, dt0, max_force_dim, max_iter,
                                   0 if linesearch == "backtracking" else 1, conv, C.byref(st), C.byref(dto))
    out = _stats(st)
    out["dt_max"] = dto.value
    return x, out


def oracle_graddescent(pot_holder, x0, alpha=0.1, max_iter=100000, linesearch="backtracking", convergence=None):
    x = np.array(x0, dtype=np.float64)
    st = OrcStats()
    conv = pot_holder.pot.convergence if convergence is None else convergence
    lib().orc_graddescent_minimise(C.byref(pot_holder.pot), _d(x), alpha, max_iter,
                                   0 if linesearch == "backtracking" else 1, conv, C.byref(st))
    return x, _stats(st)


def oracle_dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_dot(_d(a), _d(b), a.size)


# ---------------------------------------------------------------- Lennard-Jones
class OracleLJ:
    def __init__(self, ndim, nparticle, sigma=1.0, epsilon=1.0):
        self.lj = OrcLJ(ndim, nparticle, sigma, epsilon)
        self.ndof = ndim * nparticle
        self.pot = OrcPot()
        lib().orc_lj_potential(C.byref(self.lj), self.ndof, C.byref(self.pot))

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        e = C.c_double(0)
        lib().orc_lj_energy_gradient(C.byref(self.lj), _d(x), C.byref(e), None)
        return e.value

    def gradient(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        lib().orc_lj_energy_gradient(C.byref(self.lj), _d(x), None, _d(g))
        return g


class OracleLJPBC:
    def __init__(self, nparticle, box, rc=2.5, sigma=1.0, epsilon=1.0):
        self.lj = OrcLJPBC(nparticle, sigma, epsilon, rc, box)
        self.ndof = 3 * nparticle
        self.pot = OrcPot()
        lib().orc_ljpbc_potential(C.byref(self.lj), C.byref(self.pot))

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        e = C.c_double(0)
        lib().orc_ljpbc_energy_gradient(C.byref(self.lj), _d(x), C.byref(e), None)
        return e.value

    def gradient(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        lib().orc_ljpbc_energy_gradient(C.byref(self.lj), _d(x), None, _d(g))
        return g


# ---------------------------------------------------------------- BarAndHinge
class OracleBaH:
    def __init__(self, nnode, bars, bar_k, bar_l0, hinges, hinge_k, hinge_t0, force=None,
                 force_per_node=False, fixed=None):
        self._keep = []

        def keep(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a
        bars = keep(np.asarray(bars).reshape(-1), np.int32)
        hinges = keep(np.asarray(hinges).reshape(-1), np.int32)
        self.b = OrcBaH()
        self.b.nnode = nnode
        self.b.nbar = bars.size // 2
        self.b.bar = bars.ctypes.data_as(_ip)
        self.b.bar_k = _d(keep(bar_k, np.float64))
        self.b.bar_l0 = _d(keep(bar_l0, np.float64))
        self.b.nhinge = hinges.size // 4
        self.b.hinge = hinges.ctypes.data_as(_ip)
        self.b.hinge_k = _d(keep(hinge_k, np.float64))
        self.b.hinge_t0 = _d(keep(hinge_t0, np.float64))
        if force is not None:
            self.b.force = _d(keep(force, np.float64))
            self.b.force_per_node = 1 if force_per_node else 0
        if fixed is not None:
            fx = keep(np.asarray(fixed, dtype=np.uint8), np.uint8)
            self.b.fixed = fx.ctypes.data_as(C.POINTER(C.c_ubyte))
        self.ndof = 3 * nnode
        self.pot = OrcPot()
        lib().orc_bah_potential(C.byref(self.b), C.byref(self.pot))

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        e = C.c_double(0)
        lib().orc_bah_energy_gradient(C.byref(self.b), _d(x), C.byref(e), None)
        return e.value

    def gradient(self, x, constrained=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        lib().orc_bah_energy_gradient(C.byref(self.b), _d(x), None, _d(g))
        if constrained and self.b.fixed:
            fx = np.ctypeslib.as_array(self.b.fixed, shape=(self.ndof,))
            g[fx.astype(bool)] = 0
        return g


# ---------------------------------------------------------------- reference (_ref)
class RefPFExtOpts(C.Structure):
    _fields_ = [("solid", C.POINTER(C.c_ubyte)), ("contact_angle", _dp), ("n_ca", C.c_int), ("pressure", _dp),
                ("force3", _dp), ("force_fluids", _ip), ("n_force_fluids", C.c_int), ("confinement", _dp),
                ("volume_fixed", C.c_int), ("vol_const", C.c_double), ("volume", _dp),
                ("fix_fluid", C.POINTER(C.c_ubyte)), ("density_const", C.c_double)]


def _ref_opts(nfluid, keep, solid=None, contact_angle=None, pressure=None, force=None, force_fluids=(),
              confinement=None, volume_fixed=False, vol_const=0.01, volume=None, fix_fluid=None,
              density_const=1.0):
    o = RefPFExtOpts()
    if solid is not None:
        keep["solid"] = np.ascontiguousarray(solid, dtype=np.uint8)
        o.solid = keep["solid"].ctypes.data_as(C.POINTER(C.c_ubyte))
    if contact_angle is not None:
        keep["ca"] = np.ascontiguousarray(contact_angle, dtype=np.float64)
        o.contact_angle = _d(keep["ca"])
        o.n_ca = keep["ca"].size
    if pressure is not None:
        keep["p"] = np.ascontiguousarray(pressure, dtype=np.float64)
        o.pressure = _d(keep["p"])
    if force is not None:
        keep["f"] = np.ascontiguousarray(force, dtype=np.float64)
        keep["ff"] = np.ascontiguousarray(list(force_fluids) or [0], dtype=np.int32)
        o.force3 = _d(keep["f"])
        o.force_fluids = keep["ff"].ctypes.data_as(_ip)
        o.n_force_fluids = len(force_fluids)
    if confinement is not None:
        keep["c"] = np.ascontiguousarray(confinement, dtype=np.float64)
        o.confinement = _d(keep["c"])
    o.volume_fixed = 1 if volume_fixed else 0
    o.vol_const = vol_const
    if volume is not None:
        keep["v"] = np.ascontiguousarray(volume, dtype=np.float64)
        o.volume = _d(keep["v"])
    if fix_fluid is not None:
        keep["fix"] = np.ascontiguousarray([1 if f else 0 for f in fix_fluid], dtype=np.uint8)
        o.fix_fluid = keep["fix"].ctypes.data_as(C.POINTER(C.c_ubyte))
    o.density_const = density_const
    return o


class RefResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("n_energy", C.c_long),
                ("n_gradient", C.c_long), ("energy", C.c_double), ("grad_rms", C.c_double)]


_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def reflib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.ref_pf_state_create.restype = C.c_void_p
        L.ref_pf_state_create.argtypes = [_ip, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp]
        L.ref_state_eval.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.ref_pf_params.argtypes = [_ip, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, _dp]
        L.ref_state_minimise.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_int, C.c_double, _dp, C.POINTER(RefResult)]
        L.ref_lj_state_create.restype = C.c_void_p
        L.ref_lj_state_create.argtypes = [C.c_int, C.c_double, C.c_double, _dp, C.c_int]
        L.ref_bah_state_create.restype = C.c_void_p
        L.ref_bah_state_create.argtypes = [_ip, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                           C.c_double, _ip, C.c_int, _dp, _dp, C.c_int]
        L.ref_bah_elements.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp, _ip, _ip, _dp, _dp, C.c_int]
        L.ref_bah_set_hinge_theta0.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.ref_pf_state_create_ex.restype = C.c_void_p
        L.ref_pf_state_create_ex.argtypes = [_ip, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp,
                                             C.POINTER(RefPFExtOpts)]
        L.ref_pf_components.argtypes = [C.c_void_p, _dp]
        L.ref_pf_geometry.argtypes = [C.c_void_p, _dp, _dp]
        L.ref_state_constraints.argtypes = [C.c_void_p, _dp]
        L.ref_pf_diffuse_solid.argtypes = [_ip, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int, C.c_int,
                                           C.POINTER(RefPFExtOpts), C.POINTER(C.c_ubyte), C.c_int, C.c_int, _dp]
        L.ref_pfu_state_create_ex.restype = C.c_void_p
        L.ref_pfu_state_create_ex.argtypes = [_ip, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp,
                                              C.POINTER(RefPFExtOpts)]
        L.ref_state_nelements.argtypes = [C.c_void_p]
        L.ref_state_nelements.restype = C.c_long
        L.ref_state_destroy.argtypes = [C.c_void_p]
        L.ref_state_ndof.argtypes = [C.c_void_p]
        L.ref_state_ndof.restype = C.c_long
        _ref = L
    return _ref


class RefState:
    """A reference State (sjavis/minim compiled from source) behind a handle."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("reference State construction failed")
        self.h = handle
        self.ndof = reflib().ref_state_ndof(self.h)

    def __del__(self):
        try:
            reflib().ref_state_destroy(self.h)
        except Exception:
            pass

    @staticmethod
    def phasefield(grid, x, nfluid=1, resolution=1.0, surface_tension=(1.0,), interface_size=(), density="hard"):
        st = np.ascontiguousarray(surface_tension, dtype=np.float64)
        iw = np.ascontiguousarray(interface_size if len(interface_size) else [0.0], dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        h = reflib().ref_pf_state_create((C.c_int * 3)(*grid), nfluid, resolution, _d(st), len(st), _d(iw),
                                         len(interface_size), DENSITY[density], _d(x))
        return RefState(h)

    @staticmethod
    def phasefield_ext(grid, x, nfluid=1, resolution=1.0, surface_tension=(1.0,), interface_size=(),
                       density="hard", **opts):
        keep = {}
        o = _ref_opts(nfluid, keep, **opts)
        st = np.ascontiguousarray(surface_tension, dtype=np.float64)
        iw = np.ascontiguousarray(interface_size if len(interface_size) else [0.0], dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        h = reflib().ref_pf_state_create_ex((C.c_int * 3)(*grid), nfluid, resolution, _d(st), len(st), _d(iw),
                                            len(interface_size), DENSITY[density], _d(x), C.byref(o))
        r = RefState(h)
        r.ngrid = int(np.prod(grid))
        return r

    @staticmethod
    def phasefield_unstructured(grid, x, nfluid=1, resolution=1.0, surface_tension=(1.0,), interface_size=(),
                                density="hard", **opts):
        keep = {}
        o = _ref_opts(nfluid, keep, **opts)
        st = np.ascontiguousarray(surface_tension, dtype=np.float64)
        iw = np.ascontiguousarray(interface_size if len(interface_size) else [0.0], dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        h = reflib().ref_pfu_state_create_ex((C.c_int * 3)(*grid), nfluid, resolution, _d(st), len(st), _d(iw),
                                             len(interface_size), DENSITY[density], _d(x), C.byref(o))
        r = RefState(h)
        r.ngrid = int(np.prod(grid))
        return r

    def nelements(self):
        return reflib().ref_state_nelements(self.h)

    def components(self):
        out = np.zeros(14)
        assert reflib().ref_pf_components(self.h, _d(out)) == 0
        return {n: [out[2 * i], out[2 * i + 1]] for i, n in enumerate(COMPONENTS)}

    def geometry(self):
        nv, sa = np.zeros(self.ngrid), np.zeros(self.ngrid)
        reflib().ref_pf_geometry(self.h, _d(nv), _d(sa))
        return nv, sa

    def constrain(self, data):
        d = np.array(data, dtype=np.float64)
        assert reflib().ref_state_constraints(self.h, _d(d)) == 0
        return d

    @staticmethod
    def lj(ndim, x, sigma=1.0, epsilon=1.0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return RefState(reflib().ref_lj_state_create(ndim, sigma, epsilon, _d(x), x.size))

    @staticmethod
    def bah(tri, x, kb, kh, l0=None, t0=None, fixed_dofs=(), force3=None):
        tri = np.ascontiguousarray(np.asarray(tri).reshape(-1), dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        fx = np.ascontiguousarray(fixed_dofs if len(fixed_dofs) else [0], dtype=np.int32)
        f3 = None if force3 is None else _d(np.ascontiguousarray(force3, dtype=np.float64))
        h = reflib().ref_bah_state_create(tri.ctypes.data_as(_ip), tri.size // 3, kb, kh,
                                          0 if l0 is None else 1, 0.0 if l0 is None else l0,
                                          0 if t0 is None else 1, 0.0 if t0 is None else t0,
                                          fx.ctypes.data_as(_ip), len(fixed_dofs), f3, _d(x), x.size)
        return RefState(h)

    def bah_elements(self):
        cap = max(16, 4 * self.ndof)
        nb, nh = C.c_int(0), C.c_int(0)
        bars = np.zeros(2 * cap, np.int32)
        hinges = np.zeros(4 * cap, np.int32)
        bk, bl, hk, ht = (np.zeros(cap) for _ in range(4))
        rc = reflib().ref_bah_elements(self.h, C.byref(nb), bars.ctypes.data_as(_ip), _d(bk), _d(bl), C.byref(nh),
                                       hinges.ctypes.data_as(_ip), _d(hk), _d(ht), cap)
        assert rc == 0
        nb, nh = nb.value, nh.value
        return dict(bars=bars[: 2 * nb].reshape(-1, 2), bar_k=bk[:nb], bar_l0=bl[:nb],
                    hinges=hinges[: 4 * nh].reshape(-1, 4), hinge_k=hk[:nh], hinge_t0=ht[:nh])

    def set_hinge_theta0(self, ih, t0):
        assert reflib().ref_bah_set_hinge_theta0(self.h, ih, t0) == 0

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        e = C.c_double(0)
        assert reflib().ref_state_eval(self.h, _d(x), C.byref(e), None) == 0
        return e.value

    def gradient(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        assert reflib().ref_state_eval(self.h, _d(x), None, _d(g)) == 0
        return g

    def energy_gradient(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros_like(x)
        e = C.c_double(0)
        assert reflib().ref_state_eval(self.h, _d(x), C.byref(e), _d(g)) == 0
        return e.value, g

    def minimise(self, kind="lbfgs", m=5, max_iter=100000, param_a=0.0, param_b=1.0,
                 linesearch="backtracking", convergence=0.0):
        x = np.zeros(self.ndof)
        r = RefResult()
        k = {"lbfgs": 0, "fire": 1, "graddescent": 2}[kind]
        rc = reflib().ref_state_minimise(self.h, k, m, max_iter, param_a, param_b,
                                         0 if linesearch == "backtracking" else 1, convergence, _d(x), C.byref(r))
        assert rc == 0
        return x, dict(iterations=r.iterations, converged=bool(r.converged), n_energy=r.n_energy,
                       n_gradient=r.n_gradient, energy=r.energy, grad_rms=r.grad_rms)


def ref_diffuse_solid(grid, solid_in, nfluid=2, ifluid=0, two_step=False, resolution=1.0, surface_tension=(1.0,),
                      interface_size=(), density="hard", **opts):
    """PhaseField::diffuseSolid of a potential configured like phasefield_ext."""
    keep = {}
    o = _ref_opts(nfluid, keep, **opts)
    st = np.ascontiguousarray(surface_tension, dtype=np.float64)
    iw = np.ascontiguousarray(interface_size if len(interface_size) else [0.0], dtype=np.float64)
    s = np.ascontiguousarray(solid_in, dtype=np.uint8)
    out = np.zeros(int(np.prod(grid)) * nfluid)
    rc = reflib().ref_pf_diffuse_solid((C.c_int * 3)(*grid), nfluid, resolution, _d(st), len(st), _d(iw),
                                       len(interface_size), DENSITY[density], C.byref(o),
                                       s.ctypes.data_as(C.POINTER(C.c_ubyte)), ifluid, 1 if two_step else 0,
                                       _d(out))
    assert rc == 0
    return out
