"""Python mirror of minim's State / Potential / Minimiser API over the C ABI.

Names, setters, defaults and error behaviour follow the reference's C++ API
(include/State.h, include/Potential.h, include/potentials/*.h,
include/minimisers/*.h); every computation runs in libminim_amd.so on the GPU.
"""
import ctypes as C
import math

import numpy as np

from . import _capi as A

_DENSITY = {"gradient": 0, "hard": 0, "energy penalty": 1, "soft": 1, "none": 2, "fixed": 3}
_PREC = {"fp64": A.MINIM_FP64, "fp32": A.MINIM_FP32, A.MINIM_FP64: A.MINIM_FP64, A.MINIM_FP32: A.MINIM_FP32}


def _dptr(a):
    return a.ctypes.data_as(A.dp)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def _out(out, n):
    if out is None:
        return np.zeros(n)
    if not (isinstance(out, np.ndarray) and out.dtype == np.float64 and out.flags.c_contiguous and out.size == n):
        raise ValueError("out must be a contiguous float64 array of %d entries" % n)
    return out


class _Handle:
    """Owns one C handle and frees it with the given destructor."""

    def __init__(self, ptr, destroy):
        self.ptr = ptr
        self._destroy = destroy

    def __del__(self):
        if self.ptr:
            try:
                self._destroy(self.ptr)
            except Exception:
                pass
            self.ptr = None


# ----------------------------------------------------------------- potentials
class Potential:
    """Base of the device potentials (include/Potential.h:16-108)."""

    def __init__(self):
        self.fixedDofs = []

    def setConstraints(self, iFix):
        """Potential::setConstraints(vector<int> iFix) (src/Potential.cpp:80-85)."""
        self.fixedDofs = list(self.fixedDofs) + [int(i) for i in iFix]
        return self

    def newState(self, coords, precision="fp64"):
        return State(self, coords, precision)

    def _create(self):
        raise NotImplementedError

    def handle(self, coords):
        """A fresh, initialised C potential for the given coordinates."""
        L = A.load()
        h = self._create()
        if self.fixedDofs:
            idx = np.ascontiguousarray(self.fixedDofs, dtype=np.int32)
            A.check(L.minim_potential_set_fixed(h.ptr, idx.ctypes.data_as(A.ip), idx.size))
        x = _f64(coords)
        A.check(L.minim_potential_init(h.ptr, _dptr(x), x.size))
        return h


class PhaseField(Potential):
    """include/potentials/PhaseField.h:13-123: bulk, gradient and density terms,
    fixed fluids, solid nodes and wetting, pressure, fixed volumes, external
    forces, confinement and per-site surface tensions."""

    potentialType = 2

    def __init__(self):
        super().__init__()
        self.nFluid = 1
        self.resolution = 1.0
        self.gridSize = None
        self.interfaceSize = []
        self.surfaceTension = [1.0]
        self.densityConstraint = 0
        self.densityConst = 1.0
        self.fixFluid = []
        self.stencilXChunk = 0
        self.stencilTargetBlocks = 0
        self.stencilKernel = 0
        self.pressure = []
        self.volumeFixed = False
        self.volConst = 0.01
        self.volume = []
        self.solid = []
        self.contactAngle = []
        self.force = None
        self.forceFluids = []
        self.confinementStrength = []

    def setNFluid(self, n):
        self.nFluid = int(n)
        return self

    def setGridSize(self, g):
        self.gridSize = [int(v) for v in g]
        return self

    def setResolution(self, r):
        self.resolution = float(r)
        return self

    def setInterfaceSize(self, w):
        # the scalar form sizes the vector by the current nFluid (PhaseField.cpp:764-767)
        self.interfaceSize = [float(w)] * self.nFluid if np.isscalar(w) else [float(v) for v in w]
        return self

    def setSurfaceTension(self, s):
        """A value, a vector (nParams, or nParams per site), or fn(x, y, z) ->
        the site's nParams values (PhaseField.cpp:769-797)."""
        if callable(s):
            self.surfaceTension = [float(v) for site in self._sites() for v in s(*site)]
        else:
            self.surfaceTension = [float(s)] * self.nFluid if np.isscalar(s) else [float(v) for v in s]
        return self

    def _sites(self):
        if self.gridSize is None:
            raise ValueError("PhaseField: set the grid size first.")
        nx, ny, nz = self.gridSize
        return [(x, y, z) for x in range(nx) for y in range(ny) for z in range(nz)]

    def setPressure(self, p):
        self.pressure = [float(v) for v in p]
        return self

    def setVolume(self, volume, volConst=0.01):
        self.volumeFixed = True
        self.volConst = float(volConst)
        self.volume = [float(v) for v in volume]
        return self

    def setVolumeFixed(self, volumeFixed, volConst=0.01):
        self.volumeFixed = bool(volumeFixed)
        self.volConst = float(volConst)
        return self

    def setSolid(self, solid):
        """Flags per site, or fn(x, y, z) -> bool (PhaseField.cpp:836-852)."""
        if callable(solid):
            self.solid = [1 if solid(*site) else 0 for site in self._sites()]
        else:
            self.solid = [1 if v else 0 for v in np.asarray(solid).reshape(-1)]
        return self

    def setContactAngle(self, a):
        """A value for every site, a vector, or fn(x, y, z) (PhaseField.cpp:854-873)."""
        if callable(a):
            self.contactAngle = [float(a(*site)) for site in self._sites()]
        elif np.isscalar(a):
            self.contactAngle = [float(a)] * int(np.prod(self.gridSize))
        else:
            self.contactAngle = [float(v) for v in a]
        return self

    def setForce(self, force, iFluid=()):
        if len(force) != 3:
            raise ValueError("PhaseField: Invalid size of force array.")
        self.force = [float(v) for v in force]
        self.forceFluids = [int(i) for i in iFluid]
        return self

    def setConfinement(self, strength):
        self.confinementStrength = [float(v) for v in strength]
        return self

    def setDensityConstraint(self, method):
        if method not in _DENSITY:
            raise ValueError("PhaseField: Invalid density constraint. Allowed methods are: fixed, none, "
                             "gradient / hard, or energy penalty / soft")
        self.densityConstraint = _DENSITY[method]
        return self

    def setFixFluid(self, iFluid, fix=True):
        if len(self.fixFluid) != self.nFluid:
            self.fixFluid = [False] * self.nFluid
        self.fixFluid[iFluid] = bool(fix)
        return self

    def _params(self):
        """The setters' values as minim_pf_params (arrays kept alive on self)."""
        if self.gridSize is None or len(self.gridSize) != 3:
            raise ValueError("PhaseField: gridSize must have three entries.")
        p = A.PFParams()
        for i in range(3):
            p.grid[i] = self.gridSize[i]
        p.n_fluid = self.nFluid
        p.resolution = self.resolution
        self._st = _f64(self.surfaceTension)
        p.surface_tension = _dptr(self._st)
        p.n_surface_tension = self._st.size
        self._iw = _f64(self.interfaceSize if self.interfaceSize else [0.0])
        p.interface_size = _dptr(self._iw)
        p.n_interface_size = len(self.interfaceSize)
        if p.n_interface_size == 1 and self.nFluid != 1:
            p.n_interface_size = 1
        p.density_constraint = self.densityConstraint
        p.density_const = self.densityConst
        if self.fixFluid:
            self._fx = (C.c_ubyte * self.nFluid)(*[1 if f else 0 for f in self.fixFluid])
            p.fix_fluid = self._fx
        p.stencil_xchunk = self.stencilXChunk
        p.stencil_target_blocks = self.stencilTargetBlocks
        p.stencil_kernel = self.stencilKernel
        if len(self.solid):
            self._solid = (C.c_ubyte * len(self.solid))(*self.solid)
            p.solid = self._solid
        if len(self.contactAngle):
            self._ca = _f64(self.contactAngle)
            p.contact_angle = _dptr(self._ca)
            p.n_contact_angle = self._ca.size
        if len(self.pressure):
            if len(self.pressure) != self.nFluid:
                raise ValueError("PhaseField: Invalid size of pressure array.")
            self._pres = _f64(self.pressure)
            p.pressure = _dptr(self._pres)
        if self.force is not None:
            self._force = _f64(self.force)
            p.force = _dptr(self._force)
            self._ff = np.ascontiguousarray(self.forceFluids or [0], dtype=np.int32)
            p.force_fluids = self._ff.ctypes.data_as(A.ip)
            p.n_force_fluids = len(self.forceFluids)
        if len(self.confinementStrength):
            if len(self.confinementStrength) != self.nFluid:
                raise ValueError("PhaseField: Invalid size of confinement strength array.")
            self._conf = _f64(self.confinementStrength)
            p.confinement = _dptr(self._conf)
        p.volume_fixed = 1 if self.volumeFixed else 0
        p.vol_const = self.volConst
        if len(self.volume):
            if len(self.volume) != self.nFluid:
                raise ValueError("PhaseField: Invalid size of volume array.")
            self._vol = _f64(self.volume)
            p.volume = _dptr(self._vol)
        return p

    def _create(self):
        p = self._params()
        h = C.c_void_p()
        A.check(A.load().minim_pf_create(C.byref(p), C.byref(h)))
        return _Handle(h.value, A.load().minim_potential_destroy)

    COMPONENTS = ("fluid", "surface", "pressure", "density constraint", "force", "confinement", "volume constraint")

    def autotune(self, coords, precision="fp32", blocks=(256, 512, 1024, 2048), marches=(0, 4, 8, 16, 32), reps=10):
        """Launch-shape autotuner of the evaluation kernel this field uses
        (PhaseField::autotune, minim_pf_autotune): one energy+gradient launch
        (HIP events, kernel only) is timed for every pair of target block count
        and x-march length (0 = automatic) in the library's own device buffers;
        the fastest is kept for this precision in the library's process-wide
        cache, which every later handle of the same shape and kernel finds at
        init.  Returns {(blocks, march): ms} (shapes the grid cannot take are
        left out).  The shape changes the tiling only: gradients are the same
        bit for bit, energies differ by the order of the block sums."""
        L = A.load()
        x = _f64(coords)
        h = self.handle(x)
        tv = np.ascontiguousarray(blocks, dtype=np.int32)
        mv = np.ascontiguousarray(marches, dtype=np.int32)
        times = np.empty(tv.size * mv.size)
        bt, bm = C.c_int(), C.c_int()
        A.check(L.minim_pf_autotune(h.ptr, _PREC[precision], _dptr(x), x.size, tv.ctypes.data_as(A.ip), tv.size,
                                    mv.ctypes.data_as(A.ip), mv.size, int(reps), _dptr(times), C.byref(bt),
                                    C.byref(bm)))
        out = {}
        for i, b in enumerate(tv.tolist()):
            for j, mch in enumerate(mv.tolist()):
                if times[i * mv.size + j] >= 0:
                    out[(b, mch)] = float(times[i * mv.size + j])
        return out

    def tunedShape(self, precision, coords=None, handle=None):
        """(target blocks, x-march) a handle of this field uses for a
        precision after autotuning; (0, 0) = the heuristic."""
        h = handle if handle is not None else self.handle(coords)
        t, mch = C.c_int(), C.c_int()
        A.check(A.load().minim_pf_tuned_shape(h.ptr, _PREC[precision], C.byref(t), C.byref(mch)))
        return t.value, mch.value

    def energyComponents(self, coords, init_coords=None):
        """PhaseField::energyComponents (PhaseField.cpp:706-744): {name: [energy,
        |gradient|]} at coords, for the potential initialised at init_coords
        (default: coords)."""
        x = _f64(coords)
        h = self.handle(x if init_coords is None else init_coords)
        out = np.zeros(14)
        A.check(A.load().minim_pf_energy_components(h.ptr, _dptr(x), x.size, _dptr(out)))
        return {n: [out[2 * i], out[2 * i + 1]] for i, n in enumerate(self.COMPONENTS)}

    def nodeVolumes(self, coords=None):
        """(nodeVol, surfaceArea) per site after init (PhaseField.cpp:190-214)."""
        ng = int(np.prod(self.gridSize))
        h = self.handle(np.zeros(ng * self.nFluid) if coords is None else coords)
        nv, sa = np.zeros(ng), np.zeros(ng)
        A.check(A.load().minim_pf_node_volumes(h.ptr, _dptr(nv), _dptr(sa)))
        return nv, sa

    def diffuseSolid(self, solid, iFluid=0, twoStep=False):
        """PhaseField::diffuseSolid (PhaseField.cpp:998-1035) on a copy of this
        potential."""
        s = np.ascontiguousarray(np.asarray(solid).reshape(-1), dtype=np.uint8)
        out = np.zeros(int(np.prod(self.gridSize)) * self.nFluid)
        h = self._create()
        A.check(A.load().minim_pf_diffuse_solid(h.ptr, s.ctypes.data_as(C.POINTER(C.c_ubyte)), int(iFluid),
                                                1 if twoStep else 0, _dptr(out)))
        return out


    def coefficients(self, coords=None):
        """(kappa, kappaP, convergence) as PhaseField::init computes them."""
        n = int(np.prod(self.gridSize)) * self.nFluid
        h = self.handle(np.zeros(n) if coords is None else coords)
        k = np.zeros(self.nFluid)
        kp = np.zeros(self.nFluid)
        A.check(A.load().minim_pf_coefficients(h.ptr, _dptr(k), _dptr(kp)))
        conv = C.c_double()
        A.check(A.load().minim_potential_convergence(h.ptr, C.byref(conv)))
        return k, kp, conv.value


class PhaseFieldUnstructured(PhaseField):
    """include/potentials/PhaseFieldUnstructured.h:13-104: the element form of
    the phase field (one element per site and fluid plus one-DOF pressure,
    surface, force and confinement elements), with the same setters as
    PhaseField except per-site surface tensions.  The fixed volume is an
    energy term (PhaseFieldUnstructured.cpp:271-297)."""

    potentialType = 1

    def setSurfaceTension(self, s):
        if callable(s):
            raise ValueError("PhaseFieldUnstructured: per-site surface tensions are not supported")
        return super().setSurfaceTension(s)

    def _create(self):
        p = self._params()
        h = C.c_void_p()
        A.check(A.load().minim_pfu_create(C.byref(p), C.byref(h)))
        return _Handle(h.value, A.load().minim_potential_destroy)

    def energyComponents(self, coords, init_coords=None):
        raise NotImplementedError("PhaseFieldUnstructured has no energyComponents")

    def nodeVolumes(self, coords=None):
        """nodeVol per site after init (PhaseFieldUnstructured.cpp:252-268)."""
        ng = int(np.prod(self.gridSize))
        h = self.handle(np.zeros(ng * self.nFluid) if coords is None else coords)
        nv = np.zeros(ng)
        A.check(A.load().minim_pf_node_volumes(h.ptr, _dptr(nv), None))
        return nv

    def nElements(self, coords=None):
        """Number of energy elements init builds."""
        return self.elementList(coords)[0]

    def elementList(self, coords=None, handle=None):
        """(number of elements, device bytes of the list) after init."""
        if handle is None:
            ng = int(np.prod(self.gridSize))
            handle = self.handle(np.zeros(ng * self.nFluid) if coords is None else coords)
        n, b = C.c_size_t(), C.c_size_t()
        A.check(A.load().minim_pfu_element_list(handle.ptr, C.byref(n), C.byref(b)))
        return n.value, b.value


def site_types(grid, solid):
    """PhaseField::getType of every site (host): -1 solid, 0 bulk, 1..7."""
    s = np.ascontiguousarray(np.asarray(solid).reshape(-1), dtype=np.uint8)
    g = (C.c_int * 3)(*[int(v) for v in grid])
    t = np.zeros(s.size, np.int32)
    A.check(A.load().minim_pf_site_types(g, s.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                         t.ctypes.data_as(C.POINTER(C.c_int))))
    return t


class LjNd(Potential):
    """include/potentials/LjNd.h:12-44; setCutoff adds a truncated-shifted periodic mode."""

    potentialType = 1

    def __init__(self, nDim, sigma=1.0, epsilon=1.0):
        super().__init__()
        self.nDim = int(nDim)
        self.sigma = float(sigma)
        self.epsilon = float(epsilon)
        self.cutoff = 0.0
        self.box = None
        self.skin = 0.3

    def setSigma(self, s):
        self.sigma = float(s)
        return self

    def setEpsilon(self, e):
        self.epsilon = float(e)
        return self

    def setCutoff(self, rc, box, skin=0.3):
        self.cutoff = float(rc)
        self.box = [float(b) for b in (box if np.ndim(box) else [box] * 3)]
        self.skin = float(skin)
        return self

    def listBuilds(self, state):
        """LjNd::listBuilds: neighbour-list builds so far of the potential
        behind state (cutoff mode)."""
        n = C.c_long()
        A.check(A.load().minim_lj_list_builds(state.ptr, C.byref(n)))
        return n.value

    def maxNeighbours(self, state):
        """LjNd::maxNeighbours: the longest neighbour list of the last build
        of the potential behind state (cutoff mode)."""
        n = C.c_int()
        A.check(A.load().minim_lj_max_neighbours(state.ptr, C.byref(n)))
        return n.value

    def _create(self):
        L = A.load()
        h = C.c_void_p()
        A.check(L.minim_lj_create(self.nDim, self.sigma, self.epsilon, C.byref(h)))
        hd = _Handle(h.value, L.minim_potential_destroy)
        if self.cutoff > 0:
            b = _f64(self.box)
            A.check(L.minim_lj_set_cutoff(hd.ptr, self.cutoff, _dptr(b), self.skin))
        return hd


class Lj2d(LjNd):
    def __init__(self, sigma=1.0, epsilon=1.0):
        super().__init__(2, sigma, epsilon)


class Lj3d(LjNd):
    def __init__(self, sigma=1.0, epsilon=1.0):
        super().__init__(3, sigma, epsilon)


class BarAndHinge(Potential):
    """include/potentials/BarAndHinge.h:11-70 with explicit rigidities (types 0, 1, 2)."""

    potentialType = 1

    def __init__(self):
        super().__init__()
        self.triangles = None
        self.bonds = None
        self.hingesList = None
        self.kBond = None
        self.kHinge = None
        self.length0 = None
        self.theta0 = None
        self.force = None

    def setTriangulation(self, tri):
        self.triangles = np.asarray(tri, dtype=np.int32).reshape(-1, 3)
        return self

    def setBondList(self, b):
        self.bonds = np.asarray(b, dtype=np.int32).reshape(-1, 2)
        return self

    def setHingeList(self, h):
        self.hingesList = np.asarray(h, dtype=np.int32).reshape(-1, 4)
        return self

    def setRigidity(self, kBond, kHinge):
        self.kBond = kBond
        self.kHinge = kHinge
        return self

    def setLength0(self, l0):
        self.length0 = l0
        return self

    def setTheta0(self, t0):
        self.theta0 = t0
        return self

    def setForce(self, f):
        self.force = _f64(f)
        return self

    def _create(self):
        raise RuntimeError("BarAndHinge handles are created by handle()")

    def handle(self, coords):
        """Explicit lists are passed as given; a triangulation is handed to the
        library, which derives the bars and hinges in the reference's order."""
        L = A.load()
        x = _f64(coords)
        nnode = x.size // 3
        if self.kBond is None or self.kHinge is None:
            raise ValueError("BarAndHinge requires the rigidities to be set")
        self._keep = []
        p = A.BaHParams()
        p.n_node = nnode
        p.l0 = math.nan
        p.theta0 = math.nan
        explicit = self.bonds is not None and self.hingesList is not None
        if explicit:
            bb = np.ascontiguousarray(self.bonds, dtype=np.int32).reshape(-1)
            hh = np.ascontiguousarray(self.hingesList, dtype=np.int32).reshape(-1)
            p.n_bar, p.bars = bb.size // 2, bb.ctypes.data_as(A.ip)
            p.n_hinge, p.hinges = hh.size // 4, hh.ctypes.data_as(A.ip)
            self._keep += [bb, hh]
        else:
            if self.triangles is None:
                raise ValueError("BarAndHinge: The triangulation has not been set.")
            tt = np.ascontiguousarray(self.triangles, dtype=np.int32).reshape(-1)
            p.triangles, p.n_triangle = tt.ctypes.data_as(A.ip), tt.size // 3
            self._keep.append(tt)

        def arr(v):
            a = _f64(v)
            self._keep.append(a)
            return _dptr(a)

        if np.isscalar(self.kBond) and np.isscalar(self.kHinge):
            p.k_bar, p.k_hinge = float(self.kBond), float(self.kHinge)
        else:
            if not explicit:
                raise ValueError("per-element rigidities need explicit bond and hinge lists")
            p.bar_k = arr(np.broadcast_to(self.kBond, (p.n_bar,)))
            p.hinge_k = arr(np.broadcast_to(self.kHinge, (p.n_hinge,)))
        if self.length0 is not None:
            if np.isscalar(self.length0):
                p.l0 = float(self.length0)
            else:
                p.bar_l0 = arr(self.length0)
        if self.theta0 is not None:
            if np.isscalar(self.theta0):
                p.theta0 = float(self.theta0)
            else:
                p.hinge_theta0 = arr(self.theta0)
        if self.force is not None:
            self._keep.append(self.force)
            p.force = _dptr(self.force)
            p.force_per_node = 1 if self.force.size == x.size else 0
        h = C.c_void_p()
        A.check(L.minim_bah_create(C.byref(p), C.byref(h)))
        hd = _Handle(h.value, L.minim_potential_destroy)
        if self.fixedDofs:
            idx = np.ascontiguousarray(self.fixedDofs, dtype=np.int32)
            A.check(L.minim_potential_set_fixed(hd.ptr, idx.ctypes.data_as(A.ip), idx.size))
        A.check(L.minim_potential_init(hd.ptr, _dptr(x), x.size))
        return hd


def triangulation_lists(tri):
    """Bond and hinge lists of a triangulation in the reference's order
    (BarAndHinge::computeBondList / computeHingeList, BarAndHinge.cpp:70-147),
    built in O(T) from an edge map."""
    tri = np.asarray(tri, dtype=np.int64).reshape(-1, 3)
    st = np.sort(tri, axis=1)
    pairs = np.concatenate([st[:, [0, 1]], st[:, [0, 2]], st[:, [1, 2]]])
    bonds = np.unique(pairs, axis=0).astype(np.int32)
    edge = {}
    for it, t in enumerate(tri.tolist()):
        for k in range(3):
            a, b = t[k], t[(k + 1) % 3]
            edge.setdefault((min(a, b), max(a, b)), []).append(it)
    hinges = []
    tl = tri.tolist()
    for it1, t1 in enumerate(tl):
        cand = set()
        for k in range(3):
            a, b = t1[k], t1[(k + 1) % 3]
            for it2 in edge[(min(a, b), max(a, b))]:
                if it2 > it1:
                    cand.add(it2)
        for it2 in sorted(cand):
            t2 = tl[it2]
            s1 = [a in t2 for a in t1]
            if sum(s1) != 2:
                continue
            a = s1.index(False)
            nodes = [t1[a], t1[(a + 1) % 3], t1[(a + 2) % 3], [b for b in t2 if b not in t1][0]]
            hinges.append(nodes)
    return bonds, np.asarray(hinges, dtype=np.int32).reshape(-1, 4)


# ----------------------------------------------------------------- State
class State:
    """include/State.h:12-76 with coordinates resident on the GPU."""

    def __init__(self, pot, coords, precision="fp64"):
        L = A.load()
        x = _f64(coords)
        self.pot = pot
        self._ph = pot.handle(x)
        h = C.c_void_p()
        A.check(L.minim_state_create(self._ph.ptr, _dptr(x), x.size, _PREC[precision], C.byref(h)))
        self._h = _Handle(h.value, L.minim_state_destroy)
        self.ndof = x.size

    @property
    def ptr(self):
        return self._h.ptr

    def copy(self):
        L = A.load()
        h = C.c_void_p()
        A.check(L.minim_state_copy(self.ptr, C.byref(h)))
        s = State.__new__(State)
        s.pot, s._ph, s.ndof = self.pot, self._ph, self.ndof
        s._h = _Handle(h.value, L.minim_state_destroy)
        return s

    def _info(self):
        n, p, c, f = C.c_size_t(), C.c_int(), C.c_double(), C.c_int()
        A.check(A.load().minim_state_info(self.ptr, C.byref(n), C.byref(p), C.byref(c), C.byref(f)))
        return n.value, p.value, c.value, bool(f.value)

    @property
    def convergence(self):
        return self._info()[2]

    @convergence.setter
    def convergence(self, v):
        A.check(A.load().minim_state_set_convergence(self.ptr, float(v)))

    @property
    def isFailed(self):
        return self._info()[3]

    @property
    def precision(self):
        return "fp64" if self._info()[1] == A.MINIM_FP64 else "fp32"

    def setPrecision(self, p):
        A.check(A.load().minim_state_set_precision(self.ptr, _PREC[p]))

    def failed(self):
        A.check(A.load().minim_state_failed(self.ptr))

    def energy(self, coords=None):
        e = C.c_double()
        x = None if coords is None else _f64(coords)
        A.check(A.load().minim_state_energy(self.ptr, None if x is None else _dptr(x), C.byref(e)))
        return e.value

    def gradient(self, coords=None, out=None):
        g = _out(out, self.ndof)
        x = None if coords is None else _f64(coords)
        A.check(A.load().minim_state_gradient(self.ptr, None if x is None else _dptr(x), _dptr(g)))
        return g

    def componentEnergy(self, component):
        """State::componentEnergy: energy of the potential's elements of one type."""
        e = C.c_double()
        A.check(A.load().minim_state_component_energy(self.ptr, int(component), C.byref(e)))
        return e.value

    def energyGradient(self, coords=None, out=None):
        """(energy, gradient); out: an optional float64 array of ndof entries
        that receives the gradient (reused across calls, no new pages)."""
        g = _out(out, self.ndof)
        e = C.c_double()
        x = None if coords is None else _f64(coords)
        A.check(A.load().minim_state_energy_gradient(self.ptr, None if x is None else _dptr(x), C.byref(e),
                                                     _dptr(g)))
        return e.value, g

    # single-GPU equivalents of the parallel variants (src/State.cpp:99-274)
    blockEnergy = procEnergy = allEnergy = energy
    blockGradient = procGradient = allGradient = gradient

    def coords(self, new=None):
        if new is not None:
            x = _f64(new)
            if x.size != self.ndof:
                raise ValueError("State: coordinates have the wrong size.")
            A.check(A.load().minim_state_set_coords(self.ptr, _dptr(x)))
            return None
        out = np.zeros(self.ndof)
        A.check(A.load().minim_state_get_coords(self.ptr, _dptr(out)))
        return out

    blockCoords = coords

    def applyConstraints(self, data):
        d = np.array(data, dtype=np.float64)
        A.check(A.load().minim_state_apply_constraints(self.ptr, _dptr(d)))
        return d

    def device_coords(self):
        return A.load().minim_state_device_coords(self.ptr)

    def setStream(self, stream_ptr):
        A.check(A.load().minim_state_set_stream(self.ptr, stream_ptr))


# ----------------------------------------------------------------- minimisers
class Minimiser:
    """include/Minimiser.h:13-33."""

    kind = None

    def __init__(self):
        self.maxIter = 100000
        self.linesearch = "backtracking"
        self.iter = 0
        self.stats = {}
        self._h = None

    def setMaxIter(self, n):
        self.maxIter = int(n)
        self._h = None
        return self

    def setLinesearch(self, method):
        if method not in ("backtracking", "none"):
            raise ValueError("Invalid line search method.")
        self.linesearch = method
        self._h = None
        return self

    def _params(self):
        p = A.MinimiserParams()
        p.kind = self.kind
        p.max_iter = self.maxIter
        p.linesearch = A.MINIM_LS_BACKTRACKING if self.linesearch == "backtracking" else A.MINIM_LS_NONE
        return p

    def _handle(self):
        if self._h is None:
            h = C.c_void_p()
            A.check(A.load().minim_minimiser_create(C.byref(self._params()), C.byref(h)))
            self._h = _Handle(h.value, A.load().minim_minimiser_destroy)
        return self._h

    def _finish(self, r):
        self.iter = r.iterations
        self.stats = dict(iterations=r.iterations, converged=bool(r.converged), n_energy=r.n_energy,
                          n_gradient=r.n_gradient, n_energy_reused=r.n_energy_reused, iter_fp32=r.iter_fp32,
                          iter_fp64=r.iter_fp64, energy=r.energy, grad_rms=r.grad_rms, dt_max=r.dt_max,
                          grad_max=r.grad_max, seconds=r.seconds,
                          crossings={"1e-%d" % (k + 1): (r.cross_seconds[k], r.cross_iter[k]) for k in range(8)
                                     if r.cross_iter[k] >= 0})

    def minimise(self, state, adjust=None, download=True):
        """Minimiser::minimise(state, adjustState) / minimise(state, "<fields>-<N>")
        (src/Minimiser.cpp:27-69).  With download=False the result stays in the
        State's device coordinates and None is returned."""
        L = A.load()
        h = self._handle()
        out = np.zeros(state.ndof) if download or isinstance(adjust, str) else None
        r = A.Result()
        if isinstance(adjust, str):
            A.check(L.minim_minimise_log(h.ptr, state.ptr, adjust.encode(), _dptr(out), C.byref(r)))
        else:
            cb = A.ADJUST_FN(0)
            if adjust is not None:
                def _cb(it, sp, user):
                    adjust(it, state)
                cb = A.ADJUST_FN(_cb)
            A.check(L.minim_minimise(h.ptr, state.ptr, cb, None, _dptr(out) if out is not None else None,
                                     C.byref(r)))
        self._finish(r)
        return out


class Lbfgs(Minimiser):
    """include/minimisers/Lbfgs.h:14-38 (+ precision schedule)."""

    kind = A.MINIM_LBFGS

    def __init__(self):
        super().__init__()
        self.m = 5
        self.maxStep = 0.0
        self.initStep = 1.0
        self.precisionMode = "state"
        self.switchRms = 0.0
        self.batchControl = "device"
        self.batchBudget = 0
        self.batchBlock = 0

    def setBatchControl(self, mode):
        """Lbfgs::setBatchControl: "device" (the default: device decisions, the
        iteration's launches issued directly), "graph" (the same captured as a
        HIP graph and replayed), "host", or "resident" (one workgroup runs each
        instance's whole minimisation in one launch; FP64 only)."""
        if mode not in ("graph", "device", "host", "resident"):
            raise ValueError("Lbfgs: batch control must be graph, device, host or resident.")
        self.batchControl = mode
        self._h = None
        return self

    def setBatchBlock(self, elements_per_thread):
        """Lbfgs::setBatchBlock: elements per thread of the batch passes (4, 8
        or 16; 0 = the autotuned size for the batch's shape, else 8, or 4 below
        2^20 elements in all)."""
        if elements_per_thread not in (0, 4, 8, 16):
            raise ValueError("Lbfgs: the batch block must be 4, 8 or 16 elements per thread (or 0).")
        self.batchBlock = int(elements_per_thread)
        self._h = None
        return self

    def autotuneBatch(self, states, blocks=(4, 8, 16), replays=32):
        """Lbfgs::autotuneBatch: times the first `replays` iterations of
        minimiseBatch(states) per block size (the States are not changed) and
        stores the fastest for batches of the same shape.  Returns {block:
        seconds}."""
        L = A.load()
        n = len(states)
        arr = (C.c_void_p * max(n, 1))(*[s.ptr for s in states])
        bv = np.ascontiguousarray(blocks, dtype=np.int32)
        secs = np.zeros(len(bv))
        best = C.c_int()
        A.check(L.minim_lbfgs_autotune_batch(self._handle().ptr, arr, n, bv.ctypes.data_as(A.ip), len(bv),
                                             int(replays), secs.ctypes.data_as(A.dp), C.byref(best)))
        return {int(b): float(t) for b, t in zip(bv, secs)}

    def setBatchBudget(self, sites):
        """Lbfgs::setBatchBudget: under device control, instances are admitted
        smallest first while the running ones hold at most this many sites,
        until half have finished (0: every instance from the start)."""
        if sites < 0:
            raise ValueError("Lbfgs: the batch budget must be non-negative.")
        self.batchBudget = int(sites)
        self._h = None
        return self

    def batchReport(self):
        """How the last minimiseBatch ran: iterations, replays, host_syncs,
        graph_launches, captures (graph captures: one more each time the
        launch grid shrank to the instances still running)."""
        it = C.c_int()
        rp, hs, gl, cp = C.c_long(), C.c_long(), C.c_long(), C.c_long()
        blk = C.c_int()
        A.check(A.load().minim_batch_report(self._handle().ptr, C.byref(it), C.byref(rp), C.byref(hs), C.byref(gl),
                                            C.byref(cp), C.byref(blk)))
        return dict(control=self.batchControl, iterations=it.value, replays=rp.value, host_syncs=hs.value,
                    graph_launches=gl.value, captures=cp.value, block=blk.value)

    def setM(self, m):
        self.m = int(m)
        self._h = None
        return self

    def setMaxStep(self, v):
        self.maxStep = float(v)
        self._h = None
        return self

    def setInitStep(self, v):
        self.initStep = float(v)
        self._h = None
        return self

    def setPrecision(self, mode, switchRms=0.0):
        if mode not in ("state", "fp64", "fp32", "mixed"):
            raise ValueError("Lbfgs: precision must be state, fp64, fp32 or mixed.")
        self.precisionMode = mode
        self.switchRms = float(switchRms)
        self._h = None
        return self

    def autotune(self, state, blocks=(256, 512, 1024, 2048, 4096), reps=10):
        """Launch sizes of the direction and Gram passes timed on this State's
        problem (Lbfgs::autotune, minim_lbfgs_autotune); the fastest of each is
        kept in the library's process-wide cache for every later minimisation
        of the same shape.  Returns {"direction": {blocks: ms}, "gram": {...},
        "best": (direction, gram)}."""
        bv = np.ascontiguousarray(blocks, dtype=np.int32)
        times = np.empty(2 * bv.size)
        bd, bg = C.c_int(), C.c_int()
        A.check(A.load().minim_lbfgs_autotune(self._handle().ptr, state.ptr, bv.ctypes.data_as(A.ip), bv.size,
                                              int(reps), _dptr(times), C.byref(bd), C.byref(bg)))
        return {"direction": {int(b): float(times[2 * i]) for i, b in enumerate(bv)},
                "gram": {int(b): float(times[2 * i + 1]) for i, b in enumerate(bv)},
                "best": (bd.value, bg.value)}

    def minimiseBatch(self, states):
        """Lbfgs::minimiseBatch: independent minimisations of many States (one-fluid
        PhaseFields on periodic grids, FP64) in one launch per phase.  Each
        result stays in its State's device coordinates; returns one stats dict
        per State, in order.  The reference runs such instances as States on
        disjoint rank subsets (include/State.h:23)."""
        L = A.load()
        h = self._handle()
        n = len(states)
        arr = (C.c_void_p * max(n, 1))(*[s.ptr for s in states])
        res = (A.Result * max(n, 1))()
        A.check(L.minim_minimise_batch(h.ptr, arr, n, res))
        out = []
        for i in range(n):
            r = res[i]
            out.append(dict(iterations=r.iterations, converged=bool(r.converged), n_energy=r.n_energy,
                            n_gradient=r.n_gradient, n_energy_reused=r.n_energy_reused, iter_fp32=r.iter_fp32,
                            iter_fp64=r.iter_fp64, energy=r.energy, grad_rms=r.grad_rms, seconds=r.seconds,
                            start_seconds=r.start_seconds))
        return out

    def _params(self):
        p = super()._params()
        p.m = self.m
        p.max_step = self.maxStep
        p.init_step = self.initStep
        p.precision_mode = {"state": 0, "fp64": 1, "fp32": 2, "mixed": 3}[self.precisionMode]
        p.switch_rms = self.switchRms
        p.batch_control = {"device": 0, "graph": 1, "host": 2, "resident": 3}[self.batchControl]
        p.batch_budget = self.batchBudget
        p.batch_block = self.batchBlock
        return p


class Fire(Minimiser):
    """include/minimisers/Fire.h:9-36.  Fire(dtMax) disables the line search (Fire.cpp:11-13)."""

    kind = A.MINIM_FIRE

    def __init__(self, dtMax=None):
        super().__init__()
        self.dtMax = 0.0
        if dtMax is not None:
            self.dtMax = float(dtMax)
            self.linesearch = "none"

    def setDtMax(self, v):
        self.dtMax = float(v)
        self._h = None
        return self

    def setDt0(self, v):
        """Initial time step (no reference counterpart; 0 = dtMax as in Fire.cpp:43)."""
        self.dt0 = float(v)
        self._h = None
        return self

    def setMaxForceConvergence(self, dim):
        """Converge on the largest per-particle force norm (groups of dim DOFs) instead of the RMS."""
        self.maxForceDim = int(dim)
        self._h = None
        return self

    def _params(self):
        p = super()._params()
        p.dt_max = self.dtMax
        p.dt0 = getattr(self, "dt0", 0.0)
        p.max_force_dim = getattr(self, "maxForceDim", 0)
        return p

    def minimise(self, state, adjust=None, download=True):
        out = super().minimise(state, adjust, download)
        self.dtMax = self.stats["dt_max"]  # Fire::init stores its estimate (Fire.cpp:29-37)
        return out


class GradDescent(Minimiser):
    """include/minimisers/GradDescent.h:9-23."""

    kind = A.MINIM_GRADDESCENT

    def __init__(self):
        super().__init__()
        self.alpha = 0.1

    def setAlpha(self, a):
        self.alpha = float(a)
        self._h = None
        return self

    def _params(self):
        p = super()._params()
        p.alpha = self.alpha
        return p


# ----------------------------------------------------------------- kernel level
def dot(precision, a_ptr, b_ptr, n, stream=None):
    """CommGrid::dotProduct on device vectors (FP64 result)."""
    out = C.c_double()
    A.check(A.load().minim_dot(_PREC[precision], a_ptr, b_ptr, n, C.byref(out), stream))
    return out.value


def potential_eval(handle, precision, x_ptr, ndof, g_ptr=None, want_energy=True, d_ptr=None, t=0.0, stream=None):
    """Energy and/or gradient of an initialised C potential on device vectors."""
    e = C.c_double()
    A.check(A.load().minim_potential_eval(handle.ptr, _PREC[precision], x_ptr, ndof, d_ptr, float(t),
                                          C.byref(e) if want_energy else None, g_ptr, stream))
    return e.value if want_energy else None


def potential_eval_async(handle, precision, x_ptr, ndof, g_ptr=None, e_dev_ptr=None, d_ptr=None, t=0.0,
                         stream=None, want_energy=None):
    """Evaluation without a host synchronisation; the energy lands in e_dev_ptr
    (computed but not copied out when want_energy and e_dev_ptr is None)."""
    we = (e_dev_ptr is not None) if want_energy is None else bool(want_energy)
    A.check(A.load().minim_potential_eval_async(handle.ptr, _PREC[precision], x_ptr, ndof, d_ptr, float(t),
                                                1 if we else 0, e_dev_ptr, g_ptr, stream))


def potential_eval_mixed(handle, x_ptr, ndof, g_ptr=None, want_energy=True, d_ptr=None, t=0.0, stream=None):
    """Evaluation with FP64 coordinates and FP32 direction/gradient vectors."""
    e = C.c_double()
    A.check(A.load().minim_potential_eval_mixed(handle.ptr, x_ptr, ndof, d_ptr, float(t),
                                                C.byref(e) if want_energy else None, g_ptr, stream))
    return e.value if want_energy else None


def potential_eval_step(handle, precision_x, precision_v, x_ptr, ndof, d_ptr, t, xout_ptr, g_ptr, stream=None):
    """Fused accepted step: energy and gradient at x + t*d, x + t*d written to xout."""
    e = C.c_double()
    A.check(A.load().minim_potential_eval_step(handle.ptr, _PREC[precision_x], _PREC[precision_v], x_ptr, ndof, d_ptr,
                                               float(t), xout_ptr, C.byref(e), g_ptr, stream))
    return e.value


def potential_time_kernel(handle, precision, x_ptr, ndof, g_ptr=None, want_energy=True, reps=20, stream=None):
    """Average ms per launch of the evaluation kernel alone (HIP events)."""
    ms = C.c_double()
    A.check(A.load().minim_potential_time_kernel(handle.ptr, _PREC[precision], x_ptr, ndof, 1 if want_energy else 0,
                                                 g_ptr, int(reps), stream, C.byref(ms)))
    return ms.value


def potential_time_steps(handle, precision, x_ptr, ndof, g_ptr, e_dev_ptr, steps=200, warm_ms=100.0, stream=None):
    """Whole evaluations timed with HIP events around the loop and around each
    kernel launch inside it: dict(step_ms, kernel_median_ms, kernel_min_ms,
    kernel_max_ms, kernel_mean_ms)."""
    out = (C.c_double * 5)()
    A.check(A.load().minim_potential_time_steps(handle.ptr, _PREC[precision], x_ptr, ndof, g_ptr, e_dev_ptr,
                                                int(steps), float(warm_ms), stream, out))
    return dict(step_ms=out[0], kernel_median_ms=out[1], kernel_min_ms=out[2], kernel_max_ms=out[3],
                kernel_mean_ms=out[4])


def autotune_clear():
    """Forget every launch shape the autotuners stored in this process."""
    A.check(A.load().minim_autotune_clear())


def potential_constrain(handle, precision, v_ptr, n, stream=None):
    A.check(A.load().minim_potential_constrain(handle.ptr, _PREC[precision], v_ptr, n, stream))
