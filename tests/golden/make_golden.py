"""Write the golden fixtures of tests/golden/ from the reference library.

Runs in the build container only: it needs oracle/_ref/libminim_ref.so, the
reference (sjavis/minim) compiled from its own sources by oracle/Makefile.
Run single-threaded (OMP_NUM_THREADS=1), the reference's only bit-deterministic
mode.  Inputs are drawn from numpy's default_rng with the seeds below and are
stored in the fixtures, so nothing downstream depends on the generator.

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py [fixture ...]
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
from oracle_bind import RefState, ref_available  # noqa: E402
from mt64 import minima_input  # noqa: E402


def pf_cases():
    rng = np.random.default_rng(20240601)
    specs = [
        # grid, nfluid, surface tension, interface size, density, resolution
        ((6, 5, 4), 1, [1.0], [1.0], "hard", 1.0),
        ((8, 8, 1), 1, [1.0], [], "hard", 1.0),
        ((1, 7, 1), 1, [0.7], [1.5], "hard", 1.0),
        ((2, 3, 5), 1, [1.3], [2.0], "hard", 2.0),
        ((4, 5, 6), 3, [1.0, 0.8, 1.2], [2.0], "hard", 1.0),
        ((3, 4, 2), 2, [0.9], [1.2], "hard", 1.0),
        ((4, 4, 4), 4, [1.0, 0.9, 1.1, 0.8], [1.3], "soft", 1.0),
        ((2, 3, 2), 3, [1.0, 1.0, 1.0], [], "none", 1.0),
        ((5, 2, 3), 3, [1.0, 0.8, 1.2], [2.0], "fixed", 0.5),
    ]
    out = []
    for grid, nf, st, iw, dens, res in specs:
        n = int(np.prod(grid)) * nf
        if nf == 1:
            x = rng.uniform(-1.1, 1.1, n)
        else:
            u = rng.uniform(0.0, 1.0, (n // nf, nf))
            x = (u / u.sum(axis=1, keepdims=True)).reshape(-1) + rng.uniform(-0.05, 0.05, n)
        r = RefState.phasefield(grid, x, nf, res, st, iw, dens)
        e, g = r.energy(x), r.gradient(x)
        out.append(dict(grid=list(grid), nfluid=nf, surface_tension=st, interface_size=iw, density=dens,
                        resolution=res, x=x.tolist(), energy=e, gradient=g.tolist()))
    return out


def lbfgs_cases():
    rng = np.random.default_rng(777)
    out = []
    specs = [
        ("binary-16x16", (16, 16, 1), 1, [1.0], [1.0], 5, 20000, "backtracking", 0.0, 1.0),
        ("binary-24x20-m3", (24, 20, 1), 1, [1.0], [2.0], 3, 20000, "backtracking", 0.0, 1.0),
        ("binary-8x8x8-maxstep", (8, 8, 8), 1, [1.0], [1.5], 5, 20000, "backtracking", 0.3, 1.0),
        ("ternary-6x6x6", (6, 6, 6), 3, [1.0, 0.8, 1.2], [2.0], 5, 20000, "backtracking", 0.0, 1.0),
        ("binary-12x12-nols", (12, 12, 1), 1, [1.0], [2.0], 5, 300, "none", 0.0, 0.1),
    ]
    specs += [
        ("binary-32x16-stripes", (32, 16, 1), 1, [1.0], [1.0], 5, 20000, "backtracking", 0.0, 1.0),
        ("ternary-12x8x4-slabs", (12, 8, 4), 3, [1.0, 0.8, 1.2], [1.0], 5, 20000, "backtracking", 0.0, 1.0),
    ]
    for name, grid, nf, st, iw, m, maxit, ls, maxstep, init in specs:
        n = int(np.prod(grid)) * nf
        if "stripes" in name:
            xg = np.arange(grid[0])
            prof = np.where(np.sin(2 * np.pi * (xg + 0.5) / grid[0]) > 0, 1.0, -1.0)
            x0 = (np.repeat(prof, grid[1] * grid[2]) + rng.uniform(-0.05, 0.05, n))
        elif "slabs" in name:
            lab = (np.arange(grid[0]) * 3 // grid[0])
            c = np.zeros((grid[0], nf))
            c[np.arange(grid[0]), lab] = 1.0
            c = 0.9 * c + 0.1 / 3
            x0 = np.repeat(c[:, None, :], grid[1] * grid[2], axis=1).reshape(-1, nf)
            x0 = x0 + rng.uniform(-0.03, 0.03, x0.shape)
            x0 = (x0 / x0.sum(axis=1, keepdims=True)).reshape(-1)
        elif nf == 1:
            x0 = rng.uniform(-0.1, 0.1, n)
        else:
            u = rng.uniform(0.0, 1.0, (n // nf, nf))
            x0 = (u / u.sum(axis=1, keepdims=True)).reshape(-1)
        r = RefState.phasefield(grid, x0, nf, 1.0, st, iw, "hard")
        x, res = r.minimise("lbfgs", m=m, max_iter=maxit, param_a=maxstep, param_b=init, linesearch=ls)
        out.append(dict(name=name, grid=list(grid), nfluid=nf, surface_tension=st, interface_size=iw, m=m,
                        max_iter=maxit, linesearch=ls, max_step=maxstep, init_step=init, x0=x0.tolist(),
                        x=x.tolist(), **res))
    return out


def fire_gd_cases():
    rng = np.random.default_rng(99)
    out = []
    grid = (10, 10, 1)
    x0 = rng.uniform(-0.1, 0.1, 100)
    for kind, pa, ls, maxit in [("fire", 0.0, "backtracking", 3000), ("fire", 0.5, "none", 3000),
                                ("graddescent", 0.1, "backtracking", 3000), ("graddescent", 0.05, "none", 3000)]:
        r = RefState.phasefield(grid, x0, 1, 1.0, [1.0], [2.0], "hard")
        x, res = r.minimise(kind, max_iter=maxit, param_a=pa, linesearch=ls)
        out.append(dict(kind=kind, param=pa, linesearch=ls, max_iter=maxit, grid=list(grid), interface_size=[2.0],
                        x0=x0.tolist(), x=x.tolist(), **res))
    return out


def lj_cases():
    rng = np.random.default_rng(5)
    out = []
    for ndim, n in [(3, 12), (2, 9), (3, 30)]:
        # jittered lattice: no overlapping particles
        side = int(math.ceil(n ** (1.0 / ndim)))
        pts = np.array(np.meshgrid(*[np.arange(side)] * ndim, indexing="ij")).reshape(ndim, -1).T[:n] * 1.15
        x = (pts + rng.uniform(-0.1, 0.1, pts.shape)).reshape(-1)
        r = RefState.lj(ndim, x)
        e, g = r.energy_gradient(x)
        out.append(dict(ndim=ndim, n=n, x=x.tolist(), energy=e, gradient=g.tolist()))
    return out


def sheet(nx, ny):
    """Triangulated square lattice: nodes i*ny+j, two triangles per square."""
    tri = []
    for i in range(nx - 1):
        for j in range(ny - 1):
            a, b, c, d = i * ny + j, (i + 1) * ny + j, (i + 1) * ny + j + 1, i * ny + j + 1
            tri.append([a, b, c])
            tri.append([a, c, d])
    return tri


def bah_cases():
    rng = np.random.default_rng(31)
    out = []
    nx, ny = 5, 4
    tri = sheet(nx, ny)
    xy = np.array([[i, j, 0.0] for i in range(nx) for j in range(ny)], dtype=float)
    x0 = (xy + rng.normal(0, [0.05, 0.05, 0.2], xy.shape)).reshape(-1)
    fixed = [0, 1, 2, 3 * (ny - 1), 3 * (ny - 1) + 1, 3 * (ny - 1) + 2]
    for name, kb, kh, l0, t0, fx, force in [("rest-from-coords", 1.0, 0.1, None, None, [], None),
                                           ("explicit-rest", 2.0, 0.3, 1.0, math.pi, fixed, [0.0, 0.0, -0.05])]:
        r = RefState.bah(tri, x0, kb, kh, l0=l0, t0=t0, fixed_dofs=fx, force3=force)
        el = r.bah_elements()
        xs = x0 + rng.normal(0, 0.05, x0.shape)
        e, g = r.energy_gradient(xs)
        out.append(dict(name=name, tri=tri, kb=kb, kh=kh, l0=l0, t0=t0, fixed=fx, force=force, x0=x0.tolist(),
                        x=xs.tolist(), energy=e, gradient=g.tolist(),
                        bars=el["bars"].tolist(), bar_k=el["bar_k"].tolist(), bar_l0=el["bar_l0"].tolist(),
                        hinges=el["hinges"].tolist(), hinge_k=el["hinge_k"].tolist(),
                        hinge_t0=el["hinge_t0"].tolist()))
    # BarAndHinge_test.cpp:13-56: two-triangle hinge minimised by L-BFGS
    tri2 = [[0, 1, 2], [1, 3, 2]]
    xk = [-1, 0, 1, 0, -math.sqrt(2), 0, 0, math.sqrt(2), 0, 1, 0, 1]
    mins = []
    for t0 in [1e-6, math.pi / 2, math.pi, 3 * math.pi / 2, 2 * math.pi - 1e-6]:
        r = RefState.bah(tri2, xk, 100.0, 1.0)
        r.set_hinge_theta0(0, t0)
        x, res = r.minimise("lbfgs", m=5, max_iter=100000)
        mins.append(dict(theta0=t0, x=x.tolist(), **res))
    out.append(dict(name="two-triangle-hinge", tri=tri2, x0=xk, kb=100.0, kh=1.0, minima=mins))
    return out


def cfg1_cases():
    """The benchmark's CPU-runnable case: 128x128x1 binary field, phi ~ U[-0.1, 0.1]
    from mt19937_64 (mt64.cfg1_input), sigma = w = 1, L-BFGS m=5 to the
    reference tolerance.  The inputs are regenerated from the seed, so only
    the results are stored."""
    from mt64 import cfg1_input

    x0 = cfg1_input()
    r = RefState.phasefield((128, 128, 1), x0, 1, 1.0, [1.0], [1.0], "hard")
    e0 = r.energy(x0)
    x, res = r.minimise("lbfgs", m=5, max_iter=20000)
    return [dict(name="cfg1-128x128-binary", grid=[128, 128, 1], nfluid=1, surface_tension=[1.0],
                 interface_size=[1.0], seed=20240, lo=-0.1, hi=0.1, m=5, max_iter=20000, initial_energy=e0,
                 x_sum=float(x.sum()), x_sumsq=float((x * x).sum()), x_min=float(x.min()), x_max=float(x.max()),
                 **res)]


def minima_cases():
    """L-BFGS runs of the reference whose minimum has interfaces, so that the
    converged energy is far from 0 and a 1e-8 relative comparison has no
    floor to fall back on.  Stored: the recipe of the input (mt64.minima_input),
    the reference's result, and the converged field's mean profile along x
    (which places the interfaces)."""
    specs = [
        dict(name="cfg1-stripes-128x128", grid=[128, 128, 1], nfluid=1, surface_tension=[1.0],
             interface_size=[1.0], init="stripes", seed=20240, amplitude=0.1),
        dict(name="ternary-slabs-48x4x4", grid=[48, 4, 4], nfluid=3, surface_tension=[1.0, 0.8, 1.2],
             interface_size=[1.0], init="slabs", seed=777, amplitude=0.05, nslab=3),
    ]
    out = []
    for c in specs:
        x0 = minima_input(c)
        r = RefState.phasefield(c["grid"], x0, c["nfluid"], 1.0, c["surface_tension"], c["interface_size"], "hard")
        e0 = r.energy(x0)
        x, res = r.minimise("lbfgs", m=5, max_iter=20000)
        g = r.gradient(x)
        prof = x.reshape(c["grid"][0], -1, c["nfluid"]).mean(axis=1)
        out.append(dict(c, m=5, max_iter=20000, initial_energy=e0, profile=prof.tolist(),
                        grad_rms_at_x=float(math.sqrt(np.dot(g, g) / g.size)), **res))
    return out


def pf_ext_cases():
    """PhaseField with solid nodes, wetting, pressure, forces, confinement,
    per-site tensions and fixed volumes, set through the reference's setters.
    Stored per case: the options, the initial state x0 (which fixes the volume
    target and the confinement), an evaluation point x, the energy from an
    energy-only call and from an energy+gradient call (they differ when several
    fluids feel the force, PhaseField.cpp:511), the constrained gradient, a
    random vector and its projection, and energyComponents at x."""
    rng = np.random.default_rng(31337)

    def solid_of(grid, kind):
        s = np.zeros(grid, np.uint8)
        if kind == "wall-x":
            s[0] = 1
        elif kind == "channel-z":
            s[:, :, 0] = 1
            s[:, :, -1] = 1
        elif kind == "pillar":
            s[:, 0] = 1
            s[:2, :4] = 1
        elif kind == "box":
            s[1:3, 1:3, 1:3] = 1
        return s.reshape(-1)

    specs = [
        # grid, nfluid, resolution, surface tension (None: per site), interface size, density, options
        ((8, 7, 1), 1, 0.5, [1.0], [1.0], "hard",
         dict(solid="wall-x", contact_angle=True, pressure=[1.5], force=[0.3, -0.2, 0.0])),
        ((6, 5, 4), 2, 1.0, [0.9], [1.2], "hard",
         dict(solid="channel-z", pressure=[0.5, -0.3], confinement=[2.0, 0.0], volume_fixed=True, vol_const=0.02)),
        ((5, 6, 4), 3, 1.0, [1.0, 0.8, 1.2], [2.0], "hard",
         dict(solid="pillar", fix_fluid=[True, False, False], pressure=[0.0, 0.2, 0.0])),
        ((4, 4, 4), 3, 1.0, None, [1.5], "soft",
         dict(solid="box", force=[0.0, 0.0, -0.7], force_fluids=[1])),
        ((6, 1, 5), 4, 2.0, [1.0, 0.9, 1.1, 0.8], [1.3], "none",
         dict(solid="wall-x", confinement=[0.0, 1.5, 0.0, 0.5], volume=[3.0, 2.0, 1.5, 0.5], vol_const=0.05)),
        ((1, 9, 7), 1, 1.0, None, [1.0], "hard",
         dict(contact_angle=True, force=[0.0, 0.4, 0.1])),
    ]
    out = []
    for grid, nf, res, st, iw, dens, spec in specs:
        ng = int(np.prod(grid))
        opts = {}
        if "solid" in spec:
            opts["solid"] = solid_of(grid, spec["solid"])
        if spec.get("contact_angle"):
            opts["contact_angle"] = rng.uniform(30, 150, ng * nf)
        for k in ("pressure", "force", "force_fluids", "confinement", "volume", "vol_const", "volume_fixed",
                  "fix_fluid"):
            if k in spec:
                opts[k] = spec[k]
        if st is None:
            st = list(rng.uniform(0.6, 1.4, ng * nf))
        x0 = rng.uniform(-1.0, 1.0, ng) if nf == 1 else rng.uniform(0.0, 1.0, ng * nf)
        x = x0 + rng.normal(0, 0.05, x0.size)
        r = RefState.phasefield_ext(grid, x0, nf, res, st, iw, dens, **opts)
        e_only = r.energy(x)
        e_eg, _ = r.energy_gradient(x)
        g = r.gradient(x)
        v = rng.normal(0, 1, x0.size)
        pv = r.constrain(v)
        r2 = RefState.phasefield_ext(grid, x, nf, res, st, iw, dens, **opts)
        comp = r2.components()  # at x (the volume target and confinement then come from x)
        if "volume" in opts:
            # with an explicit volume target the reference's projection divides
            # by totalVolume2, which it leaves unset (PhaseField.cpp:216-231):
            # its constrained vectors are not a reference result
            g, pv = None, None
        jopts = {k: (np.asarray(v).tolist() if hasattr(v, "__len__") else v) for k, v in opts.items()}
        out.append(dict(grid=list(grid), nfluid=nf, resolution=res, surface_tension=list(st), interface_size=iw,
                        density=dens, opts=jopts, x0=x0.tolist(), x=x.tolist(), energy=e_only, energy_eg=e_eg,
                        gradient=None if g is None else g.tolist(), v=v.tolist(),
                        v_constrained=None if pv is None else pv.tolist(), components=comp))
    return out


def examples_cases():
    """The outputs of tests/examples/*_example.cpp built against the
    reference's headers and library: one record per printed line, keyed by its
    label (the numbers that follow it)."""
    import subprocess
    import tempfile

    root = os.path.dirname(os.path.dirname(HERE))
    ref = os.environ.get("MINIM_REF_DIR", "/root/reference")
    lib = os.path.join(root, "oracle", "_ref")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for ex in ("toy2d", "lj3d"):
            exe = os.path.join(tmp, ex)
            subprocess.run(["g++", "-std=c++14", "-O2", "-fopenmp", "-I", os.path.join(ref, "include"),
                            os.path.join(root, "tests", "examples", ex + "_example.cpp"), "-L", lib, "-lminim_ref",
                            "-Wl,-rpath," + lib, "-o", exe], check=True)
            r = subprocess.run([exe], check=True, capture_output=True, text=True)
            for line in r.stdout.splitlines():
                f = line.split()
                out[f[0]] = [float(v) for v in f[1:]]
    return out


def main():
    if not ref_available():
        sys.exit("oracle/_ref/libminim_ref.so is not built (make -C oracle)")
    if os.environ.get("OMP_NUM_THREADS") != "1":
        sys.exit("run with OMP_NUM_THREADS=1")
    makers = dict(pf=pf_cases, lbfgs=lbfgs_cases, fire_gd=fire_gd_cases, lj=lj_cases, bah=bah_cases, cfg1=cfg1_cases,
                  pf_ext=pf_ext_cases, minima=minima_cases, examples=examples_cases)
    only = sys.argv[1:] or list(makers)  # e.g. "make_golden.py pf_ext" writes that fixture alone
    data = {k: makers[k]() for k in only}
    for k, v in data.items():
        with open(os.path.join(HERE, "ref_%s.json" % k), "w") as f:
            json.dump(v, f)
        print(k, len(v))


if __name__ == "__main__":
    main()
