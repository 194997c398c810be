#!/usr/bin/env python3
"""Benchmark: energy+gradient evals/s and time-to-converge of the 256^3
three-fluid PhaseField, mixed precision vs FP64, on one MI355X.

One step = one energy+gradient evaluation of the whole 256^3 x 3 field (the
stencil pass of the minimiser's hot path), inputs resident in HBM.  The
headline value is the mixed path's evaluation (FP32 storage and arithmetic,
FP64 energy accumulation); the FP64 path is reported beside it, as is the
time-to-converge of full device L-BFGS (m=5) runs in both precisions and the
reference library timed on the host cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

METRIC = "energy+grad evals/s and time-to-converge, 256³ 3-fluid PF: mixed vs FP64"
HBM_PEAK_GBS = 8000.0
GRID = (256, 256, 256)
NF = 3
SIGMA = [1.0, 0.8, 1.2]
WIDTH = 2.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_input(grid, seed=2024):
    import numpy as np

    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 1.0, (int(np.prod(grid)), NF))
    return (u / u.sum(axis=1, keepdims=True)).reshape(-1)


def time_evals(m, h, prec, x, g, e_dev, n, steps, warmup, stream, with_energy_copy=True):
    import torch

    with torch.cuda.stream(stream):
        gp = None if g is None else g.data_ptr()
        for _ in range(warmup):
            m.potential_eval_async(h, prec, x.data_ptr(), n, g_ptr=gp,
                                   e_dev_ptr=e_dev.data_ptr() if with_energy_copy else None,
                                   want_energy=True, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(steps):
            m.potential_eval_async(h, prec, x.data_ptr(), n, g_ptr=gp,
                                   e_dev_ptr=e_dev.data_ptr() if with_energy_copy else None,
                                   want_energy=True, stream=stream.cuda_stream)
        t1.record(stream)
        torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps  # ms


def converge(m, x0, mode, max_iter=20000, repeats=2):
    """L-BFGS from x0 to convergence.  The first run is cold (history buffers
    allocated, result downloaded to the host); later runs reuse the
    minimiser's device buffers and leave the result on the device.  Each run
    starts from a fresh State built from x0 outside the timed region."""
    import torch

    pf = m.PhaseField().setNFluid(NF).setGridSize(list(GRID)).setSurfaceTension(SIGMA).setInterfaceSize(WIDTH)
    mn = m.Lbfgs().setM(5).setMaxIter(max_iter).setPrecision(mode)
    times = []
    for r in range(repeats):
        s = None
        s = m.State(pf, x0, "fp64")
        e0 = s.energy()
        torch.cuda.synchronize()
        t = time.perf_counter()
        mn.minimise(s, download=(r == 0))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    st = dict(mn.stats)
    st.update(seconds=times[-1], seconds_cold=times[0], initial_energy=e0, final_energy=s.energy())
    return st


def cfg1_leg(m, with_cpu=True):
    """The reference's CPU-runnable case: a 128x128x1 binary field (phi ~ U[-0.1, 0.1]
    from mt19937_64, sigma = w = 1), L-BFGS m=5 to the reference tolerance.  GPU
    time-to-converge (FP64 and mixed; warm buffers, result left on the device)
    beside the reference library's minimise on the host cores."""
    import torch
    from mt64 import cfg1_input

    x0 = cfg1_input()
    pf = m.PhaseField().setNFluid(1).setGridSize([128, 128, 1]).setSurfaceTension(1.0).setInterfaceSize(1.0)
    out = {}
    for mode in ("fp64", "mixed"):
        mn = m.Lbfgs().setM(5).setMaxIter(20000).setPrecision(mode)
        times = []
        for r in range(2):
            s = None
            s = m.State(pf, x0, "fp64")
            torch.cuda.synchronize()
            t = time.perf_counter()
            mn.minimise(s, download=(r == 0))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        out[mode] = dict(seconds=times[-1], seconds_cold=times[0], iterations=mn.stats["iterations"],
                         converged=mn.stats["converged"], energy=s.energy())
        # the same field as a batch of one: the device-controlled iteration
        # (Lbfgs.minimiseBatch), which needs one host synchronisation per 8
        # iterations instead of two per iteration
        bt = []
        for r in range(2):
            s = m.State(pf, x0, "fp64")
            torch.cuda.synchronize()
            t = time.perf_counter()
            res = mn.minimiseBatch([s])[0]
            torch.cuda.synchronize()
            bt.append(time.perf_counter() - t)
        out[mode + "_batch_of_one"] = dict(seconds=bt[-1], seconds_cold=bt[0], iterations=res["iterations"],
                                           converged=res["converged"], energy=res["energy"])
        if mode == "fp64":
            # ... and under resident control: one workgroup runs the whole
            # minimisation in one launch
            mr = m.Lbfgs().setM(5).setMaxIter(20000).setPrecision(mode).setBatchControl("resident")
            bt = []
            for r in range(2):
                s = m.State(pf, x0, "fp64")
                torch.cuda.synchronize()
                t = time.perf_counter()
                res = mr.minimiseBatch([s])[0]
                torch.cuda.synchronize()
                bt.append(time.perf_counter() - t)
            out[mode + "_batch_of_one_resident"] = dict(seconds=bt[-1], seconds_cold=bt[0],
                                                        iterations=res["iterations"], converged=res["converged"],
                                                        energy=res["energy"])
    if with_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        from oracle_bind import RefState, ref_available

        if ref_available():
            r = RefState.phasefield((128, 128, 1), x0, 1, 1.0, [1.0], [1.0], "hard")
            t = time.perf_counter()
            _, res = r.minimise("lbfgs", m=5, max_iter=20000)
            out["reference_cpu"] = dict(seconds=time.perf_counter() - t, iterations=res["iterations"],
                                        energy=res["energy"], cores=threads, kind="reference")
    return out


def time_device_evals(m, h, prec, x, g, n, steps, warmup):
    """ms per whole evaluation (kernels, reductions, energy copy) on device inputs."""
    import torch

    stream = torch.cuda.Stream()
    e_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    return time_evals(m, h, prec, x, g, e_dev, n, steps, warmup, stream)


def pillar_layout(n):
    """A pillar array on a base (the pf-pillar-collapse geometry of the reference's
    benchmarks, test/benchmark/pf-pillar-collapse/main.cpp:42-60, on an n^3
    grid): solid for y < n/16 and, for y < n/2, in square pillars of side n/4
    on a period of n/2 in x and z.  Three fluids: the diffuse solid (fixed), a
    liquid above y = 3n/4 and gas elsewhere, with a pressure on the liquid."""
    import numpy as np

    x, y, z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pillar = ((x % (n // 2)) < n // 4) & ((z % (n // 2)) < n // 4) & (y < n // 2)
    solid = ((y < n // 16) | pillar).reshape(-1).astype(np.uint8)
    liquid = (y > 3 * n // 4).reshape(-1)
    rng = np.random.default_rng(11)
    c = np.zeros((n ** 3, 3))
    c[:, 0] = solid
    c[:, 1] = np.where(liquid, 1.0, 0.0) * (1 - solid)
    c[:, 2] = np.where(liquid, 0.0, 1.0) * (1 - solid)
    c += rng.uniform(-0.02, 0.02, c.shape) * (1 - solid)[:, None]
    return solid, c.reshape(-1)


def pf_general_leg(m, n=256, steps=20, warmup=3, ref_n=48):
    """The general PhaseField kernel (solid nodes, pressure, a fixed fluid; the
    tiled kernel covers neither) at 256^3 x 3 fluids: kernel and whole-evaluation
    times in FP32 and FP64 against the HBM roofline, and the reference library's
    evaluation of the same layout at ref_n^3 on the host cores."""
    import numpy as np
    import torch

    solid, x0 = pillar_layout(n)
    pf = (m.PhaseField().setNFluid(3).setGridSize([n, n, n]).setSurfaceTension(SIGMA).setInterfaceSize(WIDTH)
          .setSolid(solid).setFixFluid(0).setPressure([0.0, 0.05, 0.0]))
    nd = x0.size
    t0 = time.perf_counter()
    h = pf.handle(x0)
    setup_s = time.perf_counter() - t0
    stream = torch.cuda.Stream()
    out = dict(workload="PhaseField %d^3 x 3 fluids, pillar array on a base (%.1f%% solid), fixed solid fluid, "
                        "pressure on the liquid, hard density constraint" % (n, 100.0 * solid.mean()),
               setup_s=setup_s)
    sites = n ** 3
    for prec, b in (("fp32", 4), ("fp64", 8)):
        dt = torch.float32 if prec == "fp32" else torch.float64
        x = torch.from_numpy(x0).to("cuda", dt)
        g = torch.empty_like(x)
        ms = time_device_evals(m, h, prec, x, g, nd, steps, warmup)
        msk = m.potential_time_kernel(h, prec, x.data_ptr(), nd, g_ptr=g.data_ptr(), reps=steps,
                                      stream=stream.cuda_stream)
        alg = sites * 3 * (b + b) + sites * 4  # coordinates in, gradient out, 4-byte forward-bond code
        out[prec] = dict(eval_ms=ms, evals_per_s=1e3 / ms, kernel_ms=msk, algorithmic_bytes=alg,
                         achieved_GBs=alg / (msk * 1e-3) / 1e9, frac_of_hbm_peak=alg / (msk * 1e-3) / 1e9 / HBM_PEAK_GBS)
        del x, g
        torch.cuda.empty_cache()
    from oracle_bind import RefState, ref_available

    if ref_available():
        threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        rs, rx = pillar_layout(ref_n)
        r = RefState.phasefield_ext((ref_n,) * 3, rx, 3, 1.0, SIGMA, [WIDTH], "hard", solid=rs,
                                    fix_fluid=[True, False, False], pressure=[0.0, 0.05, 0.0])
        r.energy_gradient(rx)
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            r.energy_gradient(rx)
        dt_ref = (time.perf_counter() - t) / reps
        out["reference_cpu"] = dict(grid=[ref_n] * 3, eval_s=dt_ref, sites_per_s=ref_n ** 3 / dt_ref, cores=threads,
                                    kind="reference")
        out["gpu_sites_per_s_fp64"] = sites / (out["fp64"]["eval_ms"] * 1e-3)
    return out


def pfu_leg(m, x0, steps=20, warmup=3, ref_n=48):
    """The element form (PhaseFieldUnstructured) on the headline workload:
    256^3 x 3 fluids, one seven-site element per site and fluid.  Whole
    evaluations and the element kernel alone in FP32 and FP64 against the HBM
    roofline, the energy beside the gather form's (the same sum arranged
    differently), and the reference library's element form at ref_n^3 on the
    host cores."""
    import numpy as np
    import torch

    pf = m.PhaseFieldUnstructured().setNFluid(NF).setGridSize(list(GRID)).setSurfaceTension(SIGMA)
    pf.setInterfaceSize(WIDTH)
    t0 = time.perf_counter()
    h = pf.handle(x0)
    setup_s = time.perf_counter() - t0
    nd = x0.size
    nel, list_bytes = pf.elementList(handle=h)
    stream = torch.cuda.Stream()
    out = dict(workload="PhaseFieldUnstructured %s x %d fluids (the headline input), %d elements, %.1f MB as "
                        "stored on the device" % ("x".join(map(str, GRID)), NF, nel, list_bytes / 1e6),
               setup_s=setup_s, elements=nel, list_bytes=list_bytes)
    gather = m.PhaseField().setNFluid(NF).setGridSize(list(GRID)).setSurfaceTension(SIGMA).setInterfaceSize(WIDTH)
    hg = gather.handle(x0)
    for prec, b in (("fp32", 4), ("fp64", 8)):
        dt = torch.float32 if prec == "fp32" else torch.float64
        x = torch.from_numpy(x0).to("cuda", dt)
        g = torch.empty_like(x)
        ms = time_device_evals(m, h, prec, x, g, nd, steps, warmup)
        msk = m.potential_time_kernel(h, prec, x.data_ptr(), nd, g_ptr=g.data_ptr(), reps=steps,
                                      stream=stream.cuda_stream)
        e_el = m.potential_eval(h, prec, x.data_ptr(), nd)
        e_ga = m.potential_eval(hg, prec, x.data_ptr(), nd)
        alg = list_bytes + nd * (b + b)  # element list as stored, coordinates in, gradient out
        out[prec] = dict(eval_ms=ms, evals_per_s=1e3 / ms, kernel_ms=msk, algorithmic_bytes=alg,
                         achieved_GBs=alg / (msk * 1e-3) / 1e9, frac_of_hbm_peak=alg / (msk * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         energy=e_el, gather_form_energy=e_ga, energy_rel_diff=abs(e_el - e_ga) / abs(e_ga))
        del x, g
        torch.cuda.empty_cache()
    del h, hg
    from oracle_bind import RefState, ref_available

    if ref_available():
        threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        rx = make_input((ref_n,) * 3)
        r = RefState.phasefield_unstructured((ref_n,) * 3, rx, NF, 1.0, SIGMA, [WIDTH], "hard")
        r.energy_gradient(rx)
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            r.energy_gradient(rx)
        dt_ref = (time.perf_counter() - t) / reps
        out["reference_cpu"] = dict(grid=[ref_n] * 3, eval_s=dt_ref, sites_per_s=ref_n ** 3 / dt_ref, cores=threads,
                                    kind="reference")
        out["gpu_sites_per_s_fp64"] = int(np.prod(GRID)) / (out["fp64"]["eval_ms"] * 1e-3)
    return out


def lj_leg(m, n_side=64, density=0.85, fire_max_iter=100000, ref_side=10):
    """Lennard-Jones at the benchmark size: n_side^3 particles at number density
    0.85 in a periodic cube, truncated-shifted at rc = 2.5 sigma with a 0.3 sigma
    skin.  Particles start on a simple cubic lattice jittered by U[-0.1, 0.1]
    (seeded); uniform-random placement puts pairs at r -> 0, whose forces
    (~1e26) no time-stepping minimiser can integrate.  Reports the evaluation
    rate (FP64, FP32) and FIRE (dt0 0.002, dtMax 0.02, no line search) to a
    largest per-particle force below 1e-6.  The reference's LjNd has no cutoff
    or box; beside it, the reference-semantics all-pairs potential is timed on
    the GPU and by the reference library on a ref_side^3 cluster."""
    import numpy as np
    import torch

    N = n_side ** 3
    a = (1.0 / density) ** (1.0 / 3.0)
    L = n_side * a
    rng = np.random.default_rng(2025)
    grid = np.stack(np.meshgrid(*[np.arange(n_side) * a] * 3, indexing="ij"), -1).reshape(-1, 3)
    x0 = (grid + rng.uniform(-0.1, 0.1, grid.shape)).reshape(-1)
    pot = m.Lj3d().setCutoff(2.5, [L, L, L], 0.3)
    out = dict(workload="LJ %d particles, density %.2f, periodic cube L=%.3f, rc 2.5, skin 0.3" % (N, density, L))
    list_n = density * 4.0 / 3.0 * np.pi * 2.8 ** 3  # neighbours within rc + skin (full list)
    for prec, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        h = pot.handle(x0)
        x = torch.tensor(x0, dtype=dt, device="cuda")
        g = torch.empty_like(x)
        ms = time_device_evals(m, h, prec, x, g, 3 * N, 50, 5)
        msk = m.potential_time_kernel(h, prec, x.data_ptr(), 3 * N, g_ptr=g.data_ptr(), reps=50)
        b = x.element_size()
        alg = N * 3 * b * 2 + 4 * list_n * N  # positions in, gradient out, neighbour list
        # VALU: about 25 flop per pair over the half list within rc + skin
        flops = 25.0 * list_n * N / 2
        vpeak = 78.6e12 if prec == "fp64" else 157.3e12
        out[prec] = dict(eval_ms=ms, evals_per_s=1e3 / ms, kernel_ms=msk, algorithmic_bytes=alg,
                         achieved_GBs=alg / (msk * 1e-3) / 1e9, frac_of_hbm_peak=alg / (msk * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         eval_achieved_GBs=alg / (ms * 1e-3) / 1e9, algorithmic_flops=flops,
                         achieved_TFLOPs=flops / (msk * 1e-3) / 1e12, frac_of_valu_peak=flops / (msk * 1e-3) / vpeak)
        del h, x, g
    s = m.State(pot, x0, "fp64")
    s.convergence = 1e-6
    f = m.Fire(0.02).setDt0(0.002).setMaxForceConvergence(3).setMaxIter(fire_max_iter)
    builds0 = pot.listBuilds(s)
    torch.cuda.synchronize()
    t = time.perf_counter()
    f.minimise(s, download=False)
    torch.cuda.synchronize()
    out["fire"] = dict(seconds=time.perf_counter() - t, iterations=f.stats["iterations"],
                       converged=f.stats["converged"], max_force=f.stats["grad_max"], energy=s.energy(),
                       max_iter=fire_max_iter, list_builds=pot.listBuilds(s) - builds0,
                       max_force_crossings={k: dict(seconds=v[0], iteration=v[1])
                                            for k, v in f.stats["crossings"].items()})
    # reference semantics (all pairs, no cutoff) at ref_n particles: GPU beside the reference library
    rn = ref_side ** 3
    ra = 1.12
    rg = np.stack(np.meshgrid(*[np.arange(ref_side) * ra] * 3, indexing="ij"), -1).reshape(-1, 3)
    xr = (rg + rng.uniform(-0.05, 0.05, rg.shape)).reshape(-1)
    hr = m.Lj3d().handle(xr)
    xt = torch.tensor(xr, dtype=torch.float64, device="cuda")
    gt = torch.empty_like(xt)
    ms = time_device_evals(m, hr, "fp64", xt, gt, xr.size, 20, 3)
    pairs = rn * (rn - 1) / 2
    out["all_pairs_reference_semantics"] = dict(particles=rn, gpu_eval_ms=ms, gpu_pairs_per_s=pairs / (ms * 1e-3))
    from oracle_bind import RefState, ref_available

    if ref_available():
        r = RefState.lj(3, xr)
        t = time.perf_counter()
        r.energy_gradient(xr)
        el = time.perf_counter() - t
        out["all_pairs_reference_semantics"].update(
            reference_cpu_eval_s=el, reference_cpu_pairs_per_s=pairs / el,
            reference_cores=int(os.environ.get("OMP_NUM_THREADS") or 1))
    return out


def sheet_triangles(nx, ny):
    """Triangulated square lattice: nodes i*ny + j, triangles (a, b, c), (a, c, d)
    per square (a = (i,j), b = (i+1,j), c = (i+1,j+1), d = (i,j+1))."""
    import numpy as np

    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    a = (i * ny + j).reshape(-1)
    b = a + ny
    c = b + 1
    d = a + 1
    return np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3).astype(np.int32)


def bah_leg(m, nside=1024, max_iter=60000, ref_side=48):
    """Bar-and-hinge sheet at the benchmark size: nside x nside nodes, two
    triangles per lattice square (2(nside-1)^2 triangles).  Nodes sit on an
    equilateral triangular lattice (node (i,j) at (i - j/2, j sqrt(3)/2)) so every
    bar has rest length 1 (setLength0(1.0)) and the sheet is flat (setTheta0(pi));
    k_bar = 1, k_hinge = 0.01.  x is compressed by 0.95, the nodes of the two
    x edges are clamped (all their DOFs fixed), z += N(0, 0.01^2) (seeded).
    Reports the evaluation rate (FP64, FP32) and L-BFGS m=10 from that start;
    beside it, the reference library's evaluation on a ref_side^2 sheet."""
    import numpy as np
    import torch

    tri = sheet_triangles(nside, nside)
    i, j = np.meshgrid(np.arange(nside), np.arange(nside), indexing="ij")
    rng = np.random.default_rng(77)
    xyz = np.stack([0.95 * (i - 0.5 * j), (np.sqrt(3) / 2) * j, rng.normal(0, 0.01, i.shape)], -1).reshape(-1, 3)
    x0 = xyz.reshape(-1)
    edge = np.where((i == 0) | (i == nside - 1), 1, 0).reshape(-1).nonzero()[0]
    fixed = np.stack([3 * edge, 3 * edge + 1, 3 * edge + 2], 1).reshape(-1)
    pot = m.BarAndHinge().setTriangulation(tri).setRigidity(1.0, 0.01).setLength0(1.0).setTheta0(np.pi)
    pot.setConstraints(fixed.tolist())
    nnode = nside * nside
    ntri = tri.shape[0]

    def counts(n):  # bars = lattice edges + diagonals; hinges = interior edges
        nb = 2 * n * (n - 1) + (n - 1) ** 2
        return nb, nb - 4 * (n - 1)

    nbar, nhinge = counts(nside)
    out = dict(workload="bar-and-hinge sheet %dx%d nodes, %d triangles" % (nside, nside, ntri))
    t = time.perf_counter()
    h = pot.handle(x0)
    out["setup_s"] = time.perf_counter() - t
    for prec, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        x = torch.tensor(x0, dtype=dt, device="cuda")
        g = torch.empty_like(x)
        ms = time_device_evals(m, h, prec, x, g, x0.size, 20, 3)
        msk = m.potential_time_kernel(h, prec, x.data_ptr(), x0.size, g_ptr=g.data_ptr(), reps=20)
        b = x.element_size()
        # node coordinates in, gradient out, bar (2 idx + k, l0) and hinge (4 idx + k, theta0) streams
        alg = nnode * 3 * b * 2 + nbar * (8 + 2 * b) + nhinge * (16 + 2 * b)
        out[prec] = dict(eval_ms=ms, evals_per_s=1e3 / ms, kernel_ms=msk, algorithmic_bytes=alg,
                         achieved_GBs=alg / (msk * 1e-3) / 1e9, frac_of_hbm_peak=alg / (msk * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         eval_achieved_GBs=alg / (ms * 1e-3) / 1e9)
        del x, g
    # FP64 throughout, and the mixed schedule (FP32 until the gradient RMS is
    # below 1e3 times the tolerance, then FP64 coordinates and arithmetic with
    # FP32 vectors)
    for key, prec in (("lbfgs_m10", "fp64"), ("lbfgs_m10_mixed", "mixed")):
        s = m.State(pot, x0, "fp64")
        lb = m.Lbfgs().setM(10).setMaxIter(max_iter).setPrecision(prec)
        e0 = s.energy()
        torch.cuda.synchronize()
        t = time.perf_counter()
        lb.minimise(s, download=False)
        torch.cuda.synchronize()
        out[key] = dict(seconds=time.perf_counter() - t, iterations=lb.stats["iterations"],
                        iter_fp32=lb.stats["iter_fp32"], converged=lb.stats["converged"],
                        grad_rms=lb.stats["grad_rms"], initial_energy=e0, energy=s.energy(), max_iter=max_iter,
                        tolerance=s.convergence,
                        grad_rms_crossings={k: dict(seconds=v[0], iteration=v[1])
                                            for k, v in lb.stats["crossings"].items()})
        del s
    from oracle_bind import RefState, ref_available

    if ref_available():
        rt = sheet_triangles(ref_side, ref_side)
        ri, rj = np.meshgrid(np.arange(ref_side), np.arange(ref_side), indexing="ij")
        rx = np.stack([0.95 * (ri - 0.5 * rj), (np.sqrt(3) / 2) * rj, rng.normal(0, 0.01, ri.shape)], -1).reshape(-1)
        r = RefState.bah(rt, rx, 1.0, 0.01, l0=1.0, t0=np.pi)
        n, t = 0, time.perf_counter()
        while n < 20 and time.perf_counter() - t < 5:
            r.energy_gradient(rx)
            n += 1
        el = (time.perf_counter() - t) / n
        rb, rh = counts(ref_side)
        out["reference_cpu"] = dict(nodes=ref_side * ref_side, triangles=rt.shape[0], eval_s=el,
                                    elements_per_s=(rb + rh) / el,
                                    cores=int(os.environ.get("OMP_NUM_THREADS") or 1))
        out["gpu_elements_per_s_fp64"] = (nbar + nhinge) / (out["fp64"]["eval_ms"] * 1e-3)
    return out


def per_state_leg(m, n_inst, workers, sides, inputs, make_pf):
    """The batch workload with one State per minimise call, spread over
    `workers` host threads, each with its own HIP stream, State objects and
    (reused) minimiser buffers, largest instances first."""
    import concurrent.futures as cf
    import threading

    import numpy as np
    import torch

    order = sorted(range(n_inst), key=lambda j: -sides[j])
    t = time.perf_counter()
    states = {j: m.State(make_pf(j), inputs(j), "fp64") for j in order}
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t
    tls = threading.local()

    def run(j):
        if not hasattr(tls, "stream"):
            torch.cuda.set_device(0)
            tls.stream = torch.cuda.Stream()
            tls.lb = m.Lbfgs().setM(5).setMaxIter(20000)
        s = states[j]
        s.setStream(tls.stream.cuda_stream)
        t0 = time.perf_counter()
        tls.lb.minimise(s, download=False)
        st = tls.lb.stats
        t1 = time.perf_counter()
        return dict(seconds=t1 - t0, done=t1 - start[0], iterations=st["iterations"], converged=st["converged"],
                    evals=st["n_energy"] + st["n_gradient"])

    # largest first: each worker allocates its minimiser buffers once, on its first instance
    start = [0.0]
    with cf.ThreadPoolExecutor(workers) as ex:
        t = time.perf_counter()
        start[0] = t
        res = list(ex.map(run, order))
        wall = time.perf_counter() - t
    secs = np.array([r["seconds"] for r in res])
    done = np.array([r["done"] for r in res])
    evals = sum(r["evals"] for r in res)
    del states
    # p50_s / p99_s: each call's own duration (from when its worker took it);
    # completion: from the start of the run until the instance was done, the
    # batch's p50_s / p99_s
    return dict(workers=workers, setup_s=setup_s, wall_s=wall, evals=evals, evals_per_s=evals / wall,
                instances_per_s=n_inst / wall, converged=int(sum(r["converged"] for r in res)),
                p50_s=float(np.percentile(secs, 50)), p99_s=float(np.percentile(secs, 99)),
                p50_completion_s=float(np.percentile(done, 50)), p99_completion_s=float(np.percentile(done, 99)),
                iterations_total=int(sum(r["iterations"] for r in res)))


def batch_leg(m, n_inst=512, workers=8, seed=4040, cpu_budget=10.0, per_state=True,
              budgets=(500000, 1000000, 2000000, 4000000)):
    """Many independent fields: n_inst 2-D binary PhaseField instances with side
    s_j = round(exp(U[ln 32, ln 512])), phi ~ U[-0.1, 0.1] and interface width
    w_j ~ U[1, 3] (all seeded), each minimised by L-BFGS m=5 to the reference
    tolerance.  Measured two ways, States set up before each timed region:
    Lbfgs.minimiseBatch (one launch per phase over every running instance; the
    per-instance time-to-converge is the wall time from the start until that
    instance converged, and also reported from its admission), and one State
    per call spread over `workers` host threads, each with its own HIP stream
    (ordered largest first; per call and from the start of the run).  Reports
    the aggregate evaluation rate and the p50/p99 per-instance times, and the
    reference library on a bounded sample of the same instances.  per_state
    False leaves out the one-State-per-call comparison.  budgets: the batch
    also runs with each of these limits on the sites of running instances
    (Lbfgs.setBatchBudget: smallest first until half have finished, then
    every remaining instance).  The FP64 batch also runs under resident control
    (one workgroup per instance runs its whole minimisation in one launch)."""

    import numpy as np
    import torch

    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), n_inst))).astype(int)
    widths = rng.uniform(1.0, 3.0, n_inst)
    seeds = rng.integers(0, 2 ** 62, n_inst)

    def inputs(j):
        return np.random.default_rng(int(seeds[j])).uniform(-0.1, 0.1, int(sides[j]) ** 2)

    def make_pf(j):
        s = int(sides[j])
        return m.PhaseField().setNFluid(1).setGridSize([s, s, 1]).setSurfaceTension(1.0).setInterfaceSize(
            float(widths[j]))

    # ---- FP64 and mixed, with the iterations driven on the device (per-instance
    # decisions in kernels, the iteration replayed as a HIP graph) and, for
    # comparison, by the host (one host synchronisation per phase)
    batched = {}
    # the batch passes' block size, tuned per precision schedule on the first
    # 32 iterations of this batch (Lbfgs.autotuneBatch, outside the timed runs)
    tuned = {}
    for prec, control in (("fp64", "device"), ("mixed", "device"), ("fp64", "resident")):
        tstates = [m.State(make_pf(j), inputs(j), "fp64") for j in range(n_inst)]
        tuned[prec if control == "device" else prec + "_resident"] = m.Lbfgs().setM(5).setMaxIter(20000).setPrecision(
            prec).setBatchControl(control).autotuneBatch(tstates, replays=32)
        del tstates
    # device control (the default) with and without budgets; the same
    # iteration replayed as a HIP graph, host control and resident control
    runs = [(p, "device", 0) for p in ("fp64", "mixed")] + [(p, "device", b) for b in budgets for p in ("fp64", "mixed")]
    runs += [(p, c, 0) for c in ("graph", "host") for p in ("fp64", "mixed")] + [("fp64", "resident", 0)]
    for prec, control, budget in runs:
        t = time.perf_counter()
        bstates = [m.State(make_pf(j), inputs(j), "fp64") for j in range(n_inst)]
        torch.cuda.synchronize()
        bsetup = time.perf_counter() - t
        lb = m.Lbfgs().setM(5).setMaxIter(20000).setPrecision(prec).setBatchControl(control).setBatchBudget(budget)
        t = time.perf_counter()
        bres = lb.minimiseBatch(bstates)
        torch.cuda.synchronize()
        bwall = time.perf_counter() - t
        rep = lb.batchReport()
        bsecs = np.array([r["seconds"] for r in bres])
        # from admission to convergence: the batch's counterpart of one call's duration
        bown = np.array([r["seconds"] - r["start_seconds"] for r in bres])
        bevals = sum(r["n_energy"] + r["n_gradient"] for r in bres)
        key = prec + ("" if control == "device" else "_%s_control" % control) + (
            "_budget_%gM_sites" % (budget / 1e6) if budget else "")
        batched[key] = dict(control=control, budget_sites=budget, host_syncs=rep["host_syncs"], replays=rep["replays"],
                             host_syncs_per_iteration=rep["host_syncs"] / max(1, rep["replays"]),
                             graph_launches=rep["graph_launches"], graph_captures=rep["captures"],
                             block_elements_per_thread=rep["block"],
                             setup_s=bsetup, wall_s=bwall, evals=bevals, evals_per_s=bevals / bwall,
                             instances_per_s=n_inst / bwall, converged=int(sum(r["converged"] for r in bres)),
                             p50_s=float(np.percentile(bsecs, 50)), p99_s=float(np.percentile(bsecs, 99)),
                             p50_since_admission_s=float(np.percentile(bown, 50)),
                             p99_since_admission_s=float(np.percentile(bown, 99)),
                             iterations_total=int(sum(r["iterations"] for r in bres)),
                             iterations_max=int(max(r["iterations"] for r in bres)),
                             iterations_fp32=int(sum(r["iter_fp32"] for r in bres)),
                             longest=[dict(side=int(sides[j]), iterations=bres[j]["iterations"],
                                           seconds=bres[j]["seconds"])
                                      for j in sorted(range(n_inst), key=lambda j: -bres[j]["iterations"])[:6]])
        del bstates

    out = dict(workload="%d 2-D binary PhaseField instances, sides %d..%d, %d sites in total; L-BFGS m=5 each"
                        % (n_inst, sides.min(), sides.max(), int((sides ** 2).sum())),
               site_bytes_per_iteration="FP64: 8 B x (4m + 10) vector streams per site; mixed: FP32 vectors",
               block_tuning_seconds_first_32_iterations={p: {str(k): v for k, v in t.items()} for p, t in tuned.items()},
               batched=batched)
    if per_state:
        out["per_state_threads"] = per_state_leg(m, n_inst, workers, sides, inputs, make_pf)
    from oracle_bind import RefState, ref_available

    if ref_available() and cpu_budget > 0:
        threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        done, el, sites_done = 0, 0.0, 0
        for j in range(n_inst):
            x0 = inputs(j)
            r = RefState.phasefield((int(sides[j]), int(sides[j]), 1), x0, 1, 1.0, [1.0], [float(widths[j])], "hard")
            t = time.perf_counter()
            r.minimise("lbfgs", m=5, max_iter=20000)
            el += time.perf_counter() - t
            done += 1
            sites_done += int(sides[j]) ** 2
            if el > cpu_budget:
                break
        out["reference_cpu"] = dict(kind="reference", cores=threads, instances=done, sites=sites_done, seconds=el,
                                    instances_per_s=done / el,
                                    sample="the first %d instances in generation order, minimised one after "
                                           "another (reference library, OpenMP %d threads)" % (done, threads))
    return out


NATIVE_REF = os.path.join(ROOT, "oracle", "_ref", "libminim_ref_native.so")
NATIVE_FLAGS = "g++ -O3 -std=c++14 -fopenmp -march=x86-64-v4 (AVX-512), serial+OpenMP build (no MPI)"


def host_cpu():
    """Model name and physical core count of the host CPU (/proc/cpuinfo)."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and core is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    if core is not None:
        cores.add((phys, core))
    return model, (len(cores) or None), os.cpu_count()


def cpu_baseline(x, budget_s=12.0):
    """The reference library (oracle/_ref, built from its own sources) timed on
    the host cores: State construction excluded, then as many
    State::energyGradient calls on the full 256^3 x 3 field as fit the budget."""
    import numpy as np

    threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle_bind import REF_SO, RefState, ref_available

    model, phys, logical = host_cpu()
    host = dict(cpu_model=model, physical_cores=phys, logical_cpus=logical, ranks=1, threads_per_rank=threads,
                build=NATIVE_FLAGS if os.path.basename(REF_SO) == os.path.basename(NATIVE_REF) else
                "g++ -O3 -std=c++14 -fopenmp -march=x86-64-v3 -ffp-contract=off (the bit-exact test build)")
    if ref_available():
        t = time.perf_counter()
        r = RefState.phasefield(GRID, x, NF, 1.0, SIGMA, [WIDTH], "hard")
        setup = time.perf_counter() - t
        n, t = 0, time.perf_counter()
        while True:
            r.energy_gradient(x)
            n += 1
            if time.perf_counter() - t > budget_s or n >= 10:
                break
        el = time.perf_counter() - t
        return dict(value=n / el, unit="evals/s", cores=threads, kind="reference",
                    sample="%d State::energyGradient calls on the 256^3x3 field (reference compiled from source, "
                           "OpenMP %d threads, 1 process); State construction %.1f s excluded" % (n, threads, setup),
                    **host)
    from oracle_bind import OraclePF

    o = OraclePF(GRID, NF, 1.0, SIGMA, [WIDTH])
    n, t = 0, time.perf_counter()
    while True:
        o.energy(x)
        o.gradient(x)
        n += 1
        if time.perf_counter() - t > budget_s:
            break
    el = time.perf_counter() - t
    host.update(threads_per_rank=1, build="oracle/oracle.c, gcc -O2")
    return dict(value=n / el, unit="evals/s", cores=1, kind="port",
                sample="%d energy+gradient evaluations of the 256^3x3 field by the serial C restatement" % n, **host)


def traffic_from_profiles():
    """HBM bytes per launch of the FP32 stencil from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    p = os.path.join(ROOT, "profiles", "pf_stencil_256cubed_traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get("fp32_eg_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="batch mode: the batched minimisations only, without the one-State-per-call comparison")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--fire-iters", type=int, default=100000, help="particles mode: FIRE iteration cap")
    ap.add_argument("--lbfgs-iters", type=int, default=60000,
                    help="sheet L-BFGS iteration cap (it converges in about 48,000 iterations)")
    ap.add_argument("--mode", choices=["full", "stencil", "converge", "particles", "batch", "general", "elements",
                                       "cfg1"],
                    default="full",
                    help="stencil: only the evaluation timings plus a 256 MiB copy (the calibration "
                         "kernel for rocprofv3 byte counters); converge: only time-to-converge")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the whole 256^3 field fits one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # the reference's CPU legs run its native-ISA build when there is one
    if os.path.exists(NATIVE_REF) and "MINIM_REF_SO" not in os.environ:
        os.environ["MINIM_REF_SO"] = NATIVE_REF

    import numpy as np
    import torch

    import minim_amd as m

    torch.cuda.set_device(0)
    n = int(np.prod(GRID)) * NF
    log("input %s x %d (%d DOF)" % (GRID, NF, n))
    x0 = make_input(GRID)
    pf = m.PhaseField().setNFluid(NF).setGridSize(list(GRID)).setSurfaceTension(SIGMA).setInterfaceSize(WIDTH)
    h = pf.handle(x0)
    stream = torch.cuda.Stream()
    x64 = torch.from_numpy(x0).cuda()
    x32 = x64.float()
    g64 = torch.empty_like(x64)
    g32 = torch.empty_like(x32)
    e_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    if args.mode == "batch":
        del x64, x32, g64, g32
        torch.cuda.empty_cache()
        res = batch_leg(m, cpu_budget=0 if args.no_cpu else 10.0, per_state=not args.batch_only)
        log(json.dumps(res, indent=1))
        print(json.dumps(res), flush=True)
        return

    if args.mode == "general":
        del x64, x32, g64, g32
        torch.cuda.empty_cache()
        res = pf_general_leg(m, steps=args.steps, warmup=args.warmup)
        log(json.dumps(res, indent=1))
        print(json.dumps(res), flush=True)
        return

    if args.mode == "elements":
        del x64, x32, g64, g32
        torch.cuda.empty_cache()
        res = pfu_leg(m, x0, steps=args.steps, warmup=args.warmup)
        log(json.dumps(res, indent=1))
        print(json.dumps(res), flush=True)
        return

    if args.mode == "cfg1":
        del x64, x32, g64, g32
        torch.cuda.empty_cache()
        res = cfg1_leg(m, with_cpu=not args.no_cpu)
        log(json.dumps(res, indent=1))
        print(json.dumps(res), flush=True)
        return

    if args.mode == "particles":
        del x64, x32, g64, g32
        torch.cuda.empty_cache()
        res = dict(lj=lj_leg(m, fire_max_iter=args.fire_iters), bah=bah_leg(m, max_iter=args.lbfgs_iters))
        log(json.dumps(res, indent=1))
        print(json.dumps(res), flush=True)
        return

    if args.mode == "converge":
        del x64, x32, g64, g32
        torch.cuda.empty_cache()
        res = {mode: converge(m, x0, mode) for mode in ("fp64", "mixed")}
        for mode, r in res.items():
            log("%s: %.3f s (cold %.3f s), %d it (%d fp32), E %.6g -> %.6g, gRMS %.3g, nE %d nG %d" % (
                mode, r["seconds"], r["seconds_cold"], r["iterations"], r["iter_fp32"], r["initial_energy"], r["final_energy"],
                r["grad_rms"], r["n_energy"], r["n_gradient"]))
        print(json.dumps(res), flush=True)
        return

    # One step = one whole evaluation (stencil kernel, cross-block reduction,
    # energy copy) on the stream.  W untimed steps, then at least 100 ms of
    # untimed steps (steady clocks), then K timed steps: HIP events bracket the
    # loop (ms_per_step) and each stencil launch inside it (kernel times).
    for _ in range(args.warmup):
        m.potential_eval_async(h, "fp32", x32.data_ptr(), n, g_ptr=g32.data_ptr(), e_dev_ptr=e_dev.data_ptr(),
                               stream=stream.cuda_stream)
    torch.cuda.synchronize()
    t32 = m.potential_time_steps(h, "fp32", x32.data_ptr(), n, g32.data_ptr(), e_dev.data_ptr(), steps=args.steps,
                                 warm_ms=100.0, stream=stream.cuda_stream)
    t64 = m.potential_time_steps(h, "fp64", x64.data_ptr(), n, g64.data_ptr(), e_dev.data_ptr(), steps=args.steps,
                                 warm_ms=100.0, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    ms32, ms32k = t32["step_ms"], t32["kernel_median_ms"]
    ms64, ms64k = t64["step_ms"], t64["kernel_median_ms"]
    # the kernel is part of the step: its per-launch time cannot exceed the step's
    assert ms32k <= ms32 and ms64k <= ms64, (t32, t64)
    log("E+G fp32 %.4f ms/step (stencil median %.4f, min %.4f, max %.4f), fp64 %.4f ms/step (stencil median %.4f)"
        % (ms32, ms32k, t32["kernel_min_ms"], t32["kernel_max_ms"], ms64, ms64k))
    if args.mode == "stencil":
        # calibration kernel for the byte counters: a plain copy of 1 GiB into
        # another 1 GiB (far beyond the on-chip caches), 1 GiB read + 1 GiB
        # written per launch
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        ms_e = m.potential_time_kernel(h, "fp32", x32.data_ptr(), n, reps=args.steps, stream=stream.cuda_stream)
        log("E-only fp32 kernel %.3f ms" % ms_e)
        print(json.dumps({"fp32": t32, "fp64": t64, "fp32_e_ms": ms_e}), flush=True)
        return

    # PCIe-inclusive rate: State::energyGradient(coords) with host FP64 vectors in
    # and out, as a caller of the reference's host API would use it
    host_ms = {}
    g_host = np.zeros(n)  # the caller's gradient array, reused across calls
    for prec in ("fp32", "fp64"):
        s = m.State(pf, x0, prec)
        s.energyGradient(x0, out=g_host)
        t = time.perf_counter()
        for _ in range(5):
            s.energyGradient(x0, out=g_host)
        host_ms[prec] = (time.perf_counter() - t) / 5 * 1e3
        del s
    log("host-boundary E+G (upload + eval + download): fp32 %.1f ms, fp64 %.1f ms" % (host_ms["fp32"], host_ms["fp64"]))

    sites = int(np.prod(GRID))
    bytes32 = sites * NF * (4 + 4)  # coordinates read once + gradient written once
    bytes64 = sites * NF * (8 + 8)
    ach32 = bytes32 / (ms32k * 1e-3) / 1e9
    ach64 = bytes64 / (ms64k * 1e-3) / 1e9

    ttc = None
    if not args.no_converge:
        del x64, x32, g64, g32
        torch.cuda.empty_cache()
        res = {}
        for mode in ("fp64", "mixed"):
            log("L-BFGS time-to-converge, %s ..." % mode)
            res[mode] = converge(m, x0, mode)
            log("  %s: %.3f s (cold %.3f s), %d it, E %.6g -> %.6g, gRMS %.3g" % (
                mode, res[mode]["seconds"], res[mode]["seconds_cold"], res[mode]["iterations"], res[mode]["initial_energy"],
                res[mode]["final_energy"], res[mode]["grad_rms"]))
        e_init = res["fp64"]["initial_energy"]
        ttc = dict(
            mixed_s=res["mixed"]["seconds"], fp64_s=res["fp64"]["seconds"],
            speedup=res["fp64"]["seconds"] / res["mixed"]["seconds"],
            mixed=res["mixed"], fp64=res["fp64"],
            energy_agreement=abs(res["mixed"]["final_energy"] - res["fp64"]["final_energy"]) / max(
                abs(res["fp64"]["final_energy"]), abs(e_init)),
            tolerance="gradient RMS < 1e-8 (reference convergence for sigma_mean=1, res=1); energies compared "
                      "with the absolute floor 1e-8*E_initial since the minimum is E~0")

    log("cfg1 (128x128x1 binary) time-to-converge ...")
    cfg1 = cfg1_leg(m, with_cpu=not args.no_cpu)
    log("  cfg1: " + json.dumps(cfg1))

    # the other workloads of the path, each beside the reference library where it can run them;
    # a failure here is recorded in the output and does not hide the headline line
    others = {}
    for name, fn in (("pf_general_256cubed_pillars", lambda: pf_general_leg(m, steps=args.steps, warmup=args.warmup)),
                     ("pf_element_form_256cubed", lambda: pfu_leg(m, x0, steps=args.steps, warmup=args.warmup)),
                     ("lj_periodic_262144", lambda: lj_leg(m, fire_max_iter=args.fire_iters)),
                     ("bar_and_hinge_1024x1024", lambda: bah_leg(m, max_iter=args.lbfgs_iters)),
                     ("batch_512_instances", lambda: batch_leg(m, cpu_budget=0 if args.no_cpu else 10.0))):
        log("%s ..." % name)
        try:
            others[name] = fn()
        except Exception as exc:  # noqa: BLE001
            others[name] = {"error": "%s: %s" % (type(exc).__name__, exc)}
        log("  %s: %s" % (name, json.dumps(others[name])))
    cpu = None if args.no_cpu else cpu_baseline(x0, args.cpu_budget)

    out = {
        "metric": METRIC,
        "value": 1e3 / ms32,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: concentrations u~U[0,1] per fluid (numpy default_rng seed 2024), normalised per site",
        "config": {"workload": "PhaseField 256^3 periodic grid x 3 fluids, surface tensions (1.0,0.8,1.2), "
                               "interface width 2, hard density constraint; one step = one energy+gradient "
                               "evaluation (FP32 storage/arithmetic, FP64 energy accumulation)",
                   "grid": list(GRID), "n_fluid": NF, "dof": n, "parallelism": "single GPU"},
        "fp64_evals_per_s": 1e3 / ms64,
        "fp64_ms_per_step": ms64,
        "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach32 / HBM_PEAK_GBS, "traffic": traffic_from_profiles(),
                     "kernel": "pf_tile_kernel<float,float,3,hard,E+G>", "kernel_ms": ms32k, "algorithmic_bytes": bytes32,
                     "kernel_ms_stats": {k: v for k, v in t32.items() if k != "step_ms"},
                     "timing": "median of %d per-launch HIP-event durations of the stencil inside the timed step "
                               "loop, after >= 100 ms of untimed steps" % args.steps,
                     "fp64": {"achieved": ach64, "frac": ach64 / HBM_PEAK_GBS, "kernel_ms": ms64k,
                              "kernel_ms_stats": {k: v for k, v in t64.items() if k != "step_ms"},
                              "algorithmic_bytes": bytes64}},
        "host_boundary": {"fp32_ms_per_eval": host_ms["fp32"], "fp64_ms_per_eval": host_ms["fp64"],
                          "what": "State::energyGradient(coords) with host FP64 coordinates in and host FP64 "
                                  "gradient out into the caller's array (PCIe transfers through pinned, "
                                  "double-buffered 16 MiB chunks included); not the headline value"},
        "time_to_converge": ttc,
        "cfg1_time_to_converge": cfg1,
        "other_workloads": others,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
