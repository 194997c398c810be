"""GPU tests of Lbfgs.minimiseBatch: many one-fluid PhaseField minimisations in
one launch per phase.

Each instance must behave as minimise(state) does on its own State with the
reference's semantics (src/minimisers/Lbfgs.cpp:41-133,
src/minimisers/linesearch.cpp:8-29, src/Minimiser.cpp:27-38):
  * converged energies within 1e-8 relative of the oracle's FP64 L-BFGS (an
    absolute floor of 1e-8 E_initial), with the energy the batch reports equal
    to the oracle's energy at the batch's own coordinates within 1e-12;
  * the first iteration (a steepest-descent step with backtracking) equal to
    the single-State path's to rounding;
  * an instance's result does not depend on which other instances share the
    launches (bit for bit), since every per-instance reduction sums the same
    block partials in the same order.
"""
import numpy as np
import pytest

from oracle_bind import OraclePF, oracle_lbfgs

pytestmark = pytest.mark.gpu


def cases(seed, sides=(8, 13, 32, 50, 64, 100), extra=((6, 7, 5), (40, 1, 1), (1, 24, 17), (2, 9, 1))):
    rng = np.random.default_rng(seed)
    out = []
    for g in [(s, s, 1) for s in sides] + list(extra):
        n = int(np.prod(g))
        out.append(dict(grid=g, w=float(rng.uniform(1.0, 3.0)), x0=rng.uniform(-0.1, 0.1, n)))
    return out


def make_state(m, c):
    pf = m.PhaseField().setNFluid(1).setGridSize(list(c["grid"])).setSurfaceTension(1.0).setInterfaceSize(c["w"])
    return m.State(pf, c["x0"])


def energy_ok(e, e_ref, e0, rel=1e-8):
    return abs(e - e_ref) <= rel * max(abs(e_ref), abs(e0))


def test_batch_vs_oracle(gpu):
    m = gpu
    cs = cases(7)
    states = [make_state(m, c) for c in cs]
    res = m.Lbfgs().setM(5).setMaxIter(20000).minimiseBatch(states)
    assert len(res) == len(cs)
    for c, s, r in zip(cs, states, res):
        o = OraclePF(c["grid"], 1, 1.0, [1.0], [c["w"]])
        xo, so = oracle_lbfgs(o, c["x0"], m=5, max_iter=20000)
        e0 = o.energy(c["x0"])
        assert so["converged"] and r["converged"], (c["grid"], r, so)
        x = s.coords()
        # the energy the batch reports is the energy of the coordinates it left
        assert abs(r["energy"] - o.energy(x)) <= 1e-12 * max(abs(r["energy"]), abs(e0)), c["grid"]
        assert energy_ok(r["energy"], so["energy"], e0), (c["grid"], r["energy"], so["energy"])
        assert r["grad_rms"] < o.convergence
        g = o.gradient(x)
        assert np.sqrt(np.mean(g * g)) < 1.0001 * o.convergence
        assert r["iter_fp64"] == r["iterations"] + 1 and r["iter_fp32"] == 0
        assert abs(r["iterations"] - so["iterations"]) <= max(10, 0.3 * so["iterations"]), (r, so)


def test_batch_mixed_vs_oracle(gpu):
    """The mixed schedule (FP32 coordinates and arithmetic, then FP64
    coordinates with FP32 vectors) per instance: converged energies within
    1e-8 of the oracle's FP64 L-BFGS, and the switch taken as the single-State
    schedule takes it."""
    m = gpu
    cs = cases(12)
    states = [make_state(m, c) for c in cs]
    res = m.Lbfgs().setM(5).setMaxIter(20000).setPrecision("mixed").minimiseBatch(states)
    for c, s, r in zip(cs, states, res):
        o = OraclePF(c["grid"], 1, 1.0, [1.0], [c["w"]])
        xo, so = oracle_lbfgs(o, c["x0"], m=5, max_iter=20000)
        e0 = o.energy(c["x0"])
        assert r["converged"], (c["grid"], r)
        x = s.coords()
        assert abs(r["energy"] - o.energy(x)) <= 1e-12 * max(abs(r["energy"]), abs(e0)), c["grid"]
        assert energy_ok(r["energy"], so["energy"], e0), (c["grid"], r["energy"], so["energy"])
        g = o.gradient(x)
        assert np.sqrt(np.mean(g * g)) < 1.0001 * o.convergence
        assert r["iter_fp32"] > 0 and r["iter_fp32"] + r["iter_fp64"] == r["iterations"] + 1
        s1 = make_state(m, c)
        lb = m.Lbfgs().setM(5).setMaxIter(20000).setPrecision("mixed")
        lb.minimise(s1)
        assert abs(r["iter_fp32"] - lb.stats["iter_fp32"]) <= max(5, 0.3 * lb.stats["iter_fp32"]), (r, lb.stats)


def test_batch_first_iteration_matches_single(gpu):
    """maxIter 0: one iteration, the first step of Lbfgs.cpp:91-96 with the line
    search.  The gradient of the batch kernel uses the single-State kernel's
    per-site arithmetic, so the step agrees to the rounding of the norms."""
    m = gpu
    cs = cases(8)
    states = [make_state(m, c) for c in cs]
    res = m.Lbfgs().setM(5).setMaxIter(0).minimiseBatch(states)
    for c, s, r in zip(cs, states, res):
        s1 = make_state(m, c)
        lb = m.Lbfgs().setM(5).setMaxIter(0)
        x1 = lb.minimise(s1)
        xb = s.coords()
        assert np.max(np.abs(xb - x1)) <= 1e-14 * max(1.0, np.max(np.abs(x1))), c["grid"]
        assert r["iterations"] == 1 and not r["converged"]
        assert abs(r["energy"] - s1.energy()) <= 1e-12 * abs(s1.energy())


def test_batch_result_independent_of_company(gpu):
    m = gpu
    cs = cases(9, sides=(16, 33, 70), extra=((5, 6, 7),))
    states = [make_state(m, c) for c in cs]
    res = m.Lbfgs().setM(4).setMaxIter(300).minimiseBatch(states)
    for i, c in enumerate(cs):
        alone = make_state(m, c)
        ra = m.Lbfgs().setM(4).setMaxIter(300).minimiseBatch([alone])[0]
        assert np.array_equal(alone.coords(), states[i].coords()), c["grid"]
        assert ra["energy"] == res[i]["energy"] and ra["iterations"] == res[i]["iterations"]
        assert ra["n_energy"] == res[i]["n_energy"]


def test_batch_settings(gpu):
    """maxStep, initStep, m and the line search 'none' are honoured: the batch
    result equals the oracle's under the same settings."""
    m = gpu
    cs = cases(11, sides=(12, 40), extra=((3, 4, 5),))
    for kw in [dict(m=3, max_step=0.05), dict(m=7, init_step=0.3), dict(m=5, linesearch="none", max_iter=60)]:
        states = [make_state(m, c) for c in cs]
        lb = m.Lbfgs().setM(kw["m"]).setMaxIter(kw.get("max_iter", 20000))
        if "max_step" in kw:
            lb.setMaxStep(kw["max_step"])
        if "init_step" in kw:
            lb.setInitStep(kw["init_step"])
        if "linesearch" in kw:
            lb.setLinesearch(kw["linesearch"])
        res = lb.minimiseBatch(states)
        for c, s, r in zip(cs, states, res):
            o = OraclePF(c["grid"], 1, 1.0, [1.0], [c["w"]])
            xo, so = oracle_lbfgs(o, c["x0"], m=kw["m"], max_iter=kw.get("max_iter", 20000),
                                  max_step=kw.get("max_step", 0.0), init_step=kw.get("init_step", 1.0),
                                  linesearch=kw.get("linesearch", "backtracking"))
            e0 = o.energy(c["x0"])
            assert r["converged"] == so["converged"], (kw, c["grid"], r, so)
            if so["converged"]:
                assert energy_ok(r["energy"], so["energy"], e0), (kw, c["grid"])
            else:
                # a fixed iteration count without line search: the same trajectory to rounding
                assert r["iterations"] == so["iterations"]
                # (without a line search rounding differences grow along the trajectory)
                assert energy_ok(r["energy"], o.energy(xo), e0), (kw, c["grid"])


def test_batch_rejects_unsupported_fields(gpu):
    m = gpu
    ok = make_state(m, cases(3, sides=(8,), extra=())[0])
    tern = m.State(m.PhaseField().setNFluid(3).setGridSize([4, 4, 4]).setSurfaceTension([1, 1, 1]),
                   np.full(192, 1 / 3))
    with pytest.raises(ValueError, match="one-fluid PhaseField"):
        m.Lbfgs().minimiseBatch([ok, tern])
    with pytest.raises(ValueError, match="FP32 throughout"):
        m.Lbfgs().setPrecision("fp32").minimiseBatch([ok])
    with pytest.raises(ValueError, match="at most 10"):
        m.Lbfgs().setM(12).minimiseBatch([ok])
    with pytest.raises(ValueError, match="twice"):
        m.Lbfgs().minimiseBatch([ok, ok])
    assert m.Lbfgs().minimiseBatch([]) == []


def test_batch_through_the_c_abi(gpu):
    """minim_minimise_batch called directly: results may be NULL, and the
    minimised coordinates are left in each State's device coordinates."""
    import ctypes as C

    from minim_amd import _capi as A

    m = gpu
    cs = cases(13, sides=(10, 24), extra=((4, 5, 6),))
    states = [make_state(m, c) for c in cs]
    lb = m.Lbfgs().setM(5).setMaxIter(20000)
    arr = (C.c_void_p * len(states))(*[s.ptr for s in states])
    assert A.load().minim_minimise_batch(lb._handle().ptr, arr, len(states), None) == 0
    for c, s in zip(cs, states):
        o = OraclePF(c["grid"], 1, 1.0, [1.0], [c["w"]])
        g = o.gradient(s.coords())
        assert np.sqrt(np.mean(g * g)) < 1.0001 * o.convergence


def run_batch(mod, cs, control, **kw):
    states = [make_state(mod, c) for c in cs]
    lb = mod.Lbfgs().setM(kw.get("m", 5)).setMaxIter(kw.get("max_iter", 20000)).setBatchControl(control)
    if "budget" in kw:
        lb.setBatchBudget(kw["budget"])
    if "precision" in kw:
        lb.setPrecision(kw["precision"])
    if "max_step" in kw:
        lb.setMaxStep(kw["max_step"])
    if "linesearch" in kw:
        lb.setLinesearch(kw["linesearch"])
    res = lb.minimiseBatch(states)
    return states, res, lb.batchReport()


@pytest.mark.parametrize("kw", [dict(), dict(precision="mixed"), dict(m=3, max_step=0.05),
                                dict(linesearch="none", max_iter=80)])
def test_batch_controls_agree_bit_for_bit(gpu, kw):
    """Device control with the iteration replayed as a HIP graph, the same
    sequence launched without the graph, and host control: every instance's
    coordinates, energy and counts are the same bit for bit (each graph replay
    is checked through the state it leaves for the next)."""
    m = gpu
    cs = cases(21, sides=(9, 16, 33, 60), extra=((5, 6, 7), (30, 1, 1), (2, 11, 13)))
    out = {}
    for control in ("graph", "device", "host"):
        print("control", control, kw, flush=True)
        states, res, rep = run_batch(m, cs, control, **kw)
        out[control] = ([s.coords() for s in states], res, rep)
    xg, rg, repg = out["graph"]
    for control in ("device", "host"):
        x, r, _ = out[control]
        for i in range(len(cs)):
            assert np.array_equal(x[i], xg[i]), (control, i)
            for k in ("energy", "iterations", "converged", "n_energy", "n_gradient", "iter_fp32", "iter_fp64"):
                assert r[i][k] == rg[i][k], (control, i, k, r[i][k], rg[i][k])
    # one host synchronisation every eight iterations under device control:
    # the flags of every 8-iteration chunk are read while the next one runs,
    # except for the first chunk and each chunk launched before a new layout
    # (the mixed schedule launches the fresh phase's iterations outside the graph)
    assert 0 < repg["replays"] - 8 * repg["host_syncs"] <= 8 * (1 + repg["captures"])
    if kw.get("precision") == "mixed":
        assert 0 < repg["graph_launches"] < repg["replays"]
    else:
        assert repg["graph_launches"] == repg["replays"]
    assert out["host"][2]["host_syncs"] >= out["host"][2]["iterations"]


def test_batch_256_instances_vs_oracle(gpu):
    """256 instances of mixed shapes (sides log-uniform in [8, 64], some 3-D,
    some with a two-stripe bias so that their minimum has interfaces), in the
    graph-driven batch: every instance converged, the energy it reports is the
    oracle's energy at its coordinates (1e-12) and the oracle's gradient there
    meets the tolerance; on a sample, the converged energy is the oracle
    L-BFGS's within 1e-8 (no floor for the interface minima)."""
    m = gpu
    rng = np.random.default_rng(256)
    cs = []
    for i in range(256):
        if i % 16 == 15:
            g = tuple(int(v) for v in rng.integers(3, 9, 3))
        else:
            sd = int(round(np.exp(rng.uniform(np.log(8), np.log(64)))))
            g = (sd, sd, 1)
        n = int(np.prod(g))
        x0 = rng.uniform(-0.1, 0.1, n)
        if i % 4 == 1 and g[0] >= 16:  # two stripes along x: interface minimum
            xg = np.arange(g[0])
            x0 = x0 + np.repeat(np.where(np.sin(2 * np.pi * (xg + 0.5) / g[0]) > 0, 0.6, -0.6), g[1] * g[2])
        cs.append(dict(grid=g, w=float(rng.uniform(1.0, 2.0)), x0=x0))
    states, res, rep = run_batch(m, cs, "graph")
    print("report", rep, flush=True)
    assert 0 < rep["replays"] - 8 * rep["host_syncs"] <= 8 * (1 + rep["captures"])
    assert rep["captures"] >= 2  # the grid shrank as instances finished
    n_iface = 0
    for i, (c, s, r) in enumerate(zip(cs, states, res)):
        o = OraclePF(c["grid"], 1, 1.0, [1.0], [c["w"]])
        x = s.coords()
        e0 = o.energy(c["x0"])
        assert r["converged"], (i, c["grid"], r)
        assert abs(r["energy"] - o.energy(x)) <= 1e-12 * max(abs(r["energy"]), abs(e0)), i
        g = o.gradient(x)
        assert np.sqrt(np.mean(g * g)) < 1.0001 * o.convergence, i
        if r["energy"] > 1:
            n_iface += 1
        if i % 8 == 1 or i % 32 == 0:
            _, so = oracle_lbfgs(o, c["x0"], m=5, max_iter=20000)
            if so["energy"] > 1:  # an interface minimum: 1e-8 relative, no floor
                assert abs(r["energy"] - so["energy"]) <= 1e-8 * so["energy"], (i, r["energy"], so["energy"])
            else:
                assert energy_ok(r["energy"], so["energy"], e0), (i, r["energy"], so["energy"])
    assert n_iface >= 10  # minima with interfaces (E > 1) among the instances


def test_batch_grid_shrinks_as_instances_finish(gpu):
    """Instances that converge early leave the launch grid: under graph control
    the iteration is captured again over the instances still running (more
    than one capture), and every instance's result equals host control's bit
    for bit (an instance's partial sums do not depend on where its blocks lie
    in the grid)."""
    m = gpu
    rng = np.random.default_rng(31)
    cs = []
    for g in [(8, 8, 1)] * 4 + [(64, 64, 1)] * 2 + [(48, 40, 1)]:
        n = int(np.prod(g))
        x0 = rng.uniform(-0.1, 0.1, n)
        if g[0] == 64:  # two stripes: an interface minimum, many iterations
            xg = np.arange(g[0])
            x0 = x0 + np.repeat(np.where(np.sin(2 * np.pi * (xg + 0.5) / g[0]) > 0, 0.6, -0.6), g[1] * g[2])
        cs.append(dict(grid=g, w=float(rng.uniform(1.0, 2.0)), x0=x0))
    out = {}
    for control in ("graph", "host"):
        print("control", control, flush=True)
        states, res, rep = run_batch(m, cs, control)
        out[control] = ([s.coords() for s in states], res, rep)
    xg, rg, repg = out["graph"]
    xh, rh, reph = out["host"]
    print("reports", repg, reph, flush=True)
    assert repg["captures"] >= 2 and reph["captures"] == 0
    its = [r["iterations"] for r in rg]
    assert max(its) > 2 * min(its)
    for i in range(len(cs)):
        assert np.array_equal(xh[i], xg[i]), i
        for k in ("energy", "iterations", "converged", "n_energy", "n_gradient"):
            assert rh[i][k] == rg[i][k], (i, k, rh[i][k], rg[i][k])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_budget_keeps_results(gpu, precision):
    """A budget of running sites admits instances from both ends of the size
    order as others finish: every instance's coordinates, energy and counts
    equal host control's (all instances at once) bit for bit, under the graph
    and without it."""
    m = gpu
    cs = cases(12, sides=(6, 9, 12, 20, 33, 48, 64), extra=((5, 6, 7), (30, 1, 1), (2, 11, 13), (40, 40, 1)))
    kw = dict(precision=precision)
    out = {}
    for control, budget in (("host", 0), ("graph", 3000), ("device", 5000)):
        print("control", control, "budget", budget, precision, flush=True)
        states, res, rep = run_batch(m, cs, control, budget=budget, **kw)
        out[control] = ([s.coords() for s in states], res, rep)
    xh, rh, _ = out["host"]
    for control in ("graph", "device"):
        x, r, rep = out[control]
        print(control, rep, flush=True)
        for i in range(len(cs)):
            assert np.array_equal(x[i], xh[i]), (control, i)
            for k in ("energy", "iterations", "converged", "n_energy", "n_gradient", "iter_fp32", "iter_fp64"):
                assert r[i][k] == rh[i][k], (control, i, k, r[i][k], rh[i][k])
    assert out["graph"][2]["captures"] >= 3  # admissions re-lay the grid
    assert all(r["converged"] for r in rh)
    # admission times: 0 for host control (every instance from the start); under
    # a budget some instances wait, and each is admitted before it finishes
    assert all(r["start_seconds"] == 0 for r in rh)
    for control in ("graph", "device"):
        r = out[control][1]
        assert all(0 <= x["start_seconds"] <= x["seconds"] for x in r), control
        assert max(x["start_seconds"] for x in r) > 0, control



def test_batch_block_sizes(gpu):
    """The batch passes at 4, 8 and 16 elements per thread: every instance
    converges to the same minimum (the block size changes only how sums are
    split, so energies agree to rounding), and for one block size graph and
    host control agree bit for bit."""
    m = gpu
    cs = cases(17, sides=(9, 33, 60, 100), extra=((5, 6, 7), (40, 1, 1)))
    out = {}
    for block, control in ((4, "graph"), (8, "graph"), (16, "graph"), (16, "host")):
        print("block", block, control, flush=True)
        states = [make_state(m, c) for c in cs]
        lb = m.Lbfgs().setM(5).setMaxIter(20000).setBatchControl(control).setBatchBlock(block)
        res = lb.minimiseBatch(states)
        assert lb.batchReport()["block"] == block
        out[(block, control)] = ([s.coords() for s in states], res)
    x16g, r16g = out[(16, "graph")]
    x16h, r16h = out[(16, "host")]
    for i in range(len(cs)):
        assert np.array_equal(x16g[i], x16h[i]), i
        assert r16g[i]["iterations"] == r16h[i]["iterations"]
    for block in (4, 8):
        x, r = out[(block, "graph")]
        for i, c in enumerate(cs):
            assert r[i]["converged"] and r16g[i]["converged"]
            e0 = OraclePF(c["grid"], 1, 1.0, [1.0], [c["w"]]).energy(c["x0"])
            assert energy_ok(r[i]["energy"], r16g[i]["energy"], e0), (block, i)


def test_batch_autotune(gpu):
    """autotuneBatch times each block size on the first iterations without
    changing the States, and the next minimiseBatch of the same shape uses
    the size it chose."""
    m = gpu
    m.autotune_clear()
    cs = cases(19, sides=(16, 48, 96), extra=((4, 5, 6),))
    states = [make_state(m, c) for c in cs]
    before = [s.coords() for s in states]
    lb = m.Lbfgs().setM(5).setMaxIter(20000)
    t = lb.autotuneBatch(states, blocks=(4, 8, 16), replays=16)
    print("tuned", t, flush=True)
    assert sorted(t) == [4, 8, 16] and all(v > 0 for v in t.values())
    for b, s in zip(before, states):
        assert np.array_equal(b, s.coords())
    best = min(t, key=t.get)
    res = lb.minimiseBatch(states)
    assert lb.batchReport()["block"] == best
    assert all(r["converged"] for r in res)
    m.autotune_clear()


@pytest.mark.parametrize("kw", [dict(), dict(m=3, max_step=0.05), dict(m=7, init_step=0.3),
                                dict(linesearch="none", max_iter=60)])
def test_batch_resident_vs_oracle(gpu, kw):
    """Resident control (one workgroup runs each instance's whole minimisation
    in one launch): every instance ends as the oracle's FP64 L-BFGS under the
    same settings does -- converged energies within 1e-8, the reported energy
    the oracle's at the coordinates left (1e-12), the gradient within the
    tolerance -- with one launch and one host synchronisation for the batch.
    Without a line search, a fixed iteration count gives the graph-driven
    batch's trajectory to rounding."""
    m = gpu
    cs = cases(23)
    states, res, rep = run_batch(m, cs, "resident", **kw)
    print("report", rep, flush=True)
    assert rep["control"] == "resident" and rep["host_syncs"] == 1 and rep["graph_launches"] == 0
    assert rep["replays"] == rep["iterations"] >= max(r["iterations"] for r in res)
    if kw.get("linesearch") == "none":
        _, rg, _ = run_batch(m, cs, "graph", **kw)
    for i, (c, s, r) in enumerate(zip(cs, states, res)):
        o = OraclePF(c["grid"], 1, 1.0, [1.0], [c["w"]])
        _, so = oracle_lbfgs(o, c["x0"], m=kw.get("m", 5), max_iter=kw.get("max_iter", 20000),
                             max_step=kw.get("max_step", 0.0), init_step=kw.get("init_step", 1.0),
                             linesearch=kw.get("linesearch", "backtracking"))
        x = s.coords()
        e0 = o.energy(c["x0"])
        assert abs(r["energy"] - o.energy(x)) <= 1e-12 * max(abs(r["energy"]), abs(e0)), c["grid"]
        assert r["converged"] == so["converged"], (c["grid"], r, so)
        if so["converged"]:
            assert energy_ok(r["energy"], so["energy"], e0), (c["grid"], r["energy"], so["energy"])
            g = o.gradient(x)
            assert np.sqrt(np.mean(g * g)) < 1.0001 * o.convergence
            assert abs(r["iterations"] - so["iterations"]) <= max(10, 0.3 * so["iterations"]), (r, so)
        else:
            assert r["iterations"] == so["iterations"] == rg[i]["iterations"]
            # (without a line search rounding differences grow along the trajectory)
            assert energy_ok(r["energy"], rg[i]["energy"], e0), (c["grid"], r["energy"], rg[i]["energy"])
            assert r["iter_fp64"] == rg[i]["iter_fp64"] and r["n_energy"] == rg[i]["n_energy"]
        if r["converged"]:
            assert r["iter_fp64"] == r["iterations"] + 1
        assert r["iter_fp32"] == 0


def test_batch_resident_first_iteration_matches_single(gpu):
    """maxIter 0 under resident control: the first step (Lbfgs.cpp:91-96 with
    the line search) equals the single-State path's to the rounding of the
    norms."""
    m = gpu
    cs = cases(8)
    states, res, _ = run_batch(m, cs, "resident", max_iter=0)
    for c, s, r in zip(cs, states, res):
        s1 = make_state(m, c)
        x1 = m.Lbfgs().setM(5).setMaxIter(0).minimise(s1)
        assert np.max(np.abs(s.coords() - x1)) <= 1e-14 * max(1.0, np.max(np.abs(x1))), c["grid"]
        assert r["iterations"] == 1 and not r["converged"]
        assert abs(r["energy"] - s1.energy()) <= 1e-12 * abs(s1.energy())


def test_batch_resident_many_instances(gpu):
    """More instances than workgroups the device holds at once (1200, sides
    4..40, some 3-D): the later workgroups start as earlier ones finish, every
    instance converges, and its result does not depend on the others (an
    instance alone gives the same coordinates bit for bit: its sums are taken
    by its own workgroup in a fixed order).  Mixed precision is refused."""
    m = gpu
    rng = np.random.default_rng(1200)
    cs = []
    for i in range(1200):
        if i % 10 == 9:
            g = tuple(int(v) for v in rng.integers(2, 8, 3))
        else:
            sd = int(rng.integers(4, 41))
            g = (sd, sd, 1)
        cs.append(dict(grid=g, w=float(rng.uniform(1.0, 2.0)), x0=rng.uniform(-0.1, 0.1, int(np.prod(g)))))
    states, res, rep = run_batch(m, cs, "resident")
    print("report", rep, flush=True)
    for i in range(0, 1200, 37):
        c, s, r = cs[i], states[i], res[i]
        assert r["converged"], (i, r)
        o = OraclePF(c["grid"], 1, 1.0, [1.0], [c["w"]])
        g = o.gradient(s.coords())
        assert np.sqrt(np.mean(g * g)) < 1.0001 * o.convergence, i
    assert all(r["converged"] for r in res)
    for i in (0, 9, 555):
        alone, ra, _ = run_batch(m, [cs[i]], "resident")
        assert np.array_equal(alone[0].coords(), states[i].coords()), i
        assert ra[0]["iterations"] == res[i]["iterations"] and ra[0]["n_energy"] == res[i]["n_energy"]
    with pytest.raises(ValueError, match="FP64 schedule only"):
        run_batch(m, cs[:2], "resident", precision="mixed")


def test_batch_resident_autotune(gpu):
    """autotuneBatch under resident control times bounded runs (at most
    `replays` iterations each) and leaves the States as they were; the next
    resident minimiseBatch uses the size it chose and converges."""
    m = gpu
    m.autotune_clear()
    cs = cases(29, sides=(16, 40, 90), extra=((4, 5, 6),))
    states = [make_state(m, c) for c in cs]
    before = [s.coords() for s in states]
    lb = m.Lbfgs().setM(5).setMaxIter(20000).setBatchControl("resident")
    t = lb.autotuneBatch(states, blocks=(4, 8), replays=8)
    print("tuned", t, flush=True)
    assert sorted(t) == [4, 8] and all(v > 0 for v in t.values())
    for b, s in zip(before, states):
        assert np.array_equal(b, s.coords())
    res = lb.minimiseBatch(states)
    assert lb.batchReport()["block"] == min(t, key=t.get)
    assert all(r["converged"] for r in res)
    m.autotune_clear()
