"""Refinement levels on the device (uqk_refine, and the level moves inside
uqk_step / uqk_run_steps) against the oracle: the levels must agree cell by
cell, the fields to the FP64 bar."""
import os

import numpy as np
import pytest

from uqcreator_amd import mesh as meshes
from uqcreator_amd import native, solver, workloads

from test_gpu_sweeps import _oracle_closure, _oracle_mesh, _rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUT = os.path.join(ROOT, "uqcreator_amd", "input")


@pytest.mark.parametrize("kind,d,degrees", [(native.BOUNDED_BARRIER, 1, [2, 3, 4, 5]), (native.EULER2D, 2, [1, 2, 3]),
                                            (native.SHALLOW_WATER2D_BARRIER, 3, [1, 2, 4])])
def test_refine_pass_matches_oracle(oracle, kind, d, degrees):
    b = native.Basis(d, degrees, 3)
    ntl = b.ntotal_per_level
    cl = native.GpuClosure(kind, b)
    cl.set_refinement(ntl)
    rng = np.random.default_rng(11)
    n = 5000
    # indicators spread over both thresholds: moments of size 1e-4 .. 1e-1
    u = rng.normal(size=(n, b.M, cl.S)) * 10.0 ** rng.uniform(-4, -1, size=(n, 1, 1))
    nt = rng.choice(ntl, size=n).astype(np.int32)
    nt[::97] = ntl[0] + 1 if ntl[0] + 1 not in ntl else 0  # no level: left alone
    print("case: refine pass", kind, d, degrees, flush=True)
    du, dnt = native.DeviceArray.from_host(u), native.DeviceArray.from_host(nt)
    cl.refine(du, dnt)
    got = dnt.download()
    want = oracle.refine(u, d, ntl, nt)
    assert np.array_equal(got, want)
    moved = (want != nt).mean()
    assert 0.2 < moved < 0.9, moved
    # with a cap on the level (refinement retardation)
    cl.set_refinement_cap(1)
    dnt.upload(nt)
    cl.refine(du, dnt)
    want1 = oracle.refine(u, d, ntl, nt, max_level=1)
    assert np.array_equal(dnt.download(), want1)
    assert want1.max() == ntl[1]
    cl.set_refinement(None)


def _run_levels_both(oracle, kind, b, m, u0, lam0, prob_kw, clo_kw, steps, above, below, graph, keep=None):
    prob = native.problem(**prob_kw)
    okw = dict(prob_kw)
    oprob = oracle.problem(okw.pop("kind"), **okw)
    cl = native.GpuClosure(kind, b, **clo_kw)
    if keep is not None:
        if "kernel" in keep:
            cl.set_kernel(keep["kernel"])
        keep["closure"] = cl
    s = solver.GpuMomentSolver(cl, prob, m, levels=True, refine_above=above, coarsen_below=below)
    s.set_state(u0, lam0)
    if graph:
        s.run_device(1e3, steps)
    else:
        s.run(1e3, steps)
    ocl = {{"epsilon": "eps"}.get(k, k): v for k, v in clo_kw.items() if k != "precision"}
    c = _oracle_closure(oracle, kind, **ocl)
    uo, lo = u0.copy(), lam0.copy()
    nto = np.full(m.n_cells, b.M, np.int32)
    res = oracle.run_levels(c, oprob, _oracle_mesh(oracle, m), b.phi, b.phiw, b.xi, uo, lo, b.ntotal_per_level, nto,
                            1e3, steps, above=above, below=below)
    assert res["rc"] == 0 and s.steps == res["steps"] == steps
    nt = s.nt_cell()
    mv = s.mean_var()
    mv_o = oracle.mean_var(c, b.phi, b.phiw, lo, nt_cell=nto)
    return nt, nto, _rel_l2(mv, mv_o, m.area)


@pytest.mark.parametrize("graph", [False, True])
def test_burgers_levels_time_loop(oracle, graph):
    """1D Burgers step, levels of degree 2..5, 40 steps"""
    b = native.Basis(1, [2, 3, 4, 5], 12)
    m = meshes.mesh1d(400, 0.0, 3.0)
    u0 = workloads.project_ic(workloads.burgers_step_ic(), m.center, b)
    print("case: Burgers levels graph", graph, flush=True)
    nt, nto, err = _run_levels_both(oracle, native.BOUNDED_BARRIER, b, m, u0, np.zeros_like(u0),
                                    dict(kind=native.BURGERS, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.5),
                                    dict(u_minus=2.9, u_plus=12.1, epsilon=1e-10), 40, 0.02, 0.004, graph)
    print("levels", np.unique(nt, return_counts=True), "rel L2", err, flush=True)
    assert np.array_equal(nt, nto)
    assert len(np.unique(nt)) >= 2
    assert err.max() < 1e-8


def test_euler2d_levels_time_loop(oracle):
    """2D Euler past the profile (64 x 16 O-grid), levels of total degree
    1, 2, 3, thresholds 2e-3 / 4e-4 so that cells move both ways, 20 steps"""
    b = native.Basis(2, [1, 2, 3], 5)
    m = meshes.naca_ogrid(64, 16)
    u0 = workloads.project_ic(workloads.euler2d_freestream_ic(), m.center, b)
    c = oracle.closure(oracle.EULER2D)
    lam0 = np.zeros_like(u0)
    for j in range(m.n_cells):
        lam0[j, 0] = oracle.DS(c, u0[j, 0])
    print("case: Euler2D NACA levels", flush=True)
    nt, nto, err = _run_levels_both(oracle, native.EULER2D, b, m, u0, lam0,
                                    dict(kind=native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.4),
                                    dict(epsilon=1e-6), 20, 2e-3, 4e-4, True)
    print("levels", np.unique(nt, return_counts=True), "rel L2", err, flush=True)
    assert np.array_equal(nt, nto)
    assert len(np.unique(nt)) >= 2
    assert err.max() < 1e-8


def test_euler2d_levels_mixed_precision(oracle):
    """The same loop in mixed precision on the workgroup kernel (chosen: the
    automatic choice keeps the generic FP64 kernel below n = 64, where it is
    faster), whose FP32 MFMA Hessian and kept factors then serve per-cell
    orders at n = 4 nTotal <= 40: levels equal to the oracle's FP64 loop
    cell by cell, E/Var within the mixed bar 1e-6."""
    b = native.Basis(2, [1, 2, 3], 5)
    m = meshes.naca_ogrid(64, 16)
    u0 = workloads.project_ic(workloads.euler2d_freestream_ic(), m.center, b)
    c = oracle.closure(oracle.EULER2D)
    lam0 = np.zeros_like(u0)
    for j in range(m.n_cells):
        lam0[j, 0] = oracle.DS(c, u0[j, 0])
    print("case: Euler2D NACA levels, mixed precision", flush=True)
    keep = {"kernel": native.KERNEL_BLOCK}
    nt, nto, err = _run_levels_both(oracle, native.EULER2D, b, m, u0, lam0,
                                    dict(kind=native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.4),
                                    dict(epsilon=1e-6, precision=native.MIXED), 20, 2e-3, 4e-4, True, keep)
    print("levels", np.unique(nt, return_counts=True), "rel L2", err, flush=True)
    assert keep["closure"].last_kernel() == native.KERNEL_BLOCK
    assert np.array_equal(nt, nto)
    assert len(np.unique(nt)) >= 2
    assert err.max() < 1e-6


def test_adaptive_input_file(oracle, tmp_path):
    """The adaptive NACA input through uqk_run_config (levels from the
    'moments' array, every cell on the finest level at the start) on a
    64 x 16 grid, 10 steps, against the oracle's loop with levels."""
    text = open(os.path.join(INPUT, "euler2d_naca_adaptive.toml")).read()
    for a, bb in [("ni = 512", "ni = 64"), ("nj = 128", "nj = 16"), ("maxSteps = 2000", "maxSteps = 10")]:
        assert a in text
        text = text.replace(a, bb)
    path = tmp_path / "naca_adaptive.toml"
    path.write_text(text)
    csv = str(tmp_path / "mv.csv")
    print("case: adaptive input", flush=True)
    rep = native.run_config(str(path), csv)
    inf, u0, lam0, center = native.config_setup(str(path))
    b = native.Basis(2, [1, 2, 3], 5)
    assert (inf["M"], inf["Q"]) == (b.M, b.Q)
    m = meshes.naca_ogrid(64, 16)
    c = _oracle_closure(oracle, native.EULER2D, eps=1e-6)
    oprob = oracle.problem(oracle.P_EULER2D, flux=oracle.FLUX_LF, cfl=0.4)
    nto = np.full(m.n_cells, b.M, np.int32)
    res = oracle.run_levels(c, oprob, _oracle_mesh(oracle, m), b.phi, b.phiw, b.xi, u0, lam0, b.ntotal_per_level, nto,
                            1e3, 10, above=2e-3, below=4e-4)
    assert res["rc"] == 0 and rep["steps"] == res["steps"] == 10
    mv = np.loadtxt(csv, delimiter=",", ndmin=2)
    mv_o = oracle.mean_var(c, b.phi, b.phiw, lam0, nt_cell=nto)
    err = _rel_l2(mv, mv_o, m.area)
    print("report", rep, "levels", np.unique(nto, return_counts=True), "rel L2", err, flush=True)
    assert len(np.unique(nto)) >= 2
    assert err.max() < 1e-8


def test_adaptive_one_shot_retardation_input(oracle, tmp_path):
    """Adaptive One-Shot IPM with refinement retardation through
    uqk_run_config (steady branch): duals converged once at the start, one
    Newton step per cell and time step after that, cells start on level 0
    under a cap that rises to level 1 once the residual falls below 1.58e-5
    (after the first step) and would rise to 2 below 1e-9 (never here).  The
    oracle runs the same schedule step by step; without the cap some cells
    would reach level 2 within these steps."""
    text = open(os.path.join(INPUT, "euler2d_naca_adaptive.toml")).read()
    for a, bb in [("ni = 512", "ni = 64"), ("nj = 128", "nj = 16"), ("maxSteps = 2000", "maxSteps = 20\nresidual = 1e-12\n"
                  "retardationResiduals = [1.58e-5, 1e-9]"), ("maxIterations = 1000", "maxIterations = 1"),
                  ("refineAbove = 2e-3", "refineAbove = 1e-4"), ("coarsenBelow = 4e-4", "coarsenBelow = 2e-5")]:
        assert a in text
        text = text.replace(a, bb)
    path = tmp_path / "naca_oneshot.toml"
    path.write_text(text)
    csv = str(tmp_path / "mv.csv")
    print("case: adaptive one-shot with retardation", flush=True)
    rep = native.run_config(str(path), csv)
    inf, u, lam, center = native.config_setup(str(path))
    b = native.Basis(2, [1, 2, 3], 5)
    m = meshes.naca_ogrid(64, 16)
    ntl = b.ntotal_per_level
    c0 = _oracle_closure(oracle, native.EULER2D, eps=1e-6, max_iter=10000, safe=1)
    st, _ = oracle.solve_closure(c0, b.phi, b.phiw, u, lam)
    assert not st.any()
    c = _oracle_closure(oracle, native.EULER2D, eps=1e-6, max_iter=1)
    oprob = oracle.problem(oracle.P_EULER2D, flux=oracle.FLUX_LF, cfl=0.4)
    om = _oracle_mesh(oracle, m)
    nt = np.full(m.n_cells, ntl[0], np.int32)
    tol, rl, caps = [1.58e-5, 1e-9], 0, []
    for k in range(20):
        caps.append(min(rl, len(ntl) - 1))
        r = oracle.run_levels(c, oprob, om, b.phi, b.phiw, b.xi, u, lam, ntl, nt, 1e3, 1, above=1e-4, below=2e-5,
                              max_level=caps[-1])
        assert r["rc"] == 0
        while rl < len(tol) and r["residual"] < tol[rl]:
            rl += 1
    assert caps[0] == 0 and caps[1:] == [1] * 19
    assert rep["steps"] == 20
    assert rep["residual"] == pytest.approx(r["residual"], rel=1e-9)
    mv = np.loadtxt(csv, delimiter=",", ndmin=2)
    mv_o = oracle.mean_var(c, b.phi, b.phiw, lam, nt_cell=nt)
    err = _rel_l2(mv, mv_o, m.area)
    print("report", rep, "levels", np.unique(nt, return_counts=True), "rel L2", err, flush=True)
    assert set(np.unique(nt).tolist()) == {ntl[0], ntl[1]}
    assert err.max() < 1e-8
