"""GPU time loop of the moment solver through the C ABI.

MomentSolver::Solve (momentsolver.cpp:79-162) per step: closure (warm start
from the previous step's duals) -> uQ = U(lambda phi) -> u = uQ phiw -> dt = min
ComputeDt -> t += dt -> ExplicitEuler::Advance -> residual of the state-0 mean.
All of a step runs in uqk_step on the device; the host only reads back dt,
the residual and the failed-cell count.
"""
import numpy as np

from . import native


class GpuMomentSolver:
    def __init__(self, closure, problem, mesh, nt_cell=None, levels=False, refine_above=0.02, coarsen_below=0.004):
        """levels=True turns on the refinement levels of the basis
        (momentsolver.cpp:104-117): every cell starts on the finest level
        unless nt_cell says otherwise, and the step moves it between levels."""
        self.cl = closure
        self.prob = problem
        self.mesh = mesh
        self.dmesh = native.DeviceMesh(mesh)
        n = mesh.n_cells
        S, M, Q = closure.S, closure.M, closure.Q
        self.u = native.DeviceArray((n, M, S))
        self.lam = native.DeviceArray((n, M, S))
        self.uq = native.DeviceArray((n, Q, S))
        self.u_scratch = native.DeviceArray((n, M, S))
        self.status = native.DeviceArray(n, np.int32)
        self.iters = native.DeviceArray(n, np.int32)
        if levels:
            ntl = closure.basis.ntotal_per_level
            closure.set_refinement(ntl, refine_above, coarsen_below)
            if nt_cell is None:
                nt_cell = np.full(n, ntl[-1], np.int32)
        else:
            closure.set_refinement(None)
        self.nt = None if nt_cell is None else native.DeviceArray.from_host(np.asarray(nt_cell, np.int32))
        self.t = 0.0
        self.steps = 0
        self.residual = None
        self.newton_steps = 0

    def set_state(self, u, lam):
        self.u.upload(u)
        self.lam.upload(lam)
        self.cl.invalidate_factors()  # the kept inverses belong to the old duals

    def step(self):
        nfail, dt, res = self.cl.step(self.prob, self.dmesh, self.u, self.lam, self.uq, self.u_scratch, self.status,
                                      self.iters, nt_cell=self.nt)
        if nfail:
            st = self.status.download()
            raise RuntimeError("closure failed in %d cells at step %d (status counts %s)"
                               % (nfail, self.steps, np.bincount(st, minlength=5).tolist()))
        self.t += dt
        self.steps += 1
        self.residual = res
        self.newton_steps += self.cl.last_closure_stats()[0]
        return dt, res

    def run(self, t_end, max_steps=-1, min_residual=-1.0):
        """momentsolver.cpp:79: while t < tEnd and residual > minResidual."""
        while self.t < t_end and (max_steps < 0 or self.steps < max_steps):
            self.step()
            if self.residual <= min_residual:
                break
        return self

    def run_device(self, t_end, max_steps=-1, use_graph=True):
        """The whole loop on the device (uqk_run_steps): the step is replayed
        from a captured hipGraph with no host synchronisation inside a step."""
        r = self.cl.run_steps(self.prob, self.dmesh, self.u, self.lam, self.uq, self.u_scratch, self.status,
                              self.iters, nt_cell=self.nt, t0=self.t, t_end=t_end, max_steps=max_steps,
                              use_graph=use_graph)
        if r["nfail"]:
            raise RuntimeError("closure failed in %d cells at step %d" % (r["nfail"], self.steps + r["steps"]))
        self.t = r["t"]
        self.steps += r["steps"]
        self.residual = r["residual"]
        self.newton_steps += r["newton_steps"]
        return self

    def nt_cell(self):
        """Per-cell nTotal (the refinement levels) as a host array, or None."""
        return None if self.nt is None else self.nt.download()

    def mean_var(self):
        mv = native.DeviceArray((self.mesh.n_cells, 2 * self.cl.S))
        self.cl.mean_var(self.lam, mv, nt_cell=self.nt)
        return mv.download()
