"""Gradients of a velocity misfit with respect to log viscosity, by the
adjoint of the converged 2D PT iteration (no reference counterpart: the
reference scripts have no adjoint).

The forward model is the 2D multiply-form iteration of
scripts/Stokes2D_ve_perf.jl:47-59 (viscous: without the elastic terms), with
the free-slip copies of :77-78, driven to convergence on the device; the
visco-elastic model runs ``nt`` physical steps (assign_τ! then the PT loop,
Stokes2D_ve.jl:149-170) and stores each step's converged stresses as a
checkpoint for the backward pass.  The objective is the misfit of the
tangential velocity (after the last step) at ``npts`` seeded points of the
top boundary row (Vx[I, ny-1], the free-slip copy of the top interior row):

    J(ln μ) = ½ Σ_k (Vx[I_k, ny-1] − obs_k)²

``Inversion.value_and_grad`` solves the forward problem, sets ∂J/∂Vx, runs the
adjoint PT solve (libpts ``pts_adjoint_solve``, or ``pts_adjoint_ve_solve``
backward over the checkpointed steps) and returns ∂J/∂ln μ at the cell
centres (``pts_grad_logmu``).  ``rrule`` packages the pair the way a
ChainRulesCore rule does: the value and a pullback.
"""
import numpy as np

from .models import StokesModel


def observation_points(nx, npts=256, seed=7):
    """Seeded Vx column indices I_k in [1, nx-1] (the non-Dirichlet faces);
    drawn without replacement when there are at least npts of them."""
    rng = np.random.default_rng(seed)
    cols = np.arange(1, nx)
    return np.sort(rng.choice(cols, size=npts, replace=cols.size < npts))


class Inversion:
    """Forward solve, misfit and adjoint gradient for one 2D grid (viscous,
    or visco-elastic over the preset's nt physical steps)."""

    def __init__(self, n=255, npts=256, obs_seed=7, eps=1e-10, adj_eps=1e-10, nout=200, iter_max=10 ** 7,
                 preset="Stokes2D_grf", device=0, **overrides):
        self.model = StokesModel(preset, n=n, device=device, **overrides)
        p = self.model.p
        if p["ndim"] != 2:
            raise ValueError("the adjoint is implemented for 2D presets")
        self.model.setup()
        self.ctx = self.model.ctx
        self.params = self.model.pt_params()
        self.nx, self.ny = p["nx"], p["ny"]
        self.points = observation_points(self.nx, npts, obs_seed)
        self.obs = np.zeros(self.points.size)
        self.eps, self.adj_eps, self.nout, self.iter_max = eps, adj_eps, nout, iter_max
        self.last = {}

    # -- viscosity --------------------------------------------------------
    def grf_logmu(self, seed):
        """ln μ of the preset's setup (GRF, smoothing) for another seed."""
        p = self.model.p
        saved = p["grf_seed"]
        p["grf_seed"] = seed
        try:
            self.model.setup()
        finally:
            p["grf_seed"] = saved
        return np.log(self.ctx.download("Mus"))

    def set_logmu(self, logmu):
        """Reset the state to the initial conditions and set μ = exp(ln μ);
        Musτ and the iteration parameters follow as in the setup."""
        p, c, m = self.model.p, self.ctx, self.model
        logmu = np.asarray(logmu, dtype=np.float64)
        if logmu.shape != (self.nx, self.ny):
            raise ValueError(f"log mu shape {logmu.shape} != {(self.nx, self.ny)}")
        c.init_inclusion(p["lx"], p["ly"], p["lz"], p["mus0"], p["musi"], p["rad2"], p["ebg"], p["rhog_in"],
                         p["rhog_out"], p["ic_igg"])
        c.upload("Mus", np.exp(logmu))
        c.copy("Mustau", "Mus")
        c.maxloc("Mustau", "Mus")
        c.bc_x("Mustau")
        c.bc_y("Mustau")
        c.iter_params(m.Vpdtau, p["G"], p["dt"], p["Re"], p["r"], m.max_l)

    # -- forward ----------------------------------------------------------
    def solve_forward(self):
        p = self.model.p
        if not p["ve"]:
            res = self.ctx.pt_solve(self.params, self.eps, self.nout, self.iter_max, p["norm_kind"])
            if not res["converged"]:
                raise RuntimeError(f"forward PT solve did not converge: {res}")
            self.last["forward"] = res
            return res
        iters = 0
        for n in range(p["nt"]):
            self.ctx.assign_tau()
            res = self.ctx.pt_solve(self.params, self.eps, self.nout, self.iter_max, p["norm_kind"])
            if not res["converged"]:
                raise RuntimeError(f"forward PT solve of step {n} did not converge: {res}")
            iters += res["iters"]
            self.ctx.checkpoint(n)
        self.last["forward"] = dict(iters=iters, steps=p["nt"], err=res["err"], converged=True)
        return self.last["forward"]

    def top_velocity(self):
        return self.ctx.download("Vx")[self.points, self.ny - 1]

    def observe(self, logmu_true):
        """Synthetic observations: the forward velocities for a hidden ln μ."""
        self.set_logmu(logmu_true)
        self.solve_forward()
        self.obs = self.top_velocity().copy()
        return self.obs

    def value(self, logmu):
        self.set_logmu(logmu)
        self.solve_forward()
        r = self.top_velocity() - self.obs
        return 0.5 * float(np.dot(r, r))

    # -- adjoint ----------------------------------------------------------
    def dJ_dVx(self):
        r = self.top_velocity() - self.obs
        g = np.zeros((self.nx + 1, self.ny))
        np.add.at(g[:, self.ny - 1], self.points, r)
        return g, 0.5 * float(np.dot(r, r))

    def gradient(self):
        """∂J/∂ln μ at the converged forward state now on the device (for the
        visco-elastic model: through all checkpointed steps)."""
        g, _ = self.dJ_dVx()
        self.ctx.adjoint_set_source(g, None)
        if self.model.p["ve"]:
            grad, res = self.ctx.adjoint_ve_solve(self.params, self.model.p["nt"], self.adj_eps, self.nout,
                                                  self.iter_max)
        else:
            res = self.ctx.adjoint_solve(self.params, self.adj_eps, self.nout, self.iter_max)
            grad = None
        if not res["converged"]:
            raise RuntimeError(f"adjoint PT solve did not converge: {res}")
        self.last["adjoint"] = res
        return grad if grad is not None else self.ctx.grad_logmu(self.params)

    def value_and_grad(self, logmu):
        self.set_logmu(logmu)
        self.solve_forward()
        _, J = self.dJ_dVx()
        return J, self.gradient()

    def close(self):
        self.model.close()


def rrule(inv, logmu):
    """(J, pullback) for J = inv.value(ln μ), as a ChainRulesCore rrule
    returns them: pullback(ΔJ) = ΔJ · ∂J/∂ln μ."""
    J, grad = inv.value_and_grad(logmu)

    def pullback(dJ):
        return dJ * grad

    return J, pullback
