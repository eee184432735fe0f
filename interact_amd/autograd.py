"""torch.autograd view of the cycle misfit, for gradient checks.

The misfit J of log slip rates at receiver patches along a replay of given
RK steps (rsfg_replay / rsfg_adjoint_gradient) is exposed as a function of
the per-patch friction parameters a - b (at fixed a, so b = a - (a - b)) and
D_c.  Its forward pass is the device replay, its backward the device
adjoint; torch only carries the values so that gradient checks can be
written with autograd.  No reference counterpart: the reference has no
sensitivities.
"""
import numpy as np
import torch

from . import lib as L


class MisfitProblem:
    """Everything of the misfit that is not a parameter: the fault system
    (assembled operator and loading), its start state, the fixed friction
    coefficients a, v0, f0, eta and law, the steps, receivers, target and
    weights."""

    def __init__(self, fs, X0, t0, a, steps, recv, target, weight, *, v0=1e-6, f0=0.6,
                 eta=None, law=1, method=L.RK5DP, prec=L.PREC_FP64, checkpoint_every=256):
        self.fs, self.X0, self.t0 = fs, np.asarray(X0, np.float64), float(t0)
        self.a = np.asarray(a, np.float64)
        self.steps = np.asarray(steps, np.float64)
        self.recv = np.asarray(recv, np.int32)
        self.target = np.asarray(target, np.float64)
        self.weight = np.asarray(weight, np.float64)
        self.fric = dict(v0=v0, f0=f0, eta=eta, law=law)
        self.method, self.prec, self.checkpoint_every = method, prec, checkpoint_every

    def _set(self, amb, dc):
        self.fs.set_friction(self.a, self.a - amb, dc, **self.fric)
        self.fs.set_state(self.X0, self.t0)

    def value(self, amb, dc):
        self._set(amb, dc)
        return self.fs.replay(self.steps, self.recv, self.target, self.weight,
                              method=self.method, prec=self.prec)["J"]

    def value_and_grad(self, amb, dc):
        """J, dJ/d(a - b) at fixed a (= -dJ/db), dJ/dD_c."""
        self._set(amb, dc)
        r = self.fs.adjoint_gradient(self.steps, self.recv, self.target, self.weight,
                                     method=self.method, prec=self.prec,
                                     checkpoint_every=self.checkpoint_every)
        return r["J"], -r["grad_b"], r["grad_dc"]


class CycleMisfit(torch.autograd.Function):
    """J(a - b, D_c) for CPU float64 tensors of per-patch values."""

    @staticmethod
    def forward(ctx, amb, dc, problem):
        a = amb.detach().cpu().numpy().astype(np.float64)
        d = dc.detach().cpu().numpy().astype(np.float64)
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            J, g_amb, g_dc = problem.value_and_grad(a, d)
            ctx.save_for_backward(torch.from_numpy(g_amb), torch.from_numpy(g_dc))
        else:
            J = problem.value(a, d)
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        g_amb, g_dc = ctx.saved_tensors
        return gJ * g_amb, gJ * g_dc, None


def cycle_misfit(amb, dc, problem):
    return CycleMisfit.apply(amb, dc, problem)
