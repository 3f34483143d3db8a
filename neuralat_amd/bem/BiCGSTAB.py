"""BiCGSTAB over a caller-supplied matvec, with the reference's semantics
(src/bem/BiCGSTAB.py:5-107): unconjugated (bilinear) rho, alpha, omega; a
fresh residual b - A x checked after every step against
||r||^2 < tol ||b||^2 or < atol; silent return on failure or rho == 0.

Dense matrices go to the library's device BiCGSTAB instead
(``neuralat_amd.bem.solver.bicgstab``); this class serves callables, whose
matvec is the caller's own code.
"""
import torch


class BiCGSTAB:
    def __init__(self, Ax_gen, device="cuda"):
        self.Ax_gen = Ax_gen
        self.device = device

    def init_params(self, b, x=None, nsteps=None, tol=1e-10, atol=1e-16):
        self.b = b.clone().detach()
        self.x = torch.zeros_like(b) if x is None else x
        self.residual_tol = tol * torch.vdot(self.b, self.b).real
        self.atol = atol
        self.nsteps = b.shape[0] if nsteps is None else nsteps
        self.status, self.r = self.check_convergence(self.x)
        one = torch.ones((), dtype=b.dtype, device=b.device)
        self.rho, self.alpha, self.omega = one.clone(), one.clone(), one.clone()
        self.v = torch.zeros_like(b)
        self.p = torch.zeros_like(b)
        self.r_hat = self.r.clone().detach()
        self.steps = 0

    def check_convergence(self, x):
        r = self.b - self.Ax_gen(x)
        rdotr = torch.vdot(r, r).real
        return bool(rdotr < self.residual_tol or rdotr < self.atol), r

    def step(self):
        rho = torch.dot(self.r, self.r_hat)
        beta = (rho / self.rho) * (self.alpha / self.omega)
        self.rho = rho
        self.p = self.r + beta * (self.p - self.omega * self.v)
        self.v = self.Ax_gen(self.p)
        self.alpha = self.rho / torch.dot(self.r_hat, self.v)
        s = self.r - self.alpha * self.v
        t = self.Ax_gen(s)
        self.omega = torch.dot(t, s) / torch.dot(t, t)
        self.x = self.x + self.alpha * self.p + self.omega * s
        self.steps += 1
        status, _ = self.check_convergence(self.x)
        if status:
            return True
        self.r = s - self.omega * t
        return False

    def solve(self, *args, **kwargs):
        self.init_params(*args, **kwargs)
        if self.status:
            return self.x
        while self.nsteps:
            if self.step():
                return self.x
            if self.rho == 0:
                break
            self.nsteps -= 1
        return self.x
