"""Differentiable sparse solve for PyTorch: x = A(kappa)^-1 f.

A(kappa) is a sparse matrix whose stored values are linear in a coefficient
vector kappa (e.g. per-face conductivities of a diffusion stencil):
    vals[j] = sum_e aw[e] * kappa[aidx[e]]   for e in aptr[j]..aptr[j+1].
forward  : assemble, ILU(0), device-resident ILU-PCG to tol (all on the GPU).
backward : the adjoint A^T lambda = dL/dx is solved with (LU)^-T-preconditioned
           CG, dL/dvals_j = -lambda_row x_col, and dL/dkappa_c =
           sum_e cw[e] dL/dvals[cidx[e]] over the coefficient map.  x is the
           forward's output (kept by autograd); no Krylov state is stored, and
           A(kappa) and its factor are rebuilt only if the problem was solved
           for other coefficients in between (gradient() without x recomputes
           the forward solve).
The ILU(0) symbolic analysis (diagonal, level sets, sweep layouts, transposed
plans) depends only on the pattern: it is built once per problem, and every
solve re-runs only the numeric factorisation of the current values.
"""
import numpy as np
import torch

from . import (assemble, gpu_sparse_matrix, grad_coef, grad_vals, make_ilu, solve_cg, solve_cg_t,
               stop_criteria)


def coef_to_entry_map(nnz, cptr, cnnz, ccoef):
    """Transposes a coefficient->entries map into entries->coefficients (ascending coefficient)."""
    cptr = np.asarray(cptr)
    counts = np.bincount(np.asarray(cnnz), minlength=nnz)
    aptr = np.zeros(nnz + 1, np.int32)
    aptr[1:] = np.cumsum(counts)
    owner = np.repeat(np.arange(len(cptr) - 1, dtype=np.int32), np.diff(cptr))
    order = np.argsort(np.asarray(cnnz), kind="stable")  # stable: ascending coefficient inside an entry
    return aptr, owner[order].astype(np.int32), np.asarray(ccoef, np.float64)[order]


class SparseLinearProblem:
    """Pattern, right-hand side and coefficient maps of x = A(kappa)^-1 f, resident on the device."""

    def __init__(self, engine, pntr, indx, cptr, cnnz, ccoef, f, rtol=1e-12, max_iter=100000):
        self.engine = engine
        nnz = len(indx)
        aptr, aidx, aw = coef_to_entry_map(nnz, cptr, cnnz, ccoef)
        ld = engine.load
        self.pntr, self.indx = ld(np.asarray(pntr, np.int32)), ld(np.asarray(indx, np.int32))
        self.cptr, self.cnnz, self.ccoef = ld(np.asarray(cptr, np.int32)), ld(np.asarray(cnnz, np.int32)), ld(ccoef)
        self.aptr, self.aidx, self.aw = ld(aptr), ld(aidx), ld(aw)
        self.f = ld(np.asarray(f, np.float64))
        self.n = len(pntr) - 1
        self.nnz = nnz
        self.vals = torch.empty(nnz, dtype=torch.float64, device=self.f.device)
        self.A = gpu_sparse_matrix(engine, self.n, self.n, self.pntr, self.indx, self.vals)
        self.rtol = rtol
        self.tol = rtol * float(torch.linalg.vector_norm(self.f))
        self.max_iter = max_iter
        self.last_iters = (0, 0)
        self.ilu = None  # analysis of the pattern, kept; refactored for each kappa
        self._kappa = None  # the coefficients vals and the factor currently hold

    def factor(self, vals):
        """ILU(0) of the current values: the pattern's analysis is built on the first call only."""
        if self.ilu is None:
            self.ilu = make_ilu(self.engine, self.pntr, self.indx, vals)
        else:
            self.ilu.factor(vals)
        return self.ilu

    def assemble(self, kappa):
        k = kappa.detach().contiguous()
        self._kappa = None  # vals and the factor are out of step until factor() has run
        assemble(self.engine, self.aptr, self.aidx, self.aw, k, self.vals)
        return self.vals

    def assemble_factor(self, kappa):
        """Assembles A(kappa) and refactors: (vals, ilu)."""
        self.factor(self.assemble(kappa))
        self._kappa = kappa.detach().contiguous().clone()
        return self.vals, self.ilu

    def state_for(self, kappa):
        """(vals, ilu) of A(kappa): rebuilt only if they hold other coefficients."""
        k = kappa.detach()
        if self._kappa is None or self._kappa.shape != k.shape or not torch.equal(self._kappa, k):
            return self.assemble_factor(kappa)
        return self.vals, self.ilu

    def forward_solve(self, kappa):
        """Returns (x, ilu) for A(kappa) x = f (always assembling and refactoring)."""
        vals, ilu = self.assemble_factor(kappa)
        x = self.engine.vector(self.n)
        it = solve_cg(self.engine, stop_criteria(self.tol, self.max_iter), self.pntr, self.indx, vals, ilu, self.f, x,
                      matrix=self.A)
        self.last_iters = (it, self.last_iters[1])
        return x, ilu

    def gradient(self, kappa, gx, x=None):
        """dL/dkappa for dL/dx = gx.  x: the forward solution for kappa when the
        caller kept it (else the forward solve is recomputed)."""
        g = gx.detach().contiguous().to(torch.float64)
        gnorm = float(torch.linalg.vector_norm(g))
        if gnorm == 0.0:  # the loss does not depend on x: lambda = 0, so is the gradient
            self.last_iters = (self.last_iters[0], 0)
            return torch.zeros(int(self.cptr.numel()) - 1, dtype=torch.float64, device=g.device)
        if x is None:
            x, ilu = self.forward_solve(kappa)
        else:
            x = x.detach()
            _, ilu = self.state_for(kappa)
        lam = self.engine.vector(self.n)
        tol = stop_criteria(self.rtol * gnorm, self.max_iter)
        it = solve_cg_t(self.engine, tol, self.A, self.vals, ilu, g, lam)
        self.last_iters = (self.last_iters[0], it)
        gvals = torch.empty(self.nnz, dtype=torch.float64, device=x.device)
        grad_vals(self.A, lam, x, gvals)
        gk = torch.empty(int(self.cptr.numel()) - 1, dtype=torch.float64, device=x.device)
        grad_coef(self.engine, self.cptr, self.cnnz, self.ccoef, gvals, gk)
        return gk


class SparseSolve(torch.autograd.Function):
    """x = A(kappa)^-1 f with the hand-written adjoint as backward."""

    @staticmethod
    def forward(ctx, kappa, problem):
        x, _ = problem.forward_solve(kappa)
        ctx.problem = problem
        ctx.save_for_backward(kappa, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        kappa, x = ctx.saved_tensors
        return ctx.problem.gradient(kappa, gx, x=x), None


def sparse_solve(kappa, problem):
    return SparseSolve.apply(kappa, problem)
