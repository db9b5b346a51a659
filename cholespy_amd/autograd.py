"""Differentiable solves: x = A^{-1} b with gradients for b and for the
stored entries of A.

The backward pass reuses the solver's factor for the adjoint solve
lam = A^{-1} dL/dx (A is symmetric) and forms dL/d(entries) = -lam x'
restricted to the caller's stored pattern with the lower-triangle rule the
factorisation uses (only i >= j entries are read; an entry below the
diagonal stands for both (i, j) and (j, i); entries above it get zero).
On GPU solvers the scatter is the HIP kernel behind
`cholespy_entry_gradient` (include/cholespy_amd.h); CPU solvers compute the
same formula with torch operations on the host.

Usage:
    solver = CholeskySolverF(n, ii, jj, values, MatrixType.COO)
    rows, cols = coo_indices(n, ii, jj, MatrixType.COO)
    x = solve(solver, b, values, rows, cols)   # differentiable in b and values
"""
from __future__ import annotations

import torch

from . import _core

MatrixType = _core.MatrixType


def coo_indices(n_rows: int, ii, jj, type) -> tuple[torch.Tensor, torch.Tensor]:
    """Row and column index (int32, on ii's device) of every stored entry,
    for COO, CSR (ii = row pointer) or CSC (ii = column pointer) input."""
    ii = torch.as_tensor(ii)
    jj = torch.as_tensor(jj, device=ii.device)
    if type == MatrixType.COO:
        rows, cols = ii, jj
    else:
        counts = (ii[1:] - ii[:-1]).long()
        outer = torch.repeat_interleave(torch.arange(n_rows, device=ii.device), counts)
        rows, cols = (outer, jj) if type == MatrixType.CSR else (jj, outer)
    return rows.to(torch.int32).contiguous(), cols.to(torch.int32).contiguous()


def _entry_gradient_host(rows, cols, lam, x):
    i, j = rows.long(), cols.long()
    g = torch.zeros(i.shape[0], dtype=torch.float64)
    lam = lam.double().reshape(lam.shape[0], -1)
    x = x.double().reshape(x.shape[0], -1)
    diag = i == j
    low = i > j
    g[diag] = -(lam[i[diag]] * x[i[diag]]).sum(-1)
    g[low] = -((lam[i[low]] * x[j[low]]).sum(-1) + (lam[j[low]] * x[i[low]]).sum(-1))
    return g


# (data pointer, version) of index tensors already checked against n: a
# training loop passes the same rows / cols every step, and the check costs
# a device synchronisation.
_checked: dict = {}


def _check_indices(rows, cols, n):
    """IndexError unless every stored entry's row and column lie in [0, n).
    (The gradient kernel never reads through an out-of-range index either:
    such an entry's gradient is NaN.)"""
    if rows.numel() == 0:
        return
    key = (rows.data_ptr(), rows._version, cols.data_ptr(), cols._version, rows.numel(), n)
    if key in _checked:
        return
    lo, hi = torch.stack([rows.min(), cols.min(), rows.max(), cols.max()]).view(2, 2).tolist()
    if min(lo) < 0 or max(hi) >= n:
        raise IndexError(f"rows / cols hold an index outside [0, {n}): [{min(lo)}, {max(hi)}]")
    if len(_checked) > 64:
        _checked.clear()
    _checked[key] = True


class SolveFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, values, solver, rows, cols):
        b_c = b.detach().contiguous()
        x = torch.empty_like(b_c)
        solver.solve(b_c, x)
        ctx.solver = solver
        ctx.save_for_backward(rows, cols, x)
        ctx.values_meta = None if values is None else (values.dtype, values.device, values.shape)
        return x

    @staticmethod
    def backward(ctx, gx):
        rows, cols, x = ctx.saved_tensors
        x = x.detach()   # the saved output comes back attached to the graph
        lam = torch.empty_like(x)
        ctx.solver.solve(gx.detach().to(x.dtype).contiguous(), lam)
        gb = lam if ctx.needs_input_grad[0] else None
        gv = None
        if ctx.needs_input_grad[1]:
            if rows is None or cols is None:
                raise ValueError("gradients for the matrix entries need rows and cols (coo_indices)")
            dtype, device, shape = ctx.values_meta
            _check_indices(rows, cols, x.shape[0])
            if x.is_cuda:
                gv = torch.empty(shape, dtype=dtype, device=device)
                _core._entry_gradient(rows, cols, lam, x, gv)
            else:
                gv = _entry_gradient_host(rows, cols, lam, x).to(dtype=dtype, device=device).reshape(shape)
        return gb, gv, None, None, None


def solve(solver, b, values=None, rows=None, cols=None):
    """x = A^{-1} b through `solver`, differentiable with respect to b and,
    when given, the stored entries `values` the solver was built from (with
    their `rows` / `cols` from coo_indices)."""
    return SolveFunction.apply(b, values, solver, rows, cols)
