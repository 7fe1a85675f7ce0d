"""Implicit Lenard-Bernstein / Dougherty collision step (vlapy/core/collisions.py)
on the GPU.

The matrix makers return the three diagonals (a, b, c) like the reference;
the solvers take them back.  Both solver names ("batched_tridiagonal" and
"naive") run the same batched partition solver on the GPU; "naive" writes
the result into f in place, as the reference's per-x dense solve does.
"""

import numpy as np
import torch

from .. import ops
from .._lib import Plan
from . import _dev


def _matrix_maker(vax, nv, nx, nu, dt, dv, op, precision="fp64"):
    vax = np.asarray(vax, dtype=np.float64)
    dtype = _dev.f_dtype(precision)
    plan = Plan(nx, nv, precision, v=vax, dv=float(dv), nu=float(nu), dt=float(dt), op=op)

    def make_arrays_for_matrix(f_xv):
        fd = _dev.to_dev(f_xv, dtype)
        dev = fd.device
        a = torch.empty(nx, nv - 1, dtype=torch.float64, device=dev)
        b = torch.empty(nx, nv, dtype=torch.float64, device=dev)
        c = torch.empty(nx, nv - 1, dtype=torch.float64, device=dev)
        ops.collision_diagonals(plan, fd, a, b, c)
        return _dev.back(a, f_xv), _dev.back(b, f_xv), _dev.back(c, f_xv)

    return make_arrays_for_matrix


def get_philharmonic_matrix_maker(vax, nv, nx, nu, dt, dv, precision="fp64"):
    """Lenard-Bernstein diagonals (collisions.py:26-83)."""
    return _matrix_maker(vax, nv, nx, nu, dt, dv, "lb", precision)


def get_dougherty_matrix_maker(vax, nv, nx, nu, dt, dv, precision="fp64"):
    """Dougherty diagonals (collisions.py:86-160)."""
    return _matrix_maker(vax, nv, nx, nu, dt, dv, "dg", precision)


def _solver(nv, in_place, precision="fp64"):
    dtype = _dev.f_dtype(precision)
    plans = {}

    def solve(a, b, c, f):
        fd = _dev.to_dev(f, dtype)
        nx = fd.shape[0]
        if fd.shape[1] != nv:
            raise ValueError("f has %d velocity cells, solver was made for %d" % (fd.shape[1], nv))
        if nx not in plans:
            plans[nx] = Plan(nx, nv, precision)
        out = fd if (in_place and _dev.is_tensor(f) and fd.data_ptr() == f.data_ptr()) else torch.empty_like(fd)
        ops.tridiag_solve(plans[nx], _dev.to_dev(a, torch.float64), _dev.to_dev(b, torch.float64),
                          _dev.to_dev(c, torch.float64), fd, out)
        res = _dev.back(out, f)
        if in_place and not _dev.is_tensor(f):
            f[...] = res
            return f
        return res

    return solve


def get_naive_solver(nx, nv=None, precision="fp64"):
    """Per-x solve writing into f (collisions.py:163-219)."""
    state = {}

    def _solver_(a, b, c, f):
        n = f.shape[1]
        if n not in state:
            state[n] = _solver(n, True, precision)
        return state[n](a, b, c, f)

    return _solver_


def get_batched_tridiag_solver(nv, precision="fp64"):
    """Batched tridiagonal solve returning a new array (collisions.py:222-265)."""
    return _solver(nv, False, precision)


def get_matrix_solver(nx, nv, solver_name="batched_tridiagonal", precision="fp64"):
    """Name dispatch (collisions.py:268-289)."""
    if solver_name == "naive":
        return get_naive_solver(nx, nv, precision)
    if solver_name == "batched_tridiagonal":
        return get_batched_tridiag_solver(nv, precision)
    raise NotImplementedError("Matrix Solver: <" + str(solver_name) + "> has not yet been implemented")


def get_batched_array_maker(vax, nv, nx, nu, dt, dv, operator="lb", precision="fp64"):
    """Name dispatch of the operator (collisions.py:292-317)."""
    if operator == "lb":
        return get_philharmonic_matrix_maker(vax, nv, nx, nu, dt, dv, precision)
    if operator == "dg":
        return get_dougherty_matrix_maker(vax, nv, nx, nu, dt, dv, precision)
    raise NotImplementedError("Collision Operator: <" + str(operator) + "> has not yet been implemented")
