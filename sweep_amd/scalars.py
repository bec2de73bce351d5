"""Finite-difference coefficients for the stencils of the time-stepping kernels.

Restates ``src/sweep/scalars.py`` of the reference:

* :func:`fd_coefficients` follows ``scalars.py:51-67`` (Taylor matrix solved in
  float64, the ``n`` right-hand coefficients kept);
* :func:`second_derivative_coefficients_fp32` follows ``scalars.py:69-93``
  (float32 centred second-derivative stencil whose coefficients sum to zero
  exactly: the rounding residual of ``c0 + 2 sum(c)`` is folded into ``c[-1]``);
* :func:`staggered_grid_coes` follows ``scalars.py:25-49``.

The kernels take their coefficients from the tables built here (the C-ABI
passes them by value), so the host is the single place these numbers come
from.  :func:`static_laplace_constants` / :func:`static_gradient_constants`
give the float32 values the reference's fixed-order CUDA stencils hold as
``constexpr`` (``operators/laplace.cuh:157-341``, ``operators/gradient.cuh:56-78``).
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np


def fd_coefficients(derivative_order: int = 2, accuracy_order: int = 2) -> np.ndarray:
    """Right half (offsets 1..n) of the centred FD weights, float64."""
    n = accuracy_order // 2
    offsets = np.arange(-n, n + 1, dtype=np.float64)
    taylor = np.vstack([offsets ** p for p in range(2 * n + 1)])
    rhs = np.zeros(2 * n + 1)
    rhs[derivative_order] = math.factorial(derivative_order)
    weights = np.linalg.solve(taylor, rhs)
    return weights[-n:]


def second_derivative_coefficients_fp32(accuracy_order: int):
    """``(c0, c)`` in float32 with ``Fraction(c0) + 2*sum(Fraction(c)) == 0``."""
    c = fd_coefficients(2, accuracy_order)
    c0 = np.float32(-2 * c.sum())
    c = c.astype(np.float32)
    resid = Fraction(float(c0)) + 2 * sum(Fraction(float(v)) for v in c)
    if resid:
        want = Fraction(float(c[-1])) - resid / 2
        c[-1] = np.float32(float(want))
        if Fraction(float(c[-1])) != want:
            raise RuntimeError(
                f"order {accuracy_order}: the zero-sum residual is not representable in c[-1]")
    return c0, c


def staggered_grid_coes(M: int) -> np.ndarray:
    """Staggered first-derivative weights a_1..a_M (float32)."""
    out = np.zeros(M, dtype=np.float32)
    for m in range(1, M + 1):
        prod = 1.0
        for n in range(1, M + 1):
            if n != m:
                num = (2 * n - 1) ** 2
                prod *= np.abs(num / (num - (2 * m - 1) ** 2))
        out[m - 1] = ((-1) ** (m + 1) / (2 * m - 1)) * prod
    return out


def laplace_and_gradient_tables(M: int):
    """``(lap, grad)``, each ``(M+1,)`` float32, in the layout of the reference's
    ``_build_fd_coefficients`` (``propagator/_c.py:1134-1156``): ``lap[0] = -c0``,
    ``lap[1:] = c``; ``grad[0] = 0``, ``grad[1:]`` the centred first-derivative
    weights."""
    c0, c = second_derivative_coefficients_fp32(2 * M)
    lap = np.zeros(M + 1, dtype=np.float32)
    lap[1:] = c
    lap[0] = -c0
    grad = np.zeros(M + 1, dtype=np.float32)
    grad[1:] = fd_coefficients(1, 2 * M).astype(np.float32)
    return lap, grad


def static_laplace_constants(M: int):
    """float32 ``[c1..cM]`` of the fixed-order forward Laplacian, the last one
    closing the zero sum in float32 arithmetic (``laplace.cuh:157-162,275-278,
    337-341``).  ``None`` for M > 4 (runtime-order path uses the table)."""
    f = np.float32
    if M == 1:
        return np.array([1.0], dtype=np.float32)
    if M == 2:
        c0, c1 = f(5.0) / f(2.0), f(4.0) / f(3.0)
        return np.array([c1, f(0.5) * c0 - c1], dtype=np.float32)
    if M == 3:
        c0, c1, c2 = f(49.0) / f(18.0), f(3.0) / f(2.0), f(-3.0) / f(20.0)
        return np.array([c1, c2, (f(0.5) * c0 - c1) - c2], dtype=np.float32)
    if M == 4:
        c0, c1, c2, c3 = f(205.0) / f(72.0), f(8.0) / f(5.0), f(-1.0) / f(5.0), f(8.0) / f(315.0)
        return np.array([c1, c2, c3, ((f(0.5) * c0 - c1) - c2) - c3], dtype=np.float32)
    return None


def static_laplace_c0(M: int):
    """float32 centre weight of the order-4 summed forward stencil."""
    return np.float32(np.float32(5.0) / np.float32(2.0)) if M == 2 else None


def static_gradient_constants(M: int):
    """float32 centred first-derivative weights the fixed-order stencils hold
    (``gradient.cuh:56-78``); order 2 is ``1/2`` applied as ``/(2h)``."""
    f = np.float32
    if M == 1:
        return np.array([0.5], dtype=np.float32)
    if M == 2:
        return np.array([f(8.0) / f(12.0), f(-1.0) / f(12.0)], dtype=np.float32)
    if M == 3:
        return np.array([f(3.0) / f(4.0), f(-3.0) / f(20.0), f(1.0) / f(60.0)], dtype=np.float32)
    if M == 4:
        return np.array([f(4.0) / f(5.0), f(-1.0) / f(5.0), f(4.0) / f(105.0), f(-1.0) / f(280.0)],
                        dtype=np.float32)
    return None
