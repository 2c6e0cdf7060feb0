"""Finite-difference weights (Fornberg 1988) and the derivative-operator tables built on them.

Host-side setup of the stencil IR.  The arithmetic follows the reference loop for loop so
that the weights agree to the last bit:

* ``calculate_weights``  — src/discretization/schemes/fornberg_calculate_weights.jl:20-78
  (including the sum-to-zero correction on the middle entry, :64).
* ``centered_difference`` — CompleteCenteredDifference,
  schemes/centered_difference/centered_diff_weights.jl:4-75.
* ``upwind_difference`` — CompleteUpwindDifference, schemes/upwind_difference/upwind_diff_weights.jl:4-91.
* ``centered_difference_nonuniform`` — CompleteCenteredDifference on nodes,
  centered_diff_weights.jl:77-153.
* ``upwind_difference_nonuniform`` — CompleteUpwindDifference on nodes, upwind_diff_weights.jl:93-171.
* ``half_centered_difference`` — CompleteHalfCenteredDifference, schemes/half_offset_weights.jl:4-70.
* ``half_centered_difference_nonuniform`` — the same on nodes, half_offset_weights.jl:72-150.
* ``boundary_interp_extrap`` — BoundaryInterpolatorExtrapolator, schemes/extrapolation_weights.jl:1-72.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np


def calculate_weights(order: int, x0: float, x) -> np.ndarray:
    """Weights approximating the ``order``-th derivative at ``x0`` from samples at ``x``.

    Fornberg (1988), Math. Comp. 51, 699-706, restated from
    src/discretization/schemes/fornberg_calculate_weights.jl:20-78.  For ``order != 0``
    the entry ``N//2`` (0-based) is corrected so that the weights sum to exactly zero
    (fornberg_calculate_weights.jl:64).
    """
    x = [float(v) for v in x]
    x0 = float(x0)
    N = len(x)
    if not order < N:
        raise ValueError("Not enough points for the requested order.")
    M = order
    c1 = 1.0
    c4 = x[0] - x0
    C = [[0.0] * (M + 1) for _ in range(N)]
    C[0][0] = 1.0
    for i in range(1, N):
        mn = min(i, M)
        c2 = 1.0
        c5 = c4
        c4 = x[i] - x0
        for j in range(0, i):
            c3 = x[i] - x[j]
            c2 *= c3
            if j == i - 1:
                for s in range(mn, 0, -1):
                    C[i][s] = c1 * (s * C[i - 1][s - 1] - c5 * C[i - 1][s]) / c2
                C[i][0] = -c1 * c5 * C[i - 1][0] / c2
            for s in range(mn, 0, -1):
                C[j][s] = (c4 * C[j][s] - s * C[j][s - 1]) / c3
            C[j][0] = c4 * C[j][0] / c3
        c1 = c2
    w = [C[k][M] for k in range(N)]
    if order != 0:
        # Julia's sum() is pairwise for arrays; for the short stencils used here (N <= 16)
        # it reduces left to right in a single block.
        s = 0.0
        for v in w:
            s += v
        w[N // 2] -= s
    return np.array(w, dtype=np.float64)


@dataclass
class DerivativeOperator:
    """Weights of one derivative operator along one axis (weights already scaled by 1/dx^d).

    Mirrors the fields of DerivativeOperator (src/discretization/derivative_operator.jl:1-15)
    that the right-hand side uses.
    """
    derivative_order: int
    approximation_order: int
    dx: float
    stencil_length: int
    stencil_coefs: np.ndarray
    boundary_stencil_length: int
    boundary_point_count: int
    low_boundary_coefs: List[np.ndarray] = field(default_factory=list)
    high_boundary_coefs: List[np.ndarray] = field(default_factory=list)
    offside: int = 0
    kind: str = "centered"      # centered | upwind | half | extrap
    tap_offset: int = 0          # interior taps at i + tap_offset + k
    nlow: int = 0                # low rows apply for i <= nlow
    nhigh: int = 0               # high rows apply for i > lref - nhigh
    hoff: int = 1                # high-row index: r = lref - i + hoff
    pn_lo: int = 0               # > 0: per-node interior rows, stencil_coefs[i - pn_lo] for node i
    pn_hi: int = 0               #      in pn_lo..pn_hi (nonuniform grids)


def _scale(w: np.ndarray, dx: float, d: int) -> np.ndarray:
    # (1 / dx^d) * w, evaluated as Julia does: the scalar 1/dx^d first, then the product.
    return (1.0 / (dx ** d)) * w if d > 0 else (1.0 / (dx ** 0)) * w


def centered_difference(d: int, p: int, dx: float) -> DerivativeOperator:
    """CompleteCenteredDifference(d, p, dx), uniform grid (centered_diff_weights.jl:4-75)."""
    if p <= 1:
        raise ValueError("approximation_order must be greater than 1.")
    L = d + p - 1 + (d + p) % 2
    BL = d + p
    half = L // 2
    dummy_x = list(range(-half, half + 1))
    left_boundary_x = list(range(0, BL))
    bpc = half
    stencil = (1.0 / dx ** d) * calculate_weights(d, 0.0, dummy_x)
    low = [(1.0 / dx ** d) * calculate_weights(d, 1.0 * x0, left_boundary_x)
           for x0 in left_boundary_x[:half]]
    high = [v[::-1] * ((-1.0) ** d) for v in low]
    return DerivativeOperator(d, p, dx, L, stencil, BL, bpc, low, high, 0, "centered",
                              tap_offset=-half, nlow=bpc, nhigh=bpc, hoff=1)


def upwind_difference(d: int, p: int, dx: float, offside: int = 0) -> DerivativeOperator:
    """CompleteUpwindDifference(d, p, dx, offside), uniform grid (upwind_diff_weights.jl:4-91)."""
    if offside < 0:
        raise ValueError("Number of offside points should be non-negative")
    L = d + p
    BL = d + p
    low_count = offside
    high_count = L - 1 - offside
    dummy_x = [0.0 - offside + k for k in range(L)]
    stencil = (1.0 / dx ** d) * calculate_weights(d, 0.0, dummy_x)
    low_x = [float(k) for k in range(BL)]
    low = [(1.0 / dx ** d) * calculate_weights(d, 1.0 * x0, low_x)
           for x0 in [float(k) for k in range(low_count)]]
    high_x = [-float(k) for k in range(BL)]
    _high = [((-1.0 / dx) ** d) * calculate_weights(d, 1.0 * x0, high_x)
             for x0 in [-float(k) for k in range(high_count)]]
    high = list(reversed(_high))
    # Row selection of _upwind_difference (upwind_difference.jl:1-28): the positive (backward)
    # operator uses low rows for i <= offside, the negative (forward) one high rows for
    # i > n - boundary_point_count with boundary_point_count = high_count.
    return DerivativeOperator(d, p, dx, L, stencil, BL, high_count, low, high, offside,
                              "upwind", tap_offset=-offside, nlow=offside, nhigh=high_count,
                              hoff=1)


def centered_difference_nonuniform(d: int, p: int, x) -> DerivativeOperator:
    """CompleteCenteredDifference(d, p, x) on the nodes x (centered_diff_weights.jl:77-153).

    Interior node i (bpc < i <= n - bpc) takes Fornberg weights at x[i] on x[i-half .. i+half];
    nodes 1..bpc take weights on the first BL nodes and nodes n-bpc+1..n on the last BL, with
    those nodes measured from x[1] (resp. x[n-BL]) by cumulative sums of the spacings, as the
    reference forms them.  High row r (1-based) belongs to node n - r + 1.
    """
    if p <= 1:
        raise ValueError("approximation_order must be greater than 1.")
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    L = d + p - 1 + (d + p) % 2
    BL = d + p
    half = L // 2
    if n < max(L, BL):
        raise ValueError("the grid has fewer nodes than the stencil")
    dx = x[1:] - x[:-1]
    low_x = np.concatenate([[0.0], np.cumsum(dx[:BL - 1])])
    high_x = np.cumsum(dx[len(dx) - BL:])
    rows = np.array([calculate_weights(d, x[i - 1], x[i - 1 - half:i + half]) for i in range(half + 1, n - half + 1)])
    low = [calculate_weights(d, low_x[r], low_x) for r in range(half)]
    high = [calculate_weights(d, high_x[BL - 1 - r], high_x) for r in range(half)]
    return DerivativeOperator(d, p, float("nan"), L, rows, BL, half, low, high, 0, "centered",
                              tap_offset=-half, nlow=half, nhigh=half, hoff=1, pn_lo=half + 1, pn_hi=n - half)


def upwind_difference_nonuniform(d: int, p: int, x, offside: int = 0) -> DerivativeOperator:
    """CompleteUpwindDifference(d, p, x, offside) on the nodes x (upwind_diff_weights.jl:93-171),
    with the row selection of _upwind_difference on nonuniform grids (upwind_difference.jl:131-162).

    The reference builds rows for nodes i = offside+1 .. n-hc (hc = L-1-offside), Fornberg
    weights at x[i] on x[i-offside .. i+L-1-offside], then stores the operator with offside 0;
    the selection reads row k = II for node II.  So the forward operator (offside 0) gives node
    II its own row, and the backward one (offside L-1) gives node II the row built for node
    II + L - 1, applied to the taps II-L+1 .. II.  Both are reproduced: per-node row of node
    II = the reference's row k = II.  High rows of the forward operator: node n - r + 1 takes
    weights at x[n-hc+r] on the last BL nodes (r = 1..hc, in the reference's order).
    """
    if offside < 0:
        raise ValueError("Number of offside points should be non-negative")
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    L = d + p
    BL = d + p
    if offside > L - 1:
        raise ValueError("Number of offside points should be less than or equal to the stencil length")
    lc, hc = offside, L - 1 - offside
    built = [calculate_weights(d, x[i - 1], x[i - 1 - offside:i - 1 + L - offside]) for i in range(lc + 1, n - hc + 1)]
    high = [calculate_weights(d, x0, x[n - BL:]) for x0 in x[n - hc:]]
    if offside == 0:
        # forward: node i in 1..n-hc, taps i .. i+L-1; high rows for i > n - hc
        return DerivativeOperator(d, p, float("nan"), L, np.array(built), BL, hc, [], high, 0, "upwind",
                                  tap_offset=0, nlow=0, nhigh=hc, hoff=1, pn_lo=1, pn_hi=n - hc)
    if offside != L - 1:
        raise NotImplementedError("nonuniform upwind operators are built with offside 0 or L-1")
    # backward: node II in L..n-L+1 (taps II-L+1 >= 1, reference row II exists)
    lo_node, hi_node = L, n - L + 1
    if hi_node < lo_node:
        raise ValueError("the grid has too few nodes for the upwind stencil")
    rows = np.array(built[lo_node - 1:hi_node])
    return DerivativeOperator(d, p, float("nan"), L, rows, BL, 0, [], [], 0, "upwind",
                              tap_offset=-(L - 1), nlow=0, nhigh=0, hoff=1, pn_lo=lo_node, pn_hi=hi_node)


def half_centered_difference(d: int, p: int, dx: float) -> DerivativeOperator:
    """CompleteHalfCenteredDifference(d, p, dx), uniform grid (half_offset_weights.jl:4-70).

    Index ``i`` of this operator is a half point between nodes i and i+1.
    """
    if p <= 1:
        raise ValueError("approximation_order must be greater than 1.")
    L = p + 2 * (d // 2) + (p % 2)
    BL = d + p
    endpoint = L // 2
    dummy_x = list(range(1 - endpoint, endpoint + 1))
    left_boundary_x = list(range(1, BL + 1))
    bpc = L // 2
    xoffset = [1.5 + k for k in range(bpc)]
    stencil = (1.0 / dx ** d) * calculate_weights(d, 0.5, dummy_x)
    low = [(1.0 / dx ** d) * calculate_weights(d, off, left_boundary_x) for off in xoffset[:L // 2]]
    high = [v[::-1] * ((-1.0) ** d) for v in low]
    return DerivativeOperator(d, p, dx, L, stencil, BL, bpc, low, high, 0, "half",
                              tap_offset=1 - L // 2, nlow=bpc, nhigh=bpc, hoff=0)


def half_centered_difference_nonuniform(d: int, p: int, x) -> DerivativeOperator:
    """CompleteHalfCenteredDifference(d, p, x) on the nodes x (half_offset_weights.jl:72-150).

    Index i is the half point hx[i] = (x[i] + x[i+1]) / 2.  Half points bpc+1 .. n-bpc take
    Fornberg weights at hx[i] on x[i-L/2+1 .. i+L/2] (row selection
    half_offset_centred_difference.jl:40-69 reads row i - bpc for index i); the first bpc half
    points take weights at hx[1..bpc] on the first BL nodes; high row r (1-based, index
    len - r) takes weights at hx[len - r] on the last BL nodes, in the reference's order.
    """
    if p <= 1:
        raise ValueError("approximation_order must be greater than 1.")
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    L = p + 2 * (d // 2) + (p % 2)
    BL = d + p
    ep = L // 2
    if n < max(L, BL) + 1:
        raise ValueError("the grid has fewer nodes than the stencil")
    hx = (x[:-1] + x[1:]) / 2
    rows = np.array([calculate_weights(d, hx[i - 1], x[i - ep:i + ep]) for i in range(ep + 1, n - ep + 1)])
    low = [calculate_weights(d, hx[k], x[:BL]) for k in range(ep)]
    high = [calculate_weights(d, hx[len(hx) - 1 - k], x[n - BL:]) for k in range(ep)]
    return DerivativeOperator(d, p, float("nan"), L, rows, BL, ep, low, high, 0, "half",
                              tap_offset=1 - ep, nlow=ep, nhigh=ep, hoff=0, pn_lo=ep + 1, pn_hi=n - ep)


def boundary_interp_extrap(p: int, dx: float) -> DerivativeOperator:
    """BoundaryInterpolatorExtrapolator(p, dx) (extrapolation_weights.jl:1-72): 0th-order
    Lagrange extrapolation excluding the node itself (weight 0 inserted at its position)."""
    if p <= 1:
        raise ValueError("approximation_order must be greater than 1.")
    L = p - 1 + p % 2
    BL = p
    half = L // 2
    dummy_x = list(range(-half, half + 1))
    left_boundary_x = list(range(0, BL))
    bpc = half
    inner = [v for v in dummy_x if v != 0]
    w = list(calculate_weights(0, 0.0, inner))
    w.insert(dummy_x.index(0), 0.0)
    stencil = np.array(w, dtype=np.float64)
    low = []
    for i, x0 in enumerate(left_boundary_x[:bpc]):
        pts = [v for v in left_boundary_x if v != x0]
        ww = list(calculate_weights(0, 1.0 * x0, pts))
        ww.insert(i, 0.0)
        low.append(np.array(ww, dtype=np.float64))
    high = [v[::-1].copy() for v in low]
    return DerivativeOperator(0, p, dx, L, stencil, BL, bpc, low, high, 0, "extrap",
                              tap_offset=-half, nlow=bpc, nhigh=bpc, hoff=1)
