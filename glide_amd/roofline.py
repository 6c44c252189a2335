"""Algorithmic work per unit of the solver's kernels (used by bench.py).

FP64 flop counts are counted from the kernels' own source by
tests/test_flop_count.py (every +, -, *, / is one flop, fma two, min/max/abs/
copysign/floor none; exp and log are counted by the operations of their
implementation in detmath.h), for an interior cell, unconstrained mass row,
gl_derivatives off.
"""

# one local Newton iteration of one cell: 5x5 block assembly (including the
# centre cell's grounding sigmoid) + LU solve + update
FLOPS_NEWTON_ITER = 742
# viscosity of one cell (computed once per sweep)
FLOPS_ETA = 128
# per cell and sweep: grounding sigmoids of the 4 edge neighbours
FLOPS_CELL_SETUP = 164
NEWTON_ITERS = 10

PEAK_HBM_GBS = 8000.0        # MI355X HBM3E
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
PEAK_BF16_TFLOPS = 2500.0    # MI355X dense BF16 MFMA


def sweep_flops(n_cells: int, n_newton: int = NEWTON_ITERS) -> int:
    return n_cells * (FLOPS_ETA + FLOPS_CELL_SETUP + n_newton * FLOPS_NEWTON_ITER)


def residual_bytes(ny: int, nx: int) -> int:
    """Compulsory HBM bytes of one residual evaluation: read U and f (state
    vectors), bed, B, beta, mask, gamma (cells); write r (state vector)."""
    nu, nv, nc = ny * (nx + 1), (ny + 1) * nx, ny * nx
    nt = nu + nv + nc
    return 8 * (3 * nt + 5 * nc)


def linearisation_bytes(ny: int, nx: int) -> int:
    """Compulsory HBM bytes of one JVP or VJP: read U and the direction (or
    lambda) (state vectors), bed, B, beta, mask (cells); write the result
    (state vector)."""
    nu, nv, nc = ny * (nx + 1), (ny + 1) * nx, ny * nx
    nt = nu + nv + nc
    return 8 * (3 * nt + 4 * nc)
