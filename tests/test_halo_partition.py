"""Index scheme of the folded-apply small-level sweep (k_vanka_halo,
glide_amd/csrc/tiles_device.h vanka_halo_tile), checked on the host.

A workgroup of 64 x 4 threads at (bx, by) runs the cells
[i0-1, i0+2] x [j0-1, j0+62] with (i0, j0) = (2 by, 62 bx) and writes the
points it owns.  The sweep writes the next state into another buffer, so
every point of the level must be owned by exactly one workgroup, and every
per-cell update a point reads must come from a cell that workgroup ran, inside
the level.  The GPU tests check the bits; this checks the partition over many
level shapes, including the ones the GPU tests do not reach.
"""
import numpy as np
import pytest

TX, TY = 64, 4
OX, OY = TX - 2, TY - 2


def cdiv(a, b):
    return (a + b - 1) // b


def owned_points(ny, nx):
    """Counts of owners per u face, v face and cell, and whether every update
    read lies in the reader's run cells and in the level."""
    cu = np.zeros((ny, nx + 1), np.int32)
    cv = np.zeros((ny + 1, nx), np.int32)
    ch = np.zeros((ny, nx), np.int32)
    reads_ok = True
    a, b = np.meshgrid(np.arange(TY), np.arange(TX), indexing="ij")
    for by in range(cdiv(ny, OY)):
        for bx in range(cdiv(nx, OX)):
            i0, j0 = by * OY, bx * OX
            ie, je = min(i0 + OY, ny), min(j0 + OX, nx)
            i, j = i0 - 1 + a, j0 - 1 + b
            own_u = (i >= i0) & (i < ie) & (j >= j0) & ((j < je) | (j == nx))
            own_v = (j >= j0) & (j < je) & (i >= i0) & ((i < ie) | (i == ny))
            own_h = (i >= i0) & (i < ie) & (j >= j0) & (j < je)
            np.add.at(cu, (i[own_u], j[own_u]), 1)
            np.add.at(cv, (i[own_v], j[own_v]), 1)
            np.add.at(ch, (i[own_h], j[own_h]), 1)
            # cells whose updates the owned points read: (i, j-1) and (i, j)
            # for u faces, (i-1, j) and (i, j) for v faces, (i, j) for H
            for own, di, dj, guard in ((own_u, 0, -1, j > 0), (own_u, 0, 0, j < nx),
                                       (own_v, -1, 0, i > 0), (own_v, 0, 0, i < ny), (own_h, 0, 0, own_h)):
                sel = own & guard
                ci, cj = i[sel] + di, j[sel] + dj
                ra, rb = ci - i0 + 1, cj - j0 + 1       # the reading workgroup's thread of that cell
                reads_ok &= bool(np.all((ra >= 0) & (ra < TY) & (rb >= 0) & (rb < TX)))
                reads_ok &= bool(np.all((ci >= 0) & (ci < ny) & (cj >= 0) & (cj < nx)))
    return cu, cv, ch, reads_ok


@pytest.mark.parametrize("ny", [1, 2, 3, 4, 5, 9, 16, 17, 32, 33])
@pytest.mark.parametrize("nx", [1, 2, 12, 61, 62, 63, 64, 123, 124, 125, 128, 150])
def test_every_point_has_one_owner(ny, nx):
    cu, cv, ch, reads_ok = owned_points(ny, nx)
    assert np.all(cu == 1) and np.all(cv == 1) and np.all(ch == 1)
    assert reads_ok


def test_viscosity_tile_covers_every_clamped_run_cell():
    """Each thread runs the nearest cell of the level (clamped) and reads the
    viscosity of it and its 8 neighbours from the tile of cells
    [i0-2, i0+TY-1] x [j0-2, j0+TX-1] (index ci - i0 + 2)."""
    for ny, nx in ((1, 1), (2, 62), (5, 63), (33, 125), (128, 128)):
        for by in range(cdiv(ny, OY)):
            for bx in range(cdiv(nx, OX)):
                i0, j0 = by * OY, bx * OX
                a, b = np.meshgrid(np.arange(TY), np.arange(TX), indexing="ij")
                ic = np.clip(i0 - 1 + a, 0, ny - 1)
                jc = np.clip(j0 - 1 + b, 0, nx - 1)
                for di in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        ea, eb = ic + di - i0 + 2, jc + dj - j0 + 2
                        assert np.all((ea >= 0) & (ea < TY + 2) & (eb >= 0) & (eb < TX + 2))
