"""Corner elevations behind the face beds (rift_corners_from_faces).

The reference keeps the face beds BX/BY as two-point means of the DEM points
and the cell bed BC as the mean of the four faces (simulator.cu:384-414,
grid.cu:531-537).  The fused step reads one corner plane instead and forms
every bed from it.  These tests pin, on the CPU:
  * the corners recovered from the faces against the closed form of the
    reconstruction (the DEM minus its own checkerboard mean),
  * the refusal of faces that are not means of one corner field,
  * the FP32 beds formed from the corner plane (period-2 extension beyond the
    corners) against the reference's beds, ghost cells and faces included.
"""
import math

import numpy as np
import pytest

from rift_amd import kp, synth


def _dem(b_ny, b_nx, seed, scale=60.0):
    side = max(b_ny, b_nx) - 1
    side = 1 << max(1, math.ceil(math.log2(side)))
    return scale * synth.fractal_dem(side, 0.6, seed)[:b_ny, :b_nx]


@pytest.mark.parametrize("shape", [(33, 40), (17, 9), (64, 65)])
def test_corners_are_the_dem_less_its_checkerboard_mean(shape):
    dem = _dem(*shape, seed=11) + 5.0
    geo = kp.geometry(dem)
    D = kp.corners_from_faces(geo["BX"], geo["BY"])
    assert D is not None and D.shape == dem.shape
    s = np.where((np.add.outer(np.arange(shape[0]), np.arange(shape[1])) & 1) == 1, -1.0, 1.0)
    expected = dem - np.mean(s * dem) * s
    np.testing.assert_allclose(D, expected, rtol=0, atol=1e-10)


def test_corner_means_reproduce_every_used_face():
    dem = _dem(29, 45, seed=3)
    geo = kp.geometry(dem)
    BX, BY = geo["BX"], geo["BY"]
    ny, nx = geo["ny"], geo["nx"]
    D = kp.corners_from_faces(BX, BY)
    # interior faces: BX(i, j) for i in [2, nx-2], j in [2, ny-3]; BY(i, j)
    # for i in [2, nx-3], j in [2, ny-2]
    bx = 0.5 * (D[:-1, :] + D[1:, :])           # (ny-4, nx-3)
    by = 0.5 * (D[:, :-1] + D[:, 1:])           # (ny-3, nx-4)
    np.testing.assert_allclose(bx, BX[2:ny - 2, 2:nx - 1], rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(by, BY[2:ny - 1, 2:nx - 2], rtol=1e-13, atol=1e-12)


def test_faces_that_are_no_corner_means_are_refused():
    dem = _dem(20, 24, seed=5)
    geo = kp.geometry(dem)
    BX, BY = geo["BX"].copy(), geo["BY"].copy()
    assert kp.corners_from_faces(BX, BY) is not None
    bad = BX.copy()
    bad[7, 9] += 1e-5                     # one interior face off its corners
    assert kp.corners_from_faces(bad, BY) is None
    bad = BY.copy()
    bad[12, 6] -= 1e-5
    assert kp.corners_from_faces(BX, bad) is None
    bad = BX.copy()
    bad[5, 1] += 1.0                      # a ghost face not repeating BX(3)
    assert kp.corners_from_faces(bad, BY) is None
    bad = BY.copy()
    bad[-2, 4] += 1.0                     # ghost row ny-1 not repeating BY(ny-3)
    assert kp.corners_from_faces(BX, bad) is None
    rng = np.random.default_rng(0)
    assert kp.corners_from_faces(rng.normal(size=BX.shape), rng.normal(size=BY.shape)) is None


def _plane_dims(nx, ny):
    nbx, nby = -(-nx // 16), -(-ny // 14)
    nstrips = -(-(nx - 4) // 60)
    P = ((max(16 * nbx + 4, 60 * nstrips + 4) + 63) // 64) * 64
    return P, 14 * nby + 4


def _ext(c, n):
    """Period-2 extension of corner index c into [0, n)."""
    if c < 0:
        return c & 1
    if c > n - 1:
        return n - 1 - ((n - 1 - c) & 1)
    return c


@pytest.mark.parametrize("shape", [(30, 41), (9, 130)])
def test_fp32_beds_from_the_corner_plane_match_the_reference_beds(shape):
    """The corner plane as the library lays it out (period 2 beyond the
    corners) and the FP32 bed expressions of the kernels give, at every cell
    a kernel reads, the reference's beds within FP32 rounding: interior and
    ghost faces (ghost faces repeat the faces two cells in), and every
    plane cell's BC equal to the BC of the nearest interior cell."""
    dem = _dem(*shape, seed=21) + 10.0
    geo = kp.geometry(dem)
    BX, BY = geo["BX"], geo["BY"]
    nx, ny = geo["nx"], geo["ny"]
    D = kp.corners_from_faces(BX, BY)
    P, R = _plane_dims(nx, ny)
    W, H = nx - 3, ny - 3
    rows = np.array([_ext(j - 2, H) for j in range(R)])
    cols = np.array([_ext(i - 2, W) for i in range(P)])
    DC = D.astype(np.float32)[np.ix_(rows, cols)]          # (R, P)
    i1 = np.r_[np.arange(1, P), P - 2]
    j1 = np.r_[np.arange(1, R), R - 2]
    half, quarter = np.float32(0.5), np.float32(0.25)
    p0 = DC + DC[:, i1]
    p1 = p0[j1, :]
    BC32 = quarter * (p0 + p1)
    BY32 = half * p0
    BX32 = half * (DC + DC[j1, :])
    # reference BC (InitGrid_k: mean of the four faces) of the nearest interior cell
    bc = 0.25 * (BX[:, 1:] + BX[:, :-1] + BY[1:, :] + BY[:-1, :])
    jc = np.clip(np.arange(R), 2, ny - 3)
    ic = np.clip(np.arange(P), 2, nx - 3)
    bc_plane = bc[np.ix_(jc, ic)]
    tol = 4 * np.finfo(np.float32).eps * (np.abs(dem).max() + 1.0)
    assert np.abs(BC32 - bc_plane).max() <= tol
    # west faces read by the kernels: interior rows, columns 0 .. nx-1
    assert np.abs(BX32[2:ny - 2, :nx] - BX[2:ny - 2, :nx]).max() <= tol
    # south faces: interior columns, rows 0 .. ny-1
    assert np.abs(BY32[:ny, 2:nx - 2] - BY[:ny, 2:nx - 2]).max() <= tol
    # a plane cell's BC is bit-identical wherever the nearest interior cell is the same
    for j in range(R):
        for i in (0, 1, nx - 2, nx - 1, P - 1):
            assert BC32[j, i] == BC32[min(max(j, 2), ny - 3), min(max(i, 2), nx - 3)]
