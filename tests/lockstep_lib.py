"""Tile-flag comparison of a device run against the FP64 oracle where a
flood front crosses dry land (test helper).

The wet flag of a tile is bit-exact given the same depths (the Track/Grow
rules are pinned on the device's own flags elsewhere); here the depths are
FP32 against FP64, and a front cell the oracle leaves at 2e-7 m FP32 may put
at or below epsilon (1.19e-7 m).  So wet flags may differ only on tiles
whose wettest interior cell is below THIN in both runs, and active sets only
where the Grow rule propagates such a difference from the step before."""
import numpy as np

THIN = 1e-5   # m: the wettest cell of a tile whose flag may differ


def tile_max_depth(h, nbx, nby):
    """Per-tile maximum depth over each 16 x 14 tile's interior cells."""
    ny, nx = h.shape
    out = np.zeros(nbx * nby)
    for by in range(nby):
        j0, j1 = by * 14 + 2, min(by * 14 + 16, ny - 2)
        if j1 <= j0:
            continue
        rows = h[j0:j1]
        for bx in range(nbx):
            i0, i1 = bx * 16 + 2, min(bx * 16 + 18, nx - 2)
            if i1 > i0:
                out[by * nbx + bx] = rows[:, i0:i1].max()
    return out


def dilate(tiles, nbx, nby):
    out = set()
    for t in tiles:
        bx, by = t % nbx, t // nbx
        out.add(t)
        for ddx, ddy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            x, y = bx + ddx, by + ddy
            if 0 <= x < nbx and 0 <= y < nby:
                out.add(y * nbx + x)
    return out


def check_tiles(g, o, prev, label):
    """Compare the device handle g with oracle o after a step; prev = the
    wet-flag differences of the step before.  Returns this step's."""
    nbx, nby = g.nbx, g.nby
    gw, ga, _ = g.tiles()
    ow, oa, _ = o.tiles()
    oact = np.zeros(g.ntiles, np.uint8)
    oact[[int(v) for v in oa if v > -1]] = 1
    oact[0] = 1
    wd = set(np.nonzero(gw != ow)[0].tolist())
    if wd:
        hg = tile_max_depth(g.field("depth"), nbx, nby)
        ho = tile_max_depth(o.depth(), nbx, nby)
        for t in wd:
            assert hg[t] < THIN and ho[t] < THIN, (label, t, hg[t], ho[t])
    ad = set(np.nonzero(ga != oact)[0].tolist())
    assert ad <= dilate(prev, nbx, nby), (label, sorted(ad - dilate(prev, nbx, nby))[:10])
    return wd
