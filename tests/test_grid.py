"""Builder invariants: maps, ghosts and numbering follow the rules of
model/hd/master/preprocess.c:343-720 (checked by brute force on the
(i,j,k) grid)."""
import numpy as np

from ems_amd import configs, structs as S
from tests import helpers as H


def _arrays(case):
    g = case.g
    n = case.sgsiz
    A = {k: S.view(getattr(g, k), n, np.int32) for k in
         ("xp1", "xm1", "yp1", "ym1", "zp1", "zm1", "s2i", "s2j", "s2k", "m2d")}
    return g, A


def test_numbering_and_maps_small_case():
    case = configs.small(seed=11, nxi=14, nyi=12, nz=6)
    g, A = _arrays(case)
    nz, ny, nx = case.nz, case.ny, case.nx
    kb = case.kbot
    wet = case.wet
    # wet cells come layer by layer from the surface; (j,i) order inside
    keys = [(-A["s2k"][c], A["s2j"][c], A["s2i"][c]) for c in wet]
    assert keys == sorted(keys)
    assert np.all(np.diff(wet) > 0)
    # surface layer first: its wet cells are 2D cells 1..v2_t
    w2 = S.view(g.w2_t, case.sgsizS, np.int32)[1:g.v2_t + 1]
    assert np.all(A["s2k"][w2] == nz - 1)
    assert np.all(A["m2d"][w2] == w2)
    idx = {}
    for c in wet:
        idx[(A["s2i"][c], A["s2j"][c], A["s2k"][c])] = c
    bpt = S.view(g.bpt, g.nbpt + 1, np.int32)[1:]
    bin_ = S.view(g.bin, g.nbpt + 1, np.int32)[1:]
    ghost_of = dict(zip(bpt, bin_))
    dirs = {"xp1": (1, 0), "xm1": (-1, 0), "yp1": (0, 1), "ym1": (0, -1)}
    opp = {"xp1": "xm1", "xm1": "xp1", "yp1": "ym1", "ym1": "yp1"}
    for c in wet:
        i, j, k = A["s2i"][c], A["s2j"][c], A["s2k"][c]
        for m, (di, dj) in dirs.items():
            nb = A[m][c]
            key = (i + di, j + dj, k)
            if k >= kb[j + dj, i + di]:
                assert nb == idx[key]
            else:
                # own ghost: maps back, self elsewhere, listed with bin = c
                assert nb not in idx.values()
                assert ghost_of[nb] == c
                assert A[opp[m]][nb] == c
                assert A[m][nb] == nb
        # vertical
        if k == nz - 1:
            assert A["zp1"][c] == c
        else:
            assert A["zp1"][c] == idx[(i, j, k + 1)]
        if k > kb[j, i]:
            assert A["zm1"][c] == idx[(i, j, k - 1)]
        else:
            gs = A["zm1"][c]
            assert A["zm1"][gs] == gs and A["zp1"][gs] == c
        assert A["m2d"][c] == idx[(i, j, nz - 1)]
    # every 3D sparse cell is either wet or a ghost, all distinct
    n_gh = g.nbpt + g.ngsed
    assert g.sgnum == len(wet) + n_gh
    assert g.v3_t == len(wet)


def test_ghost_fill_no_gradient():
    case = configs.small(seed=3)
    g = case.g
    bpt = S.view(g.bpt, g.nbpt + 1, np.int32)[1:]
    bin_ = S.view(g.bin, g.nbpt + 1, np.int32)[1:]
    for n in range(case.ntr):
        t = case.tr(n)
        assert np.array_equal(t[bpt], t[bin_])


def test_c1_sizes_and_flow():
    case = configs.c1()
    assert case.nwet == 40 * 40 * 10
    u1 = case.a3(case.wd.u1)
    u2 = case.a3(case.wd.u2)
    assert abs(max(np.abs(u1).max(), np.abs(u2).max()) - 0.5) < 1e-12
    # non-divergent per layer: fluxes of every wet cell sum to ~0
    g = case.g
    f1 = case.a3(case.wd.u1flux3d)
    f2 = case.a3(case.wd.u2flux3d)
    xp1 = S.view(g.xp1, case.sgsiz, np.int32)
    yp1 = S.view(g.yp1, case.sgsiz, np.int32)
    c = case.wet
    div = f1[xp1[c]] - f1[c] + f2[yp1[c]] - f2[c]
    assert np.abs(div).max() < 1e-9 * np.abs(f1).max()
