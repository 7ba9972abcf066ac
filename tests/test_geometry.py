"""Geometry and initial-condition builders (product, C++) against closed
forms and Python restatements of the reference host code.

* square-hole counts: 512^2 cells with half-side 0.25 / 0.22 / 0.05 of ny
  give 256^2 / 226^2 / 52^2 hole cells; the 1536x512 workload (half-side
  222 cells) gives 444^2 = 197136 (25.07 %)
* masks bit-exact against numpy predicates (2d_i4/2d_fft.cu:500-511,
  2d_i5/2d_fft.cu:516-535, 2d_i6/2d_i6.cu:469-505)
* polycrystal: grains from glibc rand() in the reference call order
  (2d_i3/2d_fft.cu:253-298), Voronoi with periodic images (:300-315), and
  the window w from the brute-force O(N * holes) loop (:335-370) --
  the product's exact distance transform must give identical w
* ICs: two-domain (2d_i4/2d_fft.cu:567-610), stripes with glibc rand
  (2d_i1/2d_fft.cu:625-657)
"""
import ctypes as C
import math
import sys
from pathlib import Path

import numpy as np
import pytest

import llg_amd as la

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import gen_fixtures as gf  # noqa: E402

libc = C.CDLL("libc.so.6")
libc.rand.restype = C.c_int
RAND_MAX = 2147483647


def rand_unit():
    return libc.rand() / RAND_MAX


@pytest.mark.parametrize("frac,side", [(0.25, 256), (0.22, 226), (0.05, 52)])
def test_square_hole_counts(frac, side):
    g = la.Geometry.square_hole(512, 512, halfside_frac=frac)
    assert int((g.active == 0).sum()) == side * side


def test_workload_hole_is_444_squared():
    g = la.Geometry.square_hole(1536, 512, halfside_frac=222 / 512)
    assert int((g.active == 0).sum()) == 444 * 444 == 197136


def test_masks_match_numpy_predicates():
    assert np.array_equal(la.Geometry.square_hole(90, 70, 0.4, 0.6, 0.2).active,
                          gf.square_hole(90, 70, 0.4, 0.6, 0.2))
    assert np.array_equal(la.Geometry.ellipse(80, 64, 0.3, 0.2).active, gf.ellipse(80, 64, 0.3, 0.2))
    ng, ny = 60, 40
    I, J = np.meshgrid(np.arange(ng), np.arange(ny))
    rect = ((np.abs(I + 0.5 - 30) <= 15) & (np.abs(J + 0.5 - 20) <= 10)).astype(np.uint8).reshape(-1)
    assert np.array_equal(la.Geometry.rectangle(ng, ny, 0.5, 0.5).active, rect)


def _poly_reference(ng, ny, G, dead_frac, seed, zb, eps):
    libc.srand(seed)
    gx = max(1, int(math.ceil(math.sqrt(G * ng / ny))))
    gy = (G + gx - 1) // gx
    dxs, dys = ng / gx, ny / gy
    S = []
    for j in range(gy):
        for i in range(gx):
            if len(S) >= G:
                break
            sx = (i + 0.15 + 0.7 * rand_unit()) * dxs
            sy = (j + 0.15 + 0.7 * rand_unit()) * dys
            S.append([sx % ng, sy % ny])
    while len(S) < G:
        S.append([rand_unit() * ng, rand_unit() * ny])
    grains = []
    for g in range(G):
        u = rand_unit()
        sign = -1.0 if rand_unit() < 0.5 else 1.0
        z = sign * u ** (1.0 / zb)
        phi = 2 * math.pi * rand_unit()
        rxy = math.sqrt(max(0.0, 1 - z * z))
        psi = 2 * math.pi * rand_unit()
        dead = 1 if rand_unit() < dead_frac else 0
        grains.append((rxy * math.cos(phi), rxy * math.sin(phi), z, math.cos(psi), math.sin(psi), dead))
    gid = np.zeros(ng * ny, np.int32)
    for j in range(ny):
        for i in range(ng):
            best, bd = 0, 1e30
            for g in range(G):
                for dyp in (-1, 0, 1):
                    for dxp in (-1, 0, 1):
                        dx = i + 0.5 - (S[g][0] + dxp * ng)
                        dy = j + 0.5 - (S[g][1] + dyp * ny)
                        d2 = dx * dx + dy * dy
                        if d2 < bd:
                            bd, best = d2, g
            gid[j * ng + i] = best
    dead = np.array([grains[g][5] for g in gid], bool)
    hx, hy = np.nonzero(dead.reshape(ny, ng))[1], np.nonzero(dead.reshape(ny, ng))[0]
    w = np.ones(ng * ny)
    for j in range(ny):
        for i in range(ng):
            c = j * ng + i
            if dead[c]:
                w[c] = 0.0
                continue
            ddx = i - hx
            ddy = j - hy
            ddx = np.where(ddx > ng // 2, ddx - ng, ddx)
            ddx = np.where(ddx < -(ng // 2), ddx + ng, ddx)
            ddy = np.where(ddy > ny // 2, ddy - ny, ddy)
            ddy = np.where(ddy < -(ny // 2), ddy + ny, ddy)
            d2 = float(np.min(ddx * ddx + ddy * ddy))
            d = math.sqrt(d2) - 0.5
            ww = 0.5 * (1 + math.tanh(d / eps))
            ww = 0.0 if ww < 1e-3 else (1.0 if ww > 1 - 1e-3 else ww)
            w[c] = ww
    return np.array(S), grains, gid, w


def test_polycrystal_matches_reference_restatement():
    ng, ny, G = 48, 36, 10
    g = la.Geometry.polycrystal(ng, ny, num_grains=G, dead_frac=0.3, seed=20251104)
    S, grains, gid, w = _poly_reference(ng, ny, G, 0.3, 20251104, 1.6, 2.2)
    got = np.array([[g.grains[k].sx, g.grains[k].sy] for k in range(G)])
    assert np.array_equal(got, S)
    assert [g.grains[k].dead for k in range(G)] == [gr[5] for gr in grains]
    assert np.array_equal(g.grain_id, gid)
    assert np.array_equal(g.w, w)
    assert np.array_equal(g.active, (w > 0).astype(np.uint8) | (~np.array([grains[k][5] for k in gid], bool)).astype(np.uint8))
    msk = g.msk.reshape(3, -1)
    assert np.array_equal(msk[2], np.array([grains[k][2] for k in gid]))


@pytest.mark.parametrize("ng,ny,G", [(200, 160, 120), (257, 97, 400), (96, 96, 1)])
def test_polycrystal_block_pruned_voronoi_equals_full_scan(ng, ny, G):
    """The builder scans, per 32x32 block, only the seed images that can be
    nearest; the ids must equal the full scan over every grain and all 3x3
    periodic images with the reference's d2 arithmetic and first-in-scan-
    order ties (2d_i3/2d_fft.cu:300-315).  Grids that are not multiples of
    the block and 400 grains (the 4096^2 workload's count) included."""
    g = la.Geometry.polycrystal(ng, ny, num_grains=G, dead_frac=0.1, seed=7)
    S = np.array([[g.grains[k].sx, g.grains[k].sy] for k in range(G)])
    I, J = np.meshgrid(np.arange(ng) + 0.5, np.arange(ny) + 0.5)
    x, y = I.ravel(), J.ravel()
    best = np.full(x.size, 1e30)
    gid = np.zeros(x.size, np.int64)
    for k in range(G):                      # the reference's loop order
        for dyp in (-1, 0, 1):
            for dxp in (-1, 0, 1):
                dx = x - (S[k, 0] + dxp * float(ng))
                dy = y - (S[k, 1] + dyp * float(ny))
                d2 = dx * dx + dy * dy
                take = d2 < best
                best = np.where(take, d2, best)
                gid = np.where(take, k, gid)
    assert np.array_equal(g.grain_id, gid)


def test_two_domain_ic():
    ng, ny = 40, 32
    g = la.Geometry.square_hole(ng, ny, halfside_frac=0.2)
    y = g.ic_two_domain(tail_my=0.0175).reshape(3, ny, ng)
    js = (7 * ny) // 8
    act = g.active.reshape(ny, ng).astype(bool)
    n = math.sqrt(1 + 0.0175 ** 2)
    for j in range(ny):
        for i in range(ng):
            if not act[j, i]:
                assert (y[:, j, i] == 0).all()
            elif j >= js:
                assert tuple(y[:, j, i]) == (-1.0, 0.0, 0.0)
            else:
                my = -0.0175 if i < ng // 2 else 0.0175
                assert np.allclose(y[:, j, i], [1 / n, my / n, 0.0], rtol=0, atol=1e-16)


def test_stripes_ic_uses_glibc_rand_order():
    ng, ny = 20, 6
    y = la.Geometry.full(ng, ny).ic_stripes(0.25, 0.75, 0.01, 12345).reshape(3, ny, ng)
    libc.srand(12345)
    for j in range(ny):
        for i in range(ng):
            mx = 1.0 if 5 <= i < 15 else -1.0
            my = 0.01 * (2.0 * libc.rand() / RAND_MAX - 1.0)
            mz = 0.01 * (2.0 * libc.rand() / RAND_MAX - 1.0)
            nrm = math.sqrt(mx * mx + my * my + mz * mz)
            assert np.array_equal(y[:, j, i], [mx / nrm, my / nrm, mz / nrm])


def test_seeded_ics_are_unit_and_masked():
    g = la.Geometry.square_hole(64, 32, halfside_frac=0.25)
    for y in (g.ic_split_x(), g.ic_tilted_z(), g.ic_random_sphere()):
        m = y.reshape(3, -1)
        n = np.linalg.norm(m, axis=0)
        assert np.allclose(n[g.active == 1], 1.0, atol=1e-15)
        assert (n[g.active == 0] == 0).all()
    m = g.ic_split_x(0.05).reshape(3, 32, 64)
    assert (m[0, :, :32][g.active.reshape(32, 64)[:, :32] == 1] > 0.99).all()
    assert (m[0, :, 32:][g.active.reshape(32, 64)[:, 32:] == 1] < -0.99).all()


def test_frame_and_aux_file_formats(tmp_path):
    """output.txt frames, mask.txt and grain_id.txt in the reference's text
    formats (2d_i4/2d_fft.cu:393-411, 2d_i3/2d_fft.cu:390-408)."""
    L = la.lib()
    ng, ny = 4, 3
    y = np.arange(3 * ng * ny, dtype=float) / 7.0
    w = np.linspace(0, 1, ng * ny)
    p = str(tmp_path / "output.txt").encode()
    h = L.LLG_FrameWriterOpen(p)
    assert L.LLG_FrameWriterWrite(h, 0.5, 3 * ng, ny, ng, y.ctypes.data_as(C.POINTER(C.c_double)),
                                  None, 0) == 0
    assert L.LLG_FrameWriterWrite(h, 1.0, 3 * ng, ny, ng, y.ctypes.data_as(C.POINTER(C.c_double)),
                                  w.ctypes.data_as(C.POINTER(C.c_double)), 0) == 0
    assert L.LLG_FrameWriterClose(h) == 0
    lines = open(p.decode()).read().split("\n")
    assert lines[0] == "0.500000 12 3"
    assert lines[1] == "%f %f %f" % (y[0], y[12], y[24])
    assert lines[1 + ng * ny] == ""
    k = 2 + ng * ny
    assert lines[k] == "1.000000 12 3"
    assert lines[k + 5] == "%f %f %f" % (w[4] * y[4], w[4] * y[16], w[4] * y[28])
    assert L.LLG_ShouldWriteFrame(50, 10.0, 100.0, 50, 500) == 1
    assert L.LLG_ShouldWriteFrame(51, 10.0, 100.0, 50, 500) == 0
    assert L.LLG_ShouldWriteFrame(1000, 200.0, 100.0, 50, 500) == 1
    mp = str(tmp_path / "mask.txt").encode()
    assert L.LLG_WriteMaskFile(mp, ng, ny, w.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert open(mp.decode()).read().split("\n")[:3] == ["4 3", "%f" % w[0], "%f" % w[1]]
    gid = np.arange(ng * ny, dtype=np.int32)
    gp = str(tmp_path / "grain_id.txt").encode()
    assert L.LLG_WriteGrainIdFile(gp, ng, ny, gid.ctypes.data_as(C.POINTER(C.c_int))) == 0
    assert open(gp.decode()).read().split("\n")[:3] == ["4 3", "0", "1"]
