"""Pins of the CPU oracle beyond the golden fixtures: the reference's own
printed known answers and what the mathematics fixes.

* N(0) = diag(-4.5733, -4.5733, -3.4198), off-diagonal tab = -1.10e-17 at
  thick = 1 (src/irregular/2d_fft/2d_i1/scaling_logs/fft_1536x512.txt:8-9);
  strength-scaled nxx0 = -9.1465e+00 at 2.0 (same log, line 13)
* trace of the self term = -4 pi; far-field trace ~ 0 (Laplace)
* parity symmetries: xx, yy, zz even, xy odd-odd; xz, yz vanish
* FFT demag == O(N^2) direct sum (periodic and padded)
* uniform m with periodic demag: h = strength * (sum_r N(r)) m
* torque invariant: m . f = alpha (m.h)(1 - |m|^2) -> 0 on the unit sphere
* Jv == central finite difference of f (Jv omits demag and, for i4, Zeeman:
  checked with demag off and matching terms)
* (I - gamma J) P^-1 = I with J the finite-difference local Jacobian
* oracle FFT == numpy.fft (pocketfft)
"""
import ctypes as C
import math

import numpy as np
import pytest

import llg_amd as la
from oracle import oracle as orc


def test_self_coupling_known_answer():
    t = orc.tensor_at(1.0, 0, 0)
    assert f"{t[0]:.4e}" == "-4.5733e+00"
    assert f"{t[4]:.4e}" == "-4.5733e+00"
    assert f"{t[8]:.4e}" == "-3.4198e+00"
    assert f"{t[1]:.2e}" == "-1.10e-17"
    assert t[2] == 0.0 and t[5] == 0.0
    assert f"{2.0 * t[0]:.4e}" == "-9.1465e+00"
    assert abs(t[0] + t[4] + t[8] + 4 * math.pi) < 1e-12


def test_tensor_symmetries_and_far_field():
    for dx, dy in [(3, 1), (5, -2), (1, 4)]:
        a, b = orc.tensor_at(1.0, dx, dy), orc.tensor_at(1.0, -dx, dy)
        c = orc.tensor_at(1.0, dx, -dy)
        for k in (0, 4, 8):
            assert abs(a[k] - b[k]) <= 1e-14 * max(1.0, abs(a[k])) and abs(a[k] - c[k]) <= 1e-14
        assert abs(a[1] + b[1]) < 1e-13 and abs(a[1] + c[1]) < 1e-13
        assert abs(a[1] - a[3]) < 1e-13            # ab == ba
        for k in (2, 5, 6, 7):                      # xz, yz, zx, zy vanish
            assert abs(a[k]) < 1e-14
        # far-field: off-centre trace is tiny compared with the self term
        assert abs(a[0] + a[4] + a[8]) < 0.05


@pytest.mark.parametrize("padded", [False, True])
def test_fft_demag_equals_direct_sum(padded):
    nx, ny = 20, 14
    y = np.random.default_rng(3).standard_normal(3 * nx * ny)
    d = orc.Demag(nx, ny, 1.0, 3.0, padded)
    h = d.apply(y)
    h2 = orc.demag_direct(nx, ny, 1.0, 3.0, y, padded)
    assert np.max(np.abs(h - h2)) < 1e-12 * np.max(np.abs(h2))


def test_uniform_periodic_field_is_tensor_sum():
    nx, ny, s = 16, 12, 2.5
    T = orc.calt(1.0, nx, ny)
    m = np.array([0.3, -0.5, 0.8])
    y = np.repeat(m, nx * ny)
    h = orc.Demag(nx, ny, 1.0, s).apply(y).reshape(3, -1)
    S = T.reshape(9, -1).sum(1).reshape(3, 3)
    expect = s * S @ m
    assert np.allclose(h, expect[:, None], rtol=0, atol=1e-12 * np.abs(expect).max())


def test_oracle_fft_matches_numpy():
    L = orc.load()
    for n in (12, 30, 48, 1536, 7):
        x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
        z = x.copy()
        L.orc_fft1d(z.ctypes.data_as(C.POINTER(C.c_double)), n, -1)
        assert np.max(np.abs(z - np.fft.fft(x))) < 1e-13 * np.max(np.abs(z))


def _unit(y, n):
    m = y.reshape(3, n)
    return (m / np.linalg.norm(m, axis=0)).reshape(-1)


def test_regularised_normalisation_fixed_point():
    ng, ny = 6, 5
    sim = orc.OracleSim(la.preset("i2", ng=ng, ny=ny, demag=0.0), la.Geometry.full(ng, ny))
    y = _unit(np.random.default_rng(4).standard_normal(3 * ng * ny), ng * ny)
    y[0] = y[ng * ny] = y[2 * ng * ny] = 0.0        # a hole cell stays exactly 0
    _, yn = sim.rhs(y)
    n = np.linalg.norm(yn.reshape(3, -1), axis=0)
    assert n[0] == 0.0
    assert np.allclose(n[1:], 1.0 / 1.001, rtol=1e-14)


@pytest.mark.parametrize("name", ["i4", "i1", "plain128"])
def test_torque_invariant_on_unit_sphere(name):
    ng, ny = 12, 10
    cfg = la.preset(name, ng=ng, ny=ny) if name == "plain128" else la.preset(name, ng=ng, ny=ny, demag=0.0)
    sim = orc.OracleSim(cfg, la.Geometry.full(ng, ny))
    y = _unit(np.random.default_rng(1).standard_normal(3 * ng * ny), ng * ny)
    f, yn = sim.rhs(y)
    dot = (yn.reshape(3, -1) * f.reshape(3, -1)).sum(0)
    assert np.max(np.abs(dot)) < 1e-12


@pytest.mark.parametrize("name", ["i4nozee", "i1", "plain128", "i6"])
def test_jv_equals_finite_difference(name):
    ng, ny = 10, 8
    if name == "i4nozee":
        cfg = la.preset("i4", ng=ng, ny=ny, demag=0.0)
        cfg.rhs.happ[0] = 0.0                      # Jv omits the Zeeman term
        geom = la.Geometry.square_hole(ng, ny, halfside_frac=0.2)
    elif name == "i6":
        cfg = la.preset("i6", ng=ng, ny=ny, demag=0.0)
        geom = la.Geometry.ellipse(ng, ny, 0.35, 0.35)
    elif name == "i1":
        cfg = la.preset("i1", ng=ng, ny=ny, demag=0.0)
        cfg.normalize = la.NORM_NONE
        geom = la.Geometry.full(ng, ny)
    else:
        cfg = la.preset(name, ng=ng, ny=ny)
        geom = la.Geometry.full(ng, ny)
    sim = orc.OracleSim(cfg, geom)
    rng = np.random.default_rng(7)
    act = np.tile(geom.active.astype(float), 3)
    y = rng.standard_normal(3 * ng * ny) * act
    v = rng.standard_normal(3 * ng * ny) * act
    eps = 1e-6
    fp, _ = sim.rhs(y + eps * v)
    fm, _ = sim.rhs(y - eps * v)
    fd = (fp - fm) / (2 * eps)
    jv = sim.jtv(y, v)
    assert np.max(np.abs(jv - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))


@pytest.mark.parametrize("name", ["i4nodemag", "plain128", "i1"])
def test_block_jacobi_inverts_local_jacobian(name):
    ng, ny = 8, 6
    if name == "i4nodemag":
        cfg = la.preset("i4", ng=ng, ny=ny, demag=0.0)
        cfg.pc.happ[0] = cfg.rhs.happ[0]           # same field as the RHS
    elif name == "i1":
        cfg = la.preset("i1", ng=ng, ny=ny, demag=0.0)
        cfg.normalize = la.NORM_NONE
    else:
        cfg = la.preset(name, ng=ng, ny=ny)
    sim = orc.OracleSim(cfg, la.Geometry.full(ng, ny))
    N = ng * ny
    y = np.random.default_rng(9).standard_normal(3 * N)
    gamma = 0.05
    sim.rhs(y)
    sim.psetup(y, gamma)
    P = sim.pinv().reshape(3, 3, N)
    eps = 1e-6
    for cell in (0, 5, N // 2, N - 1):
        J = np.zeros((3, 3))
        for b in range(3):
            e = np.zeros(3 * N)
            e[b * N + cell] = eps
            fp, _ = sim.rhs(y + e)
            fm, _ = sim.rhs(y - e)
            J[:, b] = ((fp - fm) / (2 * eps)).reshape(3, N)[:, cell]
        A = np.eye(3) - gamma * J
        if name == "i1":
            # the reference's block keeps only the diagonal of the rank-1
            # linear-anisotropy derivative; with msk = (1,0,0) that is exact
            pass
        assert np.max(np.abs(A @ P[:, :, cell] - np.eye(3))) < 1e-7


def test_psolve_equals_dense_solve():
    ng, ny = 8, 6
    cfg = la.preset("i4", ng=ng, ny=ny, demag=4.0)
    geom = la.Geometry.square_hole(ng, ny, halfside_frac=0.2)
    sim = orc.OracleSim(cfg, geom)
    N = ng * ny
    y = np.random.default_rng(2).standard_normal(3 * N) * np.tile(geom.active, 3)
    sim.rhs(y)
    sim.psetup(y, 0.02)
    P = sim.pinv().reshape(3, 3, N)
    r = np.random.default_rng(3).standard_normal(3 * N)
    z = sim.psolve(r).reshape(3, N)
    for cell in range(N):
        A = np.linalg.inv(P[:, :, cell])
        expect = np.linalg.solve(A, r.reshape(3, N)[:, cell]) * geom.active[cell]
        assert np.allclose(z[:, cell], expect, rtol=1e-12, atol=1e-12)


def test_singular_block_falls_back_to_identity():
    ng, ny = 6, 4
    cfg = la.preset("plain128", ng=ng, ny=ny)
    sim = orc.OracleSim(cfg, la.Geometry.full(ng, ny))
    N = ng * ny
    y = np.zeros(3 * N)          # m = 0: J = alpha k ... choose gamma to make A singular
    sim.rhs(y)
    # with m = 0 the local J is diag(alpha k_a) (+ Zeeman cross terms vanish);
    # pick gamma so that 1 - gamma*J_22 = 0 exactly for interior cells
    k_z = cfg.pc.chk - 4.0 * cfg.pc.che
    gamma = 1.0 / (cfg.alpha * k_z)
    sim.psetup(y, gamma)
    P = sim.pinv().reshape(3, 3, N)
    interior = [j * ng + i for j in range(1, ny - 1) for i in range(1, ng - 1)]
    for c in interior:
        assert np.array_equal(P[:, :, c], np.eye(3))


def _fd_block(sim, y, N, cell, eps=1e-6):
    """local 3x3 Jacobian of f at y with respect to the cell's own m"""
    J = np.zeros((3, 3))
    for b in range(3):
        e = np.zeros(3 * N)
        e[b * N + cell] = eps
        fp, _ = sim.rhs(y + e)
        fm, _ = sim.rhs(y - e)
        J[:, b] = ((fp - fm) / (2 * eps)).reshape(3, N)[:, cell]
    return J


@pytest.mark.parametrize("demag", [0.0, 4.0])
def test_jok_two_phase_reuses_the_stored_jacobian(demag):
    """Two-phase protocol (src/2d/2d_heter/precond.cu:154-192,265-314):
    psetup(y1, g1, jok=0) builds J(y1) and reports jcur = TRUE; a later
    psetup(y2, g2, jok=1) keeps J(y1), refactors with g2 and reports
    jcur = FALSE, so P^-1 = inv(I - g2 J_FD(y1)) -- not J(y2) -- while
    jok = 0 at y2 gives inv(I - g2 J_FD(y2)).  J_FD is the central
    difference of f in the cell's own m (the preconditioner's field set
    equal to the RHS's; with demag the self term strength*N(0) enters both)."""
    ng, ny = 8, 6
    N = ng * ny
    cfg = la.preset("i4", ng=ng, ny=ny, demag=demag, pc_mode=la.PC_JOK)
    cfg.pc.happ[0] = cfg.rhs.happ[0]
    sim = orc.OracleSim(cfg, la.Geometry.full(ng, ny))
    rng = np.random.default_rng(21)
    y1 = _unit(rng.standard_normal(3 * N), N)
    y2 = _unit(y1 + 0.3 * rng.standard_normal(3 * N), N)
    g1, g2 = 0.05, 0.02
    sim.rhs(y1)
    assert sim.psetup(y1, g1, jok=0) == 1
    sim.rhs(y2)
    assert sim.psetup(y2, g2, jok=1) == 0
    P_reuse = sim.pinv().reshape(3, 3, N).copy()
    assert sim.psetup(y2, g2, jok=0) == 1
    P_fresh = sim.pinv().reshape(3, 3, N).copy()
    for cell in (0, 7, N // 2, N - 1):
        J1 = _fd_block(sim, y1, N, cell)
        J2 = _fd_block(sim, y2, N, cell)
        A1, A2 = np.eye(3) - g2 * J1, np.eye(3) - g2 * J2
        assert np.max(np.abs(A1 @ P_reuse[:, :, cell] - np.eye(3))) < 1e-7
        assert np.max(np.abs(A2 @ P_fresh[:, :, cell] - np.eye(3))) < 1e-7
        # the two differ far beyond the tolerance: reusing J(y2) would fail
        assert np.max(np.abs(A1 @ P_fresh[:, :, cell] - np.eye(3))) > 1e-3


def test_i3_torque_invariant_on_unit_sphere():
    """The windowed polycrystal RHS (per-cell axes, DMI on w*m, windowed
    demag, w output gate) is a torque: m . f = w alpha (m.h)(1 - |m|^2) = 0
    on the unit sphere (normalisation off so |m| = 1 exactly)."""
    ng, ny = 16, 12
    geom = la.Geometry.polycrystal(ng, ny, num_grains=6, dead_frac=0.34)
    cfg = la.preset("i3", ng=ng, ny=ny, demag=1.0)
    cfg.normalize = la.NORM_NONE
    sim = orc.OracleSim(cfg, geom)
    y = _unit(np.random.default_rng(8).standard_normal(3 * ng * ny), ng * ny)
    f, yn = sim.rhs(y)
    dot = (yn.reshape(3, -1) * f.reshape(3, -1)).sum(0)
    assert np.max(np.abs(f)) > 1.0
    assert np.max(np.abs(dot)) < 1e-12 * np.max(np.abs(f))
