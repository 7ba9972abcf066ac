"""Generate the committed golden fixtures (run in the build container only).

An independent numpy restatement of the reference formulas -- written
separately from oracle/llg_oracle.c -- evaluated on small seeded grids:
  * the Newell tensor `calt` (src/irregular/2d_fft/2d_i4/demag_fft.cu:27-91)
  * periodic demag via numpy.fft with the reference's centring/shift/scale
    (2d_i4/demag_fft.cu:184-363) and padded demag by direct linear sum
  * f / Jv / block-Jacobi P^-1 / P-solve for the i4 (cubic, ymsk, Zeeman),
    i1 (linear anisotropy + DMI, exact normalisation), i6 (compact) and the
    plain 128x128-form (Neumann Laplacian, LL prefactor) physics
  * the i3 windowed polycrystal (per-cell axes, DMI on w*m, w output gate,
    windowed demag gather, fixed-axis Jv and preconditioner), written term
    by term from the i3 kernels rather than through the shared helpers
  * a scipy Radau (rtol 1e-11) trajectory of the plain form on a 6x5 grid
Usage: python tests/golden/gen_fixtures.py   (writes tests/golden/*.npz)
"""
from __future__ import annotations

import numpy as np
from pathlib import Path

OUT = Path(__file__).resolve().parent


# ------------------------------------------------------------------ tensor
def ctt(b, a, sx, sy):
    sz = 0.0
    xn, xp, yn, yp, zn, zp = sx - a, sx + a, sy - a, sy + a, sz - b, sz + b
    s = np.sqrt
    dnnn, dpnn = s(xn**2 + yn**2 + zn**2), s(xp**2 + yn**2 + zn**2)
    dnpn, dnnp = s(xn**2 + yp**2 + zn**2), s(xn**2 + yn**2 + zp**2)
    dppn, dnpp = s(xp**2 + yp**2 + zn**2), s(xn**2 + yp**2 + zp**2)
    dpnp, dppp = s(xp**2 + yn**2 + zp**2), s(xp**2 + yp**2 + zp**2)
    at, lg = np.arctan, np.log
    d = {}
    d[1] = (at(zn*yn/(xn*dnnn)) - at(zp*yn/(xn*dnnp)) - at(zn*yp/(xn*dnpn)) + at(zp*yp/(xn*dnpp))
            - at(zn*yn/(xp*dpnn)) + at(zp*yn/(xp*dpnp)) + at(zn*yp/(xp*dppn)) - at(zp*yp/(xp*dppp)))
    d[2] = (lg((dnnn-zn)/(dnnp-zp)) - lg((dpnn-zn)/(dpnp-zp)) - lg((dnpn-zn)/(dnpp-zp))
            + lg((dppn-zn)/(dppp-zp)))
    d[5] = (at(zn*xn/(yn*dnnn)) - at(zp*xn/(yn*dnnp)) - at(zn*xp/(yn*dpnn)) + at(zp*xp/(yn*dpnp))
            - at(zn*xn/(yp*dnpn)) + at(zp*xn/(yp*dnpp)) + at(zn*xp/(yp*dppn)) - at(zp*xp/(yp*dppp)))
    d[9] = (at(xn*yn/(zn*dnnn)) - at(xp*yn/(zn*dpnn)) - at(xn*yp/(zn*dnpn)) + at(xp*yp/(zn*dppn))
            - at(xn*yn/(zp*dnnp)) + at(xp*yn/(zp*dpnp)) + at(xn*yp/(zp*dnpp)) - at(xp*yp/(zp*dppp)))
    return d


def tensor_grid(thick, dx, dy):
    """xx, yy, zz, xy at integer displacements dx, dy (arrays)."""
    a, b = 0.49999, 0.5 * thick
    acc = {k: 0.0 for k in (1, 2, 5, 9)}
    for jy in range(-4, 5):
        for ix in range(-4, 5):
            d = ctt(b, a, dx + 0.1 * ix, dy + 0.1 * jy)
            for k in acc:
                acc[k] = acc[k] + d[k]
    return acc[1] / 81.0, acc[5] / 81.0, acc[9] / 81.0, acc[2] / 81.0


def demag_periodic(y, nx, ny, thick, strength):
    """Reference operator: circular convolution, centred tensor, shifted."""
    i = np.arange(nx) - nx // 2
    j = np.arange(ny) - ny // 2
    DX, DY = np.meshgrid(i.astype(float), j.astype(float))
    xx, yy, zz, xy = tensor_grid(thick, DX, DY)
    m = y.reshape(3, ny, nx)
    F = [np.fft.fft2(t) for t in (xx, xy, yy, zz)]
    M = [np.fft.fft2(m[c]) for c in range(3)]
    hx = np.real(np.fft.ifft2(F[0] * M[0] + F[1] * M[1]))
    hy = np.real(np.fft.ifft2(F[1] * M[0] + F[2] * M[1]))
    hz = np.real(np.fft.ifft2(F[3] * M[2]))
    out = np.stack([np.roll(np.roll(h, -(ny // 2), 0), -(nx // 2), 1) for h in (hx, hy, hz)])
    return strength * out.reshape(-1)


def demag_padded(y, nx, ny, thick, strength):
    """Open-boundary direct linear sum."""
    m = y.reshape(3, ny, nx)
    DX, DY = np.meshgrid(np.arange(-nx + 1, nx).astype(float), np.arange(-ny + 1, ny).astype(float))
    xx, yy, zz, xy = tensor_grid(thick, DX, DY)
    h = np.zeros((3, ny, nx))
    for tj in range(ny):
        for ti in range(nx):
            sl = (slice(tj + ny - 1, tj - 1 if tj > 0 else None, -1),
                  slice(ti + nx - 1, ti - 1 if ti > 0 else None, -1))
            kxx, kyy, kzz, kxy = xx[sl], yy[sl], zz[sl], xy[sl]
            h[0, tj, ti] = np.sum(kxx * m[0] + kxy * m[1])
            h[1, tj, ti] = np.sum(kxy * m[0] + kyy * m[1])
            h[2, tj, ti] = np.sum(kzz * m[2])
    return strength * h.reshape(-1)


# ---------------------------------------------------------------- physics
def nbr(a, ng, ny, neumann):
    """left, right, up, down neighbour values of a (ny, ng) plane."""
    if neumann:
        L = np.concatenate([a[:, :1], a[:, :-1]], 1)
        R = np.concatenate([a[:, 1:], a[:, -1:]], 1)
        U = np.concatenate([a[:1, :], a[:-1, :]], 0)
        D = np.concatenate([a[1:, :], a[-1:, :]], 0)
    else:
        L, R = np.roll(a, 1, 1), np.roll(a, -1, 1)
        U, D = np.roll(a, 1, 0), np.roll(a, -1, 0)
    return L, R, U, D


def field(m, v, P, neumann, exch_self):
    """h and (if v is given) its directional derivative dh for terms P."""
    ng = m.shape[2]
    ny = m.shape[1]
    msk, nsk = np.array(P["msk"]), np.array(P["nsk"])
    mdot = sum(m[c] * msk[c] for c in range(3))
    h, dh = [], []
    vdot = sum(v[c] * msk[c] for c in range(3)) if v is not None else None
    for a in range(3):
        L, R, U, D = nbr(m[a], ng, ny, neumann)
        ex = L + R + U + D
        if exch_self:
            ex = ex - 4.0 * m[a]
        if P["aniso"] == "cubic":
            an = msk[a] * P["chk"] * m[a] * (m[a] * m[a] - 1.0)
        else:
            an = msk[a] * (P["chk"] * mdot + P["cha"])
        h.append(P["happ"][a] + P["che"] * ex + an + P["chb"] * nsk[a] * (L + R))
        if v is not None:
            vL, vR, vU, vD = nbr(v[a], ng, ny, neumann)
            vex = vL + vR + vU + vD
            if exch_self:
                vex = vex - 4.0 * v[a]
            if P["aniso"] == "cubic":
                dan = msk[a] * P["chk"] * (3.0 * m[a] ** 2 - 1.0) * v[a]
            else:
                dan = msk[a] * P["chk"] * vdot
            dh.append(P["che"] * vex + dan + P["chb"] * nsk[a] * (vL + vR))
    return np.array(h), (np.array(dh) if v is not None else None)


def llg(m, h, alpha, chg):
    mh = (m * h).sum(0)
    return np.array([chg * (m[2] * h[1] - m[1] * h[2]) + alpha * (h[0] - mh * m[0]),
                     chg * (m[0] * h[2] - m[2] * h[0]) + alpha * (h[1] - mh * m[1]),
                     chg * (m[1] * h[0] - m[0] * h[1]) + alpha * (h[2] - mh * m[2])])


def jv(m, v, h, dh, alpha, chg):
    mh = (m * h).sum(0)
    dmh = (v * h).sum(0) + (m * dh).sum(0)
    return np.array([
        chg * (v[2] * h[1] + m[2] * dh[1] - v[1] * h[2] - m[1] * dh[2]) + alpha * (dh[0] - dmh * m[0] - mh * v[0]),
        chg * (v[0] * h[2] + m[0] * dh[2] - v[2] * h[0] - m[2] * dh[0]) + alpha * (dh[1] - dmh * m[1] - mh * v[1]),
        chg * (v[1] * h[0] + m[1] * dh[0] - v[0] * h[1] - m[0] * dh[1]) + alpha * (dh[2] - dmh * m[2] - mh * v[2])])


def pinv(m, h, P, n0, gamma, alpha, chg, neumann, exch_self):
    ny, ng = m.shape[1:]
    msk = np.array(P["msk"])
    k = []
    gx, gy = np.meshgrid(np.arange(ng), np.arange(ny))
    ghosts = ((gx == 0).astype(float) + (gx == ng - 1) + (gy == 0) + (gy == ny - 1)) if neumann else 0.0
    self_ex = P["che"] * (ghosts - (4.0 if exch_self else 0.0)) if (neumann or exch_self) else 0.0
    for a in range(3):
        if P["aniso"] == "cubic":
            dan = P["chk"] * (3.0 * m[a] ** 2 - 1.0) * msk[a]
        else:
            dan = msk[a] * msk[a] * P["chk"] + 0.0 * m[a]
        k.append(dan + n0[a] + self_ex)
    k = np.array(k)
    e = h + m * k
    mh = (m * h).sum(0)
    m1, m2, m3 = m
    h1, h2, h3 = h
    k1, k2, k3 = k
    e1, e2, e3 = e
    J = np.array([
        [alpha * (k1 - e1 * m1 - mh), chg * (m3 * k2 - h3) - alpha * e2 * m1, chg * (h2 - m2 * k3) - alpha * e3 * m1],
        [chg * (h3 - m3 * k1) - alpha * e1 * m2, alpha * (k2 - e2 * m2 - mh), chg * (m1 * k3 - h1) - alpha * e3 * m2],
        [chg * (m2 * k1 - h2) - alpha * e1 * m3, chg * (h1 - m1 * k2) - alpha * e2 * m3, alpha * (k3 - e3 * m3 - mh)]])
    A = np.eye(3)[:, :, None, None] - gamma * J
    Am = np.moveaxis(A, (0, 1), (2, 3))       # (ny, ng, 3, 3)
    Pi = np.linalg.inv(Am)
    return np.moveaxis(Pi, (2, 3), (0, 1))    # (3, 3, ny, ng)


def case(name, ng, ny, P_rhs, P_jv, P_pc, alpha, chg, active, y, neumann=False, exch_self=False,
         normalize=None, demag=None, gate=True, compact=False, seed=0):
    rng = np.random.default_rng(seed)
    N = ng * ny
    y = y.copy()
    m = y.reshape(3, ny, ng).copy()
    if normalize == "exact":
        n = np.sqrt((m ** 2).sum(0))
        m = np.where(n > 1e-30, m / np.where(n > 1e-30, n, 1.0), m)
    elif normalize == "reg":
        n = np.sqrt((m ** 2).sum(0))
        m = m / (n + 1e-3)
    if demag is None:
        hd = np.zeros((3, ny, ng))
        n0 = (0.0, 0.0, 0.0)
    else:
        kind, strength = demag
        flat = m.reshape(-1)
        hd = (demag_periodic(flat, ng, ny, 1.0, strength) if kind == "periodic"
              else demag_padded(flat, ng, ny, 1.0, strength)).reshape(3, ny, ng)
        t = tensor_grid(1.0, np.array(0.0), np.array(0.0))
        n0 = (t[0] * strength, t[1] * strength, t[2] * strength)
    g = active.reshape(ny, ng).astype(float)
    h, _ = field(m, None, P_rhs, neumann, exch_self)
    f = llg(m, h + hd, alpha, chg)
    f = f * g if (gate or compact) else f
    v = rng.standard_normal((3, ny, ng))
    hj, dhj = field(m, v, P_jv, neumann, exch_self)
    Jv = jv(m, v, hj, dhj, alpha, chg)
    Jv = Jv * g if (gate or compact) else Jv
    gamma = 0.0371
    hp, _ = field(m, None, P_pc, neumann, exch_self)
    Pi = pinv(m, hp + hd, P_pc, n0, gamma, alpha, chg, neumann, exch_self)
    r = rng.standard_normal((3, ny, ng))
    z = np.einsum("abij,bij->aij", Pi, r)
    z = z * g if (gate or compact) else z
    np.savez_compressed(OUT / f"{name}.npz", y=y, y_norm=m.reshape(-1), active=active,
                        f=f.reshape(-1), hd=hd.reshape(-1), v=v.reshape(-1), Jv=Jv.reshape(-1),
                        gamma=gamma, Pinv=Pi.reshape(9, -1).reshape(-1), r=r.reshape(-1),
                        z=z.reshape(-1), ng=ng, ny=ny)
    print("wrote", name)


# ------------------------------------------------------- i3 polycrystal
# Written straight from the i3 kernels (src/irregular/2d_fft/2d_i3/), term by
# term, without the shared `field` helper above: the tiled RHS's LDS tile of
# w*m, its per-cell axes and DMI and the w_self output gate
# (2d_fft.cu:451-627), the regularised normalisation (2d_fft.cu:418-428),
# the windowed demag gather (demag_fft.cu:125-139), the fixed-axis Jv
# (jtv.cu:43-96) and the block-Jacobi build and apply (precond.cu:108-241,
# 244-268; anisotropy axis and DMI direction fixed at x, no gate).
def i3_fixture(name, ng, ny, seed):
    rng = np.random.default_rng(seed)
    N = ng * ny
    c_chk, c_che, c_alpha, c_chg, c_cha, c_chb = 1.0, 4.0, 0.2, 1.0, 0.0, 0.6   # 2d_fft.cu:176-181
    j_chk, j_che, j_alpha, j_chg, j_cha, j_chb = 4.0, 4.0, 0.2, 1.0, 0.0, 0.3   # jtv.cu:30-35
    p_chk, p_che, p_alpha, p_chg, p_cha, p_chb = 4.0, 4.0, 0.2, 1.0, 0.0, 0.3   # precond.cu:58-63
    strength = 1.0
    w = rng.uniform(1e-3, 1.0, N)
    w[rng.random(N) < 0.3] = 1.0
    w[rng.random(N) < 0.1] = 1e-3
    def unit(k):
        a = rng.standard_normal((3, N))
        return (a / np.linalg.norm(a, axis=0)).reshape(-1)
    msk, nsk = unit(0), unit(1)
    y = rng.standard_normal(3 * N)

    # normalize_m_kernel: m / (|m| + 1e-3), in place on y
    m = y.reshape(3, N)
    ymp = np.sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])
    inv = 1.0 / (ymp + 1.0e-3)
    yn = np.concatenate([m[0] * inv, m[1] * inv, m[2] * inv])

    # Demag_ApplyWindowed: gather w*y, then the periodic pipeline
    wy = np.concatenate([w * yn[:N], w * yn[N:2 * N], w * yn[2 * N:]])
    hd = demag_periodic(wy, ng, ny, 1.0, strength)
    t0 = tensor_grid(1.0, np.array(0.0), np.array(0.0))
    nxx0, nyy0, nzz0 = t0[0] * strength, t0[1] * strength, t0[2] * strength

    def wrapped(a):   # a (ny, ng) -> left, right, up, down with periodic wrap
        return np.roll(a, 1, 1), np.roll(a, -1, 1), np.roll(a, 1, 0), np.roll(a, -1, 0)

    M = [yn[k * N:(k + 1) * N].reshape(ny, ng) for k in range(3)]
    W = w.reshape(ny, ng)
    MS = [msk[k * N:(k + 1) * N].reshape(ny, ng) for k in range(3)]
    NS = [nsk[k * N:(k + 1) * N].reshape(ny, ng) for k in range(3)]
    HD = [hd[k * N:(k + 1) * N].reshape(ny, ng) for k in range(3)]
    # tile of w*m and its neighbours
    T = [W * M[k] for k in range(3)]
    lx1, rx1, ux1, dx1 = wrapped(T[0])
    lx2, rx2, ux2, dx2 = wrapped(T[1])
    lx3, rx3, ux3, dx3 = wrapped(T[2])
    m1, m2, m3 = M
    mdotmsk = m1 * MS[0] + m2 * MS[1] + m3 * MS[2]
    aniso_factor = c_chk * mdotmsk + c_cha
    h1 = c_che * (lx1 + rx1 + ux1 + dx1) + MS[0] * aniso_factor + c_chb * NS[0] * (lx1 + rx1) + HD[0]
    h2 = c_che * (lx2 + rx2 + ux2 + dx2) + MS[1] * aniso_factor + c_chb * NS[1] * (lx2 + rx2) + HD[1]
    h3 = c_che * (lx3 + rx3 + ux3 + dx3) + MS[2] * aniso_factor + c_chb * NS[2] * (lx3 + rx3) + HD[2]
    mh = m1 * h1 + m2 * h2 + m3 * h3
    f = np.concatenate([
        (W * (c_chg * (m3 * h2 - m2 * h3) + c_alpha * (h1 - mh * m1))).ravel(),
        (W * (c_chg * (m1 * h3 - m3 * h1) + c_alpha * (h2 - mh * m2))).ravel(),
        (W * (c_chg * (m2 * h1 - m1 * h2) + c_alpha * (h3 - mh * m3))).ravel()])

    # jtv_kernel (unwindowed neighbours, anisotropy and DMI along x)
    v = rng.standard_normal(3 * N)
    V = [v[k * N:(k + 1) * N].reshape(ny, ng) for k in range(3)]
    yL1, yR1, yU1, yD1 = wrapped(m1)
    yL2, yR2, yU2, yD2 = wrapped(m2)
    yL3, yR3, yU3, yD3 = wrapped(m3)
    v1L, v1R, v1U, v1D = wrapped(V[0])
    v2L, v2R, v2U, v2D = wrapped(V[1])
    v3L, v3R, v3U, v3D = wrapped(V[2])
    v1, v2, v3 = V
    che_chb = j_che + j_chb
    jh1 = che_chb * (yL1 + yR1) + j_che * (yU1 + yD1) + j_chk * m1 + j_cha
    jh2 = j_che * (yL2 + yR2 + yU2 + yD2)
    jh3 = j_che * (yL3 + yR3 + yU3 + yD3)
    jmh = m1 * jh1 + m2 * jh2 + m3 * jh3
    dh1 = che_chb * (v1L + v1R) + j_che * (v1U + v1D) + j_chk * v1
    dh2 = j_che * (v2L + v2R + v2U + v2D)
    dh3 = j_che * (v3L + v3R + v3U + v3D)
    dmh = (v1 * jh1 + v2 * jh2 + v3 * jh3) + (m1 * dh1 + m2 * dh2 + m3 * dh3)
    Jv = np.concatenate([
        (j_chg * (v3 * jh2 + m3 * dh2 - v2 * jh3 - m2 * dh3) + j_alpha * (dh1 - dmh * m1 - jmh * v1)).ravel(),
        (j_chg * (v1 * jh3 + m1 * dh3 - v3 * jh1 - m3 * dh1) + j_alpha * (dh2 - dmh * m2 - jmh * v2)).ravel(),
        (j_chg * (v2 * jh1 + m2 * dh1 - v1 * jh2 - m1 * dh2) + j_alpha * (dh3 - dmh * m3 - jmh * v3)).ravel()])

    # build_J_kernel: h with the last f()'s demag field, k = (chk + N_xx(0), N_yy(0), N_zz(0))
    gamma = 0.0413
    ph1 = p_che * (yL1 + yR1 + yU1 + yD1) + (p_chk * m1 + p_cha) + p_chb * (yL1 + yR1) + HD[0]
    ph2 = p_che * (yL2 + yR2 + yU2 + yD2) + HD[1]
    ph3 = p_che * (yL3 + yR3 + yU3 + yD3) + HD[2]
    k1, k2, k3 = p_chk + nxx0, nyy0, nzz0
    e1, e2, e3 = ph1 + m1 * k1, ph2 + m2 * k2, ph3 + m3 * k3
    pmh = m1 * ph1 + m2 * ph2 + m3 * ph3
    J = np.array([
        [p_alpha * (k1 - e1 * m1 - pmh), p_chg * (m3 * k2 - ph3) - p_alpha * e2 * m1,
         p_chg * (ph2 - m2 * k3) - p_alpha * e3 * m1],
        [p_chg * (ph3 - m3 * k1) - p_alpha * e1 * m2, p_alpha * (k2 - e2 * m2 - pmh),
         p_chg * (m1 * k3 - ph1) - p_alpha * e3 * m2],
        [p_chg * (m2 * k1 - ph2) - p_alpha * e1 * m3, p_chg * (ph1 - m1 * k2) - p_alpha * e2 * m3,
         p_alpha * (k3 - e3 * m3 - pmh)]])
    A = np.eye(3)[:, :, None, None] - gamma * J
    Pi = np.moveaxis(np.linalg.inv(np.moveaxis(A, (0, 1), (2, 3))), (2, 3), (0, 1))
    r = rng.standard_normal(3 * N)
    z = np.einsum("abij,bij->aij", Pi, r.reshape(3, ny, ng)).reshape(-1)
    np.savez_compressed(OUT / f"{name}.npz", y=y, y_norm=yn, active=np.ones(N, np.uint8), w=w,
                        msk=msk, nsk=nsk, f=f, hd=hd, v=v, Jv=Jv, gamma=gamma,
                        Pinv=Pi.reshape(9, -1).reshape(-1), r=r, z=z, ng=ng, ny=ny)
    print("wrote", name)


def square_hole(ng, ny, cxf, cyf, hf):
    cx, cy, hs = cxf * (ng - 1), cyf * (ny - 1), hf * ny
    I, J = np.meshgrid(np.arange(ng), np.arange(ny))
    hole = (np.abs(I - cx) <= hs) & (np.abs(J - cy) <= hs)
    return (~hole).astype(np.uint8).reshape(-1)


def ellipse(ng, ny, rxf, ryf):
    I, J = np.meshgrid(np.arange(ng), np.arange(ny))
    dx, dy = I + 0.5 - 0.5 * ng, J + 0.5 - 0.5 * ny
    q = dx * dx / (rxf * ng) ** 2 + dy * dy / (ryf * ny) ** 2
    return (q <= 1.0).astype(np.uint8).reshape(-1)


def main():
    rng = np.random.default_rng(12345)
    # i4: square hole, cubic anisotropy, Zeeman, periodic demag 4.0
    ng, ny = 16, 12
    act = square_hole(ng, ny, 0.5, 0.5, 0.2)
    y = rng.standard_normal(3 * ng * ny) * np.tile(act, 3)
    cub = dict(che=20.0, chk=1.0, cha=0.0, chb=0.0, msk=(1, 1, 1), nsk=(1, 0, 0), aniso="cubic")
    case("i4_16x12", ng, ny, dict(cub, happ=(-0.2, 0, 0)), dict(cub, happ=(0, 0, 0)),
         dict(cub, happ=(-1.0, 0, 0)), 0.2, 1.0, act, y, demag=("periodic", 4.0))
    # i4 physics with the padded (open-boundary) demag
    case("i4pad_12x10", 12, 10, dict(cub, happ=(-0.2, 0, 0)), dict(cub, happ=(0, 0, 0)),
         dict(cub, happ=(-1.0, 0, 0)), 0.2, 1.0, square_hole(12, 10, 0.5, 0.5, 0.2),
         rng.standard_normal(360) * np.tile(square_hole(12, 10, 0.5, 0.5, 0.2), 3),
         demag=("padded", 4.0))
    # i1: linear anisotropy + DMI along x, exact normalisation, demag 2.0
    ng, ny = 14, 10
    act = np.ones(ng * ny, np.uint8)
    lin = dict(che=4.0, chk=4.0, cha=0.0, chb=0.3, msk=(1, 0, 0), nsk=(1, 0, 0), aniso="linear",
               happ=(0, 0, 0))
    case("i1_14x10", ng, ny, lin, lin, lin, 0.2, 1.0, act, rng.standard_normal(3 * ng * ny),
         normalize="exact", demag=("periodic", 2.0), gate=False)
    # i6: compact ellipse, cubic, che 10, demag 4.0
    ng, ny = 16, 16
    act = ellipse(ng, ny, 0.3, 0.3)
    cub6 = dict(cub, che=10.0, happ=(0, 0, 0))
    case("i6_16x16", ng, ny, cub6, cub6, cub6, 0.2, 1.0, act,
         rng.standard_normal(3 * ng * ny) * np.tile(act, 3), demag=("periodic", 4.0), compact=True)
    # plain form: Neumann Laplacian, uniaxial z, LL prefactor
    a = 0.02
    pl = dict(che=1.0, chk=0.1, cha=0.0, chb=0.0, msk=(0, 0, 1), nsk=(0, 0, 0), aniso="linear",
              happ=(0, 0, -0.05))
    ng, ny = 9, 7
    case("plain_9x7", ng, ny, pl, pl, pl, a / (1 + a * a), 1 / (1 + a * a),
         np.ones(ng * ny, np.uint8), rng.standard_normal(3 * ng * ny), neumann=True,
         exch_self=True, gate=False)

    # i3: windowed polycrystal (per-cell axes, DMI, w gate, windowed demag)
    i3_fixture("i3_14x10", 14, 10, seed=31)

    # Radau trajectory of the plain form on 6x5 (tight tolerance)
    from scipy.integrate import solve_ivp
    ng, ny = 6, 5
    N = ng * ny
    th = 0.1 * rng.random(N)
    ph = 2 * np.pi * rng.random(N)
    y0 = np.concatenate([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    alpha, chg = a / (1 + a * a), 1 / (1 + a * a)

    def rhs(t, yy):
        m = yy.reshape(3, ny, ng)
        h, _ = field(m, None, pl, True, True)
        return llg(m, h, alpha, chg).reshape(-1)

    ts = np.array([0.5, 1.0, 2.0, 4.0])
    sol = solve_ivp(rhs, (0, ts[-1]), y0, method="Radau", rtol=1e-11, atol=1e-13, t_eval=ts)
    np.savez_compressed(OUT / "plain_traj_6x5.npz", y0=y0, t=ts, y=sol.y.T, ng=ng, ny=ny)
    print("wrote plain_traj_6x5")


if __name__ == "__main__":
    main()
