"""The reference's whole-block MHD kernels, restated in NumPy FP64 (oracle only).

Each function follows the scalar Pyccel kernel it names, operation for operation
(same expression trees, same branch tests), vectorised with ``np.where``.  Test
infrastructure: see the package docstring.
"""

from __future__ import annotations

import numpy as np

#: Relative size below which the HLLD star denominator counts as zero
#: (``mhd.py:138``, ``mhd_sweep_pyccel.py:127``).
DEGENERATE = 1.0e-12
#: Limiter codes of the sweep (``kernels/mhd_sweep/__init__.py:13``).
LIMITER_CODES = ("minmod", "vanleer", "mc", "none")
_BIG = 1.7976931348623157e308


def limited_slope(a: np.ndarray, b: np.ndarray, limiter: int) -> np.ndarray:
    """Limited slope from backward/forward differences (``mhd_sweep_pyccel.py:25-48``)."""
    if limiter == 3:
        return 0.5 * (a + b)
    with np.errstate(all="ignore"):
        positive = a * b > 0.0
        if limiter == 0:
            slope = np.where(np.abs(a) < np.abs(b), a, b)
        elif limiter == 1:
            product = a * b
            slope = 2.0 * product / np.where(positive, a + b, 1.0)
        else:
            central = 0.5 * (a + b)
            limit = np.where(np.abs(a) < np.abs(b), 2.0 * a, 2.0 * b)
            size = np.minimum(np.abs(central), np.abs(limit))
            slope = np.where(central > 0.0, size, np.where(central < 0.0, -size, 0.0))
    return np.where(positive, slope, 0.0)


def fast_speed(rho, p, bn, bt1, bt2, gamma):
    """Fast magnetosonic speed of the sweep (``mhd_sweep_pyccel.py:51-61``)."""
    a2 = gamma * p / rho
    b2 = (bn * bn + bt1 * bt1 + bt2 * bt2) / rho
    disc = (a2 + b2) * (a2 + b2) - 4.0 * a2 * bn * bn / rho
    root = np.sqrt(np.where(disc > 0.0, disc, 0.0))
    return np.sqrt(0.5 * (a2 + b2 + root))


def fast_speed_rate(rho, p, bn, bt1, bt2, gamma):
    """Fast speed of the CFL kernel, ``(bn * bn)`` grouped (``mhd_rate_pyccel.py:17-27``)."""
    a2 = gamma * p / rho
    b2 = (bn * bn + bt1 * bt1 + bt2 * bt2) / rho
    disc = (a2 + b2) * (a2 + b2) - 4.0 * a2 * (bn * bn) / rho
    root = np.sqrt(np.where(disc > 0.0, disc, 0.0))
    return np.sqrt(0.5 * (a2 + b2 + root))


def _side(w, bn, gamma):
    """Conserved state, flux and total pressure of one side (``mhd_sweep_pyccel.py:64-98``)."""
    rho, vn, vt1, vt2, p, bt1, bt2 = w
    magnetic = 0.5 * (bn * bn + bt1 * bt1 + bt2 * bt2)
    energy = p / (gamma - 1.0) + 0.5 * rho * (vn * vn + vt1 * vt1 + vt2 * vt2) + magnetic
    pt = p + magnetic
    vb = vn * bn + vt1 * bt1 + vt2 * bt2
    u = [rho, rho * vn, rho * vt1, rho * vt2, energy, bt1, bt2]
    f = [
        rho * vn,
        rho * vn * vn + pt - bn * bn,
        rho * vn * vt1 - bn * bt1,
        rho * vn * vt2 - bn * bt2,
        (energy + pt) * vn - bn * vb,
        bt1 * vn - bn * vt1,
        bt2 * vn - bn * vt2,
    ]
    return u, f, pt


def _star(w, bn, s, d, pt, sm, pts, energy0):
    """HLLD star state beside the contact (``mhd_sweep_pyccel.py:101-145``)."""
    vn = w[1]
    rho_s = d / (s - sm)
    den = d * (s - sm) - bn * bn
    keep = ~(np.abs(den) >= DEGENERATE * (pt + bn * bn))
    safe = np.where(keep, 1.0, den)
    vt1 = np.where(keep, w[2], w[2] - bn * w[5] * (sm - vn) / safe)
    vt2 = np.where(keep, w[3], w[3] - bn * w[6] * (sm - vn) / safe)
    bt1 = np.where(keep, w[5], w[5] * (d * (s - vn) - bn * bn) / safe)
    bt2 = np.where(keep, w[6], w[6] * (d * (s - vn) - bn * bn) / safe)
    vb = vn * bn + w[2] * w[5] + w[3] * w[6]
    vb_s = sm * bn + vt1 * bt1 + vt2 * bt2
    energy = ((s - vn) * energy0 - pt * vn + pts * sm + bn * (vb - vb_s)) / (s - sm)
    us = [rho_s, rho_s * sm, rho_s * vt1, rho_s * vt2, energy, bt1, bt2]
    return us, (vt1, vt2, bt1, bt2)


#: The sign of the rotational-discontinuity term in the double-star energy of the left
#: and right states (``mhd_sweep_pyccel.py:233-238``); a module constant so that a test
#: can plant a sign error there.
DOUBLE_STAR_SENSE = (-1.0, 1.0)


def hlld_flux(left, right, bn, gamma) -> np.ndarray:
    """HLLD flux of the faces between ``left`` and ``right`` (``mhd_sweep_pyccel.py:148-251``).

    ``left``/``right``: ``(7, ...)`` normal-frame primitives (rho, v_n, v_t1, v_t2,
    p, B_t1, B_t2); ``bn`` the face-normal field.  Returns ``(7, ...)`` fluxes.
    """
    left = np.asarray(left, dtype=np.float64)
    right = np.asarray(right, dtype=np.float64)
    bn = np.asarray(bn, dtype=np.float64)
    with np.errstate(all="ignore"):
        ul, fl, ptl = _side(left, bn, gamma)
        ur, fr, ptr = _side(right, bn, gamma)
        rl, vl = left[0], left[1]
        rr, vr = right[0], right[1]
        cf = np.maximum(
            fast_speed(rl, left[4], bn, left[5], left[6], gamma),
            fast_speed(rr, right[4], bn, right[5], right[6], gamma),
        )
        sl = np.minimum(vl, vr) - cf
        sr = np.maximum(vl, vr) + cf
        dl = (sl - vl) * rl
        dr = (sr - vr) * rr
        sm = (dr * vr - dl * vl - ptr + ptl) / (dr - dl)
        pts = (dr * ptl - dl * ptr + dl * dr * (vr - vl)) / (dr - dl)
        usl, (vl1, vl2, bl1, bl2) = _star(left, bn, sl, dl, ptl, sm, pts, ul[4])
        usr, (vr1, vr2, br1, br2) = _star(right, bn, sr, dr, ptr, sm, pts, ur[4])
        sql = np.sqrt(usl[0])
        sqr = np.sqrt(usr[0])
        sign = np.where(bn > 0.0, 1.0, np.where(bn < 0.0, -1.0, 0.0))
        vt2_1 = (sql * vl1 + sqr * vr1 + (br1 - bl1) * sign) / (sql + sqr)
        vt2_2 = (sql * vl2 + sqr * vr2 + (br2 - bl2) * sign) / (sql + sqr)
        bt2_1 = (sql * br1 + sqr * bl1 + sql * sqr * (vr1 - vl1) * sign) / (sql + sqr)
        bt2_2 = (sql * br2 + sqr * bl2 + sql * sqr * (vr2 - vl2) * sign) / (sql + sqr)
        vb2 = sm * bn + vt2_1 * bt2_1 + vt2_2 * bt2_2
        ssl = sm - np.abs(bn) / sql
        ssr = sm + np.abs(bn) / sqr

        def double_star(us, vt1, vt2, bt1, bt2, sq, sense):
            vb_s = sm * bn + vt1 * bt1 + vt2 * bt2
            rho_s = us[0]
            return [
                rho_s,
                rho_s * sm,
                rho_s * vt2_1,
                rho_s * vt2_2,
                us[4] + sense * sq * (vb_s - vb2) * sign,
                bt2_1,
                bt2_2,
            ]

        uddl = double_star(usl, vl1, vl2, bl1, bl2, sql, DOUBLE_STAR_SENSE[0])
        uddr = double_star(usr, vr1, vr2, br1, br2, sqr, DOUBLE_STAR_SENSE[1])
        out = []
        for c in range(7):
            star_l = fl[c] + sl * (usl[c] - ul[c])
            star_r = fr[c] + sr * (usr[c] - ur[c])
            value = np.where(
                ssl >= 0.0,
                star_l,
                np.where(
                    sm >= 0.0,
                    fl[c] + sl * (usl[c] - ul[c]) + ssl * (uddl[c] - usl[c]),
                    np.where(
                        ssr >= 0.0,
                        fr[c] + sr * (usr[c] - ur[c]) + ssr * (uddr[c] - usr[c]),
                        star_r,
                    ),
                ),
            )
            value = np.where(sr < 0.0, fr[c], value)
            value = np.where(sl >= 0.0, fl[c], value)
            out.append(value)
    return np.stack(np.broadcast_arrays(*out))


def sweep(prim: np.ndarray, bn: np.ndarray, axis: int, limiter: int, gamma: float):
    """MUSCL + HLLD fluxes along one axis of a haloed block (``mhd_sweep_pyccel.py:254-319``).

    ``prim`` is ``(8, X, Y[, Z])``; ``axis`` the sweep axis (0, 1 or 2).  In 2D the
    normal frame takes rows ``[0, 1+a, 2-a, 3, 4, 6-a, 7]`` (``:268-289``); in 3D the
    rows of the cyclic frame (``frame_rows``, ``mhd3d.py:511-528``).  ``bn`` is the
    face-normal field with ``n+1`` faces along ``axis``.  Returns the fluxes in the
    physical layout, ``n-3`` faces along ``axis`` (face ``k`` between cells ``k+1``
    and ``k+2``) and every transverse cell.
    """
    if prim.ndim == 3:
        rows = [0, 1 + axis, 2 - axis, 3, 4, 6 - axis, 7]
    else:
        rows = frame_rows(axis)
    q = np.moveaxis(prim[rows], 1 + axis, 1)
    bnm = np.moveaxis(bn, axis, 0)
    n = q.shape[1]
    with np.errstate(all="ignore"):
        slope = np.zeros_like(q)
        slope[:, 1 : n - 1] = limited_slope(
            q[:, 1 : n - 1] - q[:, 0 : n - 2], q[:, 2:n] - q[:, 1 : n - 1], limiter
        )
        left = q[:, 1 : n - 2] + 0.5 * slope[:, 1 : n - 2]
        right = q[:, 2 : n - 1] - 0.5 * slope[:, 2 : n - 1]
        bad = (left[0] <= 0.0) | (left[4] <= 0.0) | (right[0] <= 0.0) | (right[4] <= 0.0)
        left = np.where(bad, q[:, 1 : n - 2], left)
        right = np.where(bad, q[:, 2 : n - 1], right)
    flux = hlld_flux(left, right, bnm[2 : n - 1], gamma)
    return np.ascontiguousarray(np.moveaxis(flux, 1, 1 + axis))


def frame_rows(axis: int) -> list[int]:
    """3D sweep rows in the cyclic frame (s, b, c): rho, v_s, v_b, v_c, p, B_b, B_c.

    ``mhd3d.py:517-519`` reorders prim to (rho, v_s, v_b, v_c, p, B_s, B_b, B_c) and
    the 2D sweep's x-rows ``[0, 1, 2, 3, 4, 6, 7]`` then pick the tangential field.
    """
    b, c = (axis + 1) % 3, (axis + 2) % 3
    return [0, 1 + axis, 1 + b, 1 + c, 4, 5 + b, 5 + c]


def _upwind(mass, positive, negative):
    """``ct_contact_pyccel.py:16-23``."""
    return np.where(
        mass > 0.0, positive, np.where(mass < 0.0, negative, 0.5 * (positive + negative))
    )


def ct_contact(fx_by, gy_bx, fx_rho, gy_rho, ecc) -> np.ndarray:
    """Nodal ``E_z`` by the CT-contact average (``ct_contact_pyccel.py:26-61``).

    Over the last two axes (leading axes are a batch, e.g. the planes of a 3D
    block): ``fx_*`` are x-face fluxes ``(X-3, Y)``, ``gy_*`` y-face fluxes
    ``(X, Y-3)``, ``ecc`` the cell field ``(X, Y)``; returns ``(X-3, Y-3)`` =
    ``(nx+1, ny+1)``.
    """
    X, Y = ecc.shape[-2:]
    na, nb = X - 3, Y - 3
    i = np.arange(na)[:, None] + 2
    j = np.arange(nb)[None, :] + 2
    fxb_lower = fx_by[..., i - 2, j - 1]
    fxb_upper = fx_by[..., i - 2, j]
    gyb_left = gy_bx[..., i - 1, j - 2]
    gyb_right = gy_bx[..., i, j - 2]

    def e(di, dj):
        return ecc[..., i + di, j + dj]

    dl2 = _upwind(fx_rho[..., i - 2, j - 1], gyb_left - e(-1, -1), gyb_right - e(0, -1))
    dr2 = _upwind(fx_rho[..., i - 2, j], gyb_left - e(-1, 0), gyb_right - e(0, 0))
    dl1 = _upwind(gy_rho[..., i - 1, j - 2], -fxb_lower - e(-1, -1), -fxb_upper - e(-1, 0))
    dr1 = _upwind(gy_rho[..., i, j - 2], -fxb_lower - e(0, -1), -fxb_upper - e(0, 0))
    return 0.25 * (gyb_right + gyb_left - fxb_upper - fxb_lower + dl1 + dr1 + dl2 + dr2)


def _finite(x):
    """``mhd_primitive_pyccel.py:18-20``."""
    return np.abs(x) <= _BIG


def primitive(rho, mx, my, mz, energy, bz, az, dx, dy, b0x, b0y, gamma):
    """Primitive variables, face fields, cell E_z, bad cells (``mhd_primitive_pyccel.py:23-87``).

    Returns ``(prim (8,X,Y), bxf (X+1,Y), byf (X,Y+1), ecc (X,Y), bad)``.
    """
    with np.errstate(all="ignore"):
        bxf = (az[:, 1:] - az[:, :-1]) / dy + b0x
        byf = -(az[1:, :] - az[:-1, :]) / dx + b0y
        r = rho
        u = mx / r
        v = my / r
        w = mz / r
        bx = 0.5 * (bxf[:-1] + bxf[1:])
        by = 0.5 * (byf[:, :-1] + byf[:, 1:])
        b = bz
        p = (gamma - 1.0) * (
            energy - 0.5 * r * (u * u + v * v + w * w) - 0.5 * (bx * bx + by * by + b * b)
        )
        prim = np.stack([r, u, v, w, p, bx, by, b])
        ecc = bx * v - by * u
        positive = (r > 0.0) & (p > 0.0)
        finite = np.all(_finite(prim), axis=0)
    bad = int(np.count_nonzero(~positive | ~finite))
    return prim, bxf, byf, ecc, bad


def rate(rho, mx, my, mz, energy, bz, az, halo, dx, dy, b0x, b0y, gamma):
    """Largest fast-wave rate and smallest density of the block (``mhd_rate_pyccel.py:30-83``)."""
    h = halo
    nx = rho.shape[0] - 2 * h
    ny = rho.shape[1] - 2 * h
    i = slice(h, h + nx)
    j = slice(h, h + ny)
    i1 = slice(h + 1, h + nx + 1)
    j1 = slice(h + 1, h + ny + 1)
    with np.errstate(all="ignore"):
        r = rho[i, j]
        u = mx[i, j] / r
        v = my[i, j] / r
        w = mz[i, j] / r
        bx = 0.5 * ((az[i, j1] - az[i, j]) / dy + b0x + ((az[i1, j1] - az[i1, j]) / dy + b0x))
        by = 0.5 * (-(az[i1, j] - az[i, j]) / dx + b0y + (-(az[i1, j1] - az[i, j1]) / dx + b0y))
        bzc = bz[i, j]
        p = (gamma - 1.0) * (
            energy[i, j] - 0.5 * r * (u * u + v * v + w * w) - 0.5 * (bx * bx + by * by + bzc * bzc)
        )
        rates = (np.abs(u) + fast_speed_rate(r, p, bx, by, bzc, gamma)) / dx + (
            np.abs(v) + fast_speed_rate(r, p, by, bx, bzc, gamma)
        ) / dy
        largest = float(np.max(np.where(rates > 0.0, rates, 0.0)))
        smallest = float(np.nanmin(r))
    return largest, smallest


def mc(a, b):
    """Monotonized-central limiter of two differences (``mhd_conductive_pyccel.py:25-38``)."""
    with np.errstate(all="ignore"):
        central = 0.5 * (a + b)
        limit = np.where(np.abs(a) < np.abs(b), 2.0 * a, 2.0 * b)
        size = np.minimum(np.abs(central), np.abs(limit))
        slope = np.where(central > 0.0, size, np.where(central < 0.0, -size, 0.0))
        return np.where(a * b > 0.0, slope, 0.0)


def resistive(prim, bxf, byf, dx, dy, eta, fx, fy, ez) -> None:
    """Adds the resistive terms to ``fx``, ``fy`` and ``ez`` in place
    (``kernels/mhd_resistive/mhd_resistive_pyccel.py:21-67``).

    ``J_z`` at the haloed block's nodes from the face fields; on the x-faces the B_z
    flux gains ``eta J_y`` and the energy flux ``eta (J_y B_z - J_z B_y)``, on the
    y-faces ``-eta J_x`` and ``eta (J_z B_x - J_x B_z)`` (face means of the two cells
    and of the face's two nodes); ``ez`` gains ``eta J_z``.
    """
    h = 2
    X, Y = prim.shape[1:]
    nx, ny = X - 2 * h, Y - 2 * h
    jz = np.zeros((X, Y))
    jz[1:, 1:] = (byf[1:X, 1:Y] - byf[0 : X - 1, 1:Y]) / dx - (bxf[1:X, 1:Y] - bxf[1:X, 0 : Y - 1]) / dy
    i = h + np.arange(nx + 1)[:, None]
    j = h + np.arange(ny)[None, :]
    jy = -(prim[7, i, j] - prim[7, i - 1, j]) / dx
    bz_x = 0.5 * (prim[7, i, j] + prim[7, i - 1, j])
    by_x = 0.5 * (prim[6, i, j] + prim[6, i - 1, j])
    jz_x = 0.5 * (jz[i, j] + jz[i, j + 1])
    f = i - h
    fx[6, f, j] = fx[6, f, j] + eta * jy
    fx[4, f, j] = fx[4, f, j] + eta * (jy * bz_x - jz_x * by_x)
    i = h + np.arange(nx)[:, None]
    j = h + np.arange(ny + 1)[None, :]
    jx = (prim[7, i, j] - prim[7, i, j - 1]) / dy
    bz_y = 0.5 * (prim[7, i, j] + prim[7, i, j - 1])
    bx_y = 0.5 * (prim[5, i, j] + prim[5, i, j - 1])
    jz_y = 0.5 * (jz[i, j] + jz[i + 1, j])
    f = j - h
    fy[6, i, f] = fy[6, i, f] + -eta * jx
    fy[4, i, f] = fy[4, i, f] + eta * (jz_y * bx_y - jx * bz_y)
    ez[...] = ez + eta * jz[h : h + nx + 1, h : h + ny + 1]


#: The smallest normal double (``mhd_conductive_pyccel.py:55``).
TINY = 2.2250738585072014e-308


def conductive(prim, bxf, byf, dx, dy, parallel, perpendicular, fx, fy) -> None:
    """Subtracts the anisotropic heat flux from ``fx[4]`` and ``fy[4]`` in place
    (``kernels/mhd_conductive/mhd_conductive_pyccel.py:41-100``).

    ``q_n = parallel b_n (b_n dT/dn + b_t dT/dt) + perpendicular dT/dn`` with
    ``T = p / rho``, ``b`` the unit field at the face (its length at least ``TINY``),
    ``dT/dn`` the difference across the face and ``dT/dt`` the MC-limited mean of the
    MC-limited gradients along the face of the two cells (zero in the block's first
    and last line).
    """
    X, Y = prim.shape[1:]
    with np.errstate(all="ignore"):
        t = prim[4] / prim[0]
        across = np.zeros((X, Y))
        across[:, 1 : Y - 1] = mc((t[:, 2:] - t[:, 1:-1]) / dy, (t[:, 1:-1] - t[:, :-2]) / dy)
        k = np.arange(X - 3)[:, None]
        j = np.arange(Y)[None, :]
        bn = bxf[k + 2, j]
        bt = 0.5 * (prim[6, k + 1, j] + prim[6, k + 2, j])
        bzf = 0.5 * (prim[7, k + 1, j] + prim[7, k + 2, j])
        magnitude = np.maximum(np.sqrt(bn * bn + bt * bt + bzf * bzf), TINY)
        b_n = bn / magnitude
        b_t = bt / magnitude
        normal = (t[k + 2, j] - t[k + 1, j]) / dx
        transverse = mc(across[k + 1, j], across[k + 2, j])
        fx[4] = fx[4] - (parallel * b_n * (b_n * normal + b_t * transverse) + perpendicular * normal)
        across = np.zeros((X, Y))
        across[1 : X - 1, :] = mc((t[2:, :] - t[1:-1, :]) / dx, (t[1:-1, :] - t[:-2, :]) / dx)
        i = np.arange(X)[:, None]
        k = np.arange(Y - 3)[None, :]
        bn = byf[i, k + 2]
        bt = 0.5 * (prim[5, i, k + 1] + prim[5, i, k + 2])
        bzf = 0.5 * (prim[7, i, k + 1] + prim[7, i, k + 2])
        magnitude = np.maximum(np.sqrt(bn * bn + bt * bt + bzf * bzf), TINY)
        b_n = bn / magnitude
        b_t = bt / magnitude
        normal = (t[i, k + 2] - t[i, k + 1]) / dy
        transverse = mc(across[i, k + 1], across[i, k + 2])
        fy[4] = fy[4] - (parallel * b_n * (b_n * normal + b_t * transverse) + perpendicular * normal)
