"""FP32 emulation of the kernel sequence (test utility, CPU only).

The same decomposition as ``plan_emulator`` (pruned column passes over
spectral boxes, alias-free working grids, the kernels' floors, masks and
FP64 tables), with a precision switch per GL task: an FP32 task casts each
inverse-transform input to complex64 and runs its transforms and pointwise
products in single precision (scipy.fft keeps complex64 in complex64), while
every spectrum it hands on and every GL accumulation stays FP64 -- the
arithmetic a FP32-transform / FP64-accumulate GPU path would do.  It shows
why the GL graphs cannot run their transforms in FP32 (``api._precision``).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.fft as sfft

import plan_emulator as E

EPS32 = float(np.finfo(np.float32).eps)


def to_grid(spec_box, box, gy, gx, scale, f32):
    dt = np.complex64 if f32 else np.complex128
    g = np.zeros((gy, gx), dtype=dt)
    rows = np.arange(box.my0, box.my0 + box.wy) % gy
    cols = np.arange(box.mx0, box.mx0 + box.wx) % gx
    g[np.ix_(rows, cols)] = (spec_box * scale).astype(dt)
    if f32:
        return sfft.ifft2(g) * np.float32(gy * gx)
    return np.fft.ifft2(g) * (gy * gx)


def from_grid(field, box, mask, scale, f32):
    gy, gx = field.shape
    if f32:
        f = sfft.fft2(field.astype(np.complex64)).astype(np.complex128) * scale
    else:
        f = np.fft.fft2(field) * scale
    rows = np.arange(box.my0, box.my0 + box.wy) % gy
    cols = np.arange(box.mx0, box.mx0 + box.wx) % gx
    s = np.where(mask.astype(bool), f[np.ix_(rows, cols)], 0.0)
    return s, float(np.max(np.abs(s[mask.astype(bool)]), initial=0.0))


def floor(x, level, top):
    return np.where(np.abs(x) < level * top, 0.0, x)


def order3(p, s_box, inner_f32, outer_f32, level32, level64=None):
    """eta33+, psi33+ box spectra; inner_f32[o][i] / outer_f32[o] select FP32 tasks, whose spectra
    are floored at level32 (FP64 tasks at level64, default the plan's FP64 floor)."""
    o3 = p.o3
    level64 = p.floor_level if level64 is None else level64
    nn = p.nx * p.ny
    g2x, g2y = p.g2
    g3x, g3y = p.g3
    up2, up3 = nn / (g2x * g2y), nn / (g3x * g3y)
    I, O = o3["inner_rank"], o3["outer_rank"]
    inner_codes = ["one", "nu", "nu2", "nu3", "qx_nu", "qy_nu", "qx", "qy", "q2_nu", "q2", "nu_qx", "nu_qy"]
    first_codes = ["one", "i_qx", "i_qy", "qx_nu", "qy_nu", "mi_nu", "nu_qx", "nu_qy", "mi_q2_nu", "mi_q2nu", "ma",
                   "mq2"]
    pair_codes = [(0, "one"), (0, "i_qx"), (0, "i_qy"), (2, "i_qx"), (2, "i_qy"), (2, "a"), (2, "q2"), (3, "a"),
                  (1, "one")]
    mask2, mask3 = p.mask2.astype(bool), p.mask3.astype(bool)
    eta = np.zeros((p.box3.wy, p.box3.wx), dtype=complex)
    psi = np.zeros_like(eta)
    for o in range(O):
        pair = np.zeros((4,) + mask2.shape, dtype=complex)
        for i in range(I):
            f32 = inner_f32[o][i]
            lev = level32 if f32 else level64
            d = s_box * o3["damp_inner"][o * I + i]
            v, vn, vn2, vn3, hx, hy, jx, jy, rad, q2f, njx, njy = (
                to_grid(d * E.mult(c, p.t1), p.box1, g2y, g2x, 1.0 / nn, f32) for c in inner_codes)
            srcs = [2 * v * vn2 + vn * vn - hx * hx - hy * hy, 2 * v * rad + 2 * (hx * jx + hy * jy),
                    2 * v * vn3 + 4 * vn * vn2 - 2 * hx * jx - 2 * hy * jy,
                    2 * (vn * rad + v * q2f) + 2 * (jx * jx + hx * njx + jy * jy + hy * njy)]
            sp = []
            for s in srcs:
                x, top = from_grid(s, p.box2, p.mask2, up2, f32)
                sp.append(floor(x, lev, top))
            sd, sk, sdt, skt = sp
            A = p.t2.a
            ch, sh, c = o3["inner_cosh"][i], o3["inner_sinhr"][i], o3["inner_coef"][i]
            pair[0] += c * (-A * sh * sd + ch * sk)
            pair[1] -= 1j * c * (-A * sh * sdt + ch * skt)
            pair[2] += 1j * c * (ch * sd - sh * sk)
            pair[3] += c * (ch * sdt - sh * skt)
        pair = np.where(mask2, pair, 0.0)
        pair[1] = floor(pair[1], level64, float(np.max(np.abs(pair[1][mask2]), initial=0.0)))
        f32 = outer_f32[o]
        lev = level32 if f32 else level64
        d = s_box * o3["damp_outer"][o]
        f = [to_grid(d * E.mult(c, p.t1), p.box1, g3y, g3x, 1.0 / nn, f32) for c in first_codes]
        q = [to_grid(pair[s] * E.mult(c, p.t2), p.box2, g3y, g3x, 1.0 / nn, f32) for s, c in pair_codes]
        fe, fex, fey, fpx, fpy, fpz, fpxz, fpyz, fpzz, fpzzz, fptz, fptzz = f
        pe, pex, pey, ppx, ppy, ppz, ppzz, pptz, pet = q
        k_pair = fpx * pex + fpy * pey + ppx * fex + ppy * fey - fe * ppzz - pe * fpzz
        k_dir = fe * (fpxz * fex + fpyz * fey) - 0.5 * fe * fe * fpzzz
        d_pair = fe * pptz + pe * fptz + fpx * ppx + fpy * ppy + fpz * ppz
        d_dir = 0.5 * fe * fe * fptzz + fe * (fpx * fpxz + fpy * fpyz + fpz * fpzz)
        cont = 0.5 * fe * ppz + 0.5 * pe * fpz + 0.125 * fe * fe * fpzz
        cont_t = (0.5 * (fpz * ppz + fe * pptz) + 0.5 * (pet * fpz + pe * fptz) + 0.25 * fe * fpz * fpzz
                  + 0.125 * fe * fe * fptzz)
        sp = []
        for s in (k_dir / 4 + k_pair / 2, d_dir / 4 + d_pair / 2, cont, cont_t):
            x, top = from_grid(s, p.box3, p.mask3, up3, f32)
            sp.append(floor(x, lev, top))
        fk, fd, c0, ct = sp
        A = p.t3.a
        kp = floor(fk + A * c0, lev, float(np.max(np.abs((fk + A * c0)[mask3]), initial=0.0)))
        dp = floor(fd - ct, lev, float(np.max(np.abs((fd - ct)[mask3]), initial=0.0)))
        ch, sh, c = o3["outer_cosh"][o], o3["outer_sinhr"][o], o3["outer_coef"][o]
        eta += c * (A * sh * fd - 1j * ch * fk)
        psi += c * (-1j * ch * dp - sh * kp)
    return np.where(mask3, eta, 0.0), np.where(mask3, psi, 0.0)


def eta22(p, s_box, f32, level):
    """eta22+ box spectrum with every GL node's transforms in FP32 (f32) or FP64, floored at `level`."""
    o2 = p.o2
    nn = p.nx * p.ny
    g2x, g2y = p.g2
    up2 = nn / (g2x * g2y)
    codes = ["one", "nu", "nu2", "qx_nu", "qy_nu", "q2_nu", "qx", "qy"]
    A = p.t2.a
    out = np.zeros(p.mask2.shape, dtype=complex)
    for j in range(o2["eta_rank"]):
        d = s_box * o2["eta_damp"][j]
        v, vn, vn2, hx, hy, rad, jx, jy = (to_grid(d * E.mult(c, p.t1), p.box1, g2y, g2x, 1.0 / nn, f32)
                                           for c in codes)
        sd, top_d = from_grid(2 * v * vn2 + vn * vn - hx * hx - hy * hy, p.box2, p.mask2, up2, f32)
        sk, top_k = from_grid(2 * v * rad + 2 * (hx * jx + hy * jy), p.box2, p.mask2, up2, f32)
        sd, sk = floor(sd, level, top_d), floor(sk, level, top_k)
        out += o2["eta_coef"][j] * (-A * o2["eta_sinhr"][j] * sd + o2["eta_cosh"][j] * sk)
    return out


def physical(p, spec_box, box) -> np.ndarray:
    return E.synth(spec_box, box, p.ny, p.nx, 1.0).real


def floor32(p) -> float:
    """The reference floor rule 16 eps log2(N^2) (_spectral.py:52-64) with the FP32 epsilon."""
    return 16.0 * EPS32 * max(1.0, math.log2(p.nx * p.ny))
