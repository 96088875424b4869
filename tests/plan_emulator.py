"""NumPy mirror of the HIP kernel sequence (test utility, CPU only).

Executes the same decomposition as the kernels -- pruned column passes over
spectral boxes, row passes on the alias-free working grids, the same floors,
masks and FP64 tables from ``plan.HostPlan`` -- so that the plan and the
decomposition can be checked against the oracle without a GPU.  It is test
infrastructure; the product package never imports it.
"""

from __future__ import annotations

import math

import numpy as np

from green_laplace_wave_reconstruction_amd.plan import Box, HostPlan


def mult(code: str, t) -> np.ndarray:
    table = {
        "one": lambda: np.ones_like(t.qx), "i_qx": lambda: 1j * t.qx, "i_qy": lambda: 1j * t.qy,
        "qx_nu": lambda: t.qx * t.inu, "qy_nu": lambda: t.qy * t.inu, "mi_nu": lambda: -1j * t.nu,
        "nu_qx": lambda: t.nu * t.qx, "nu_qy": lambda: t.nu * t.qy, "mi_q2_nu": lambda: -1j * t.q2 * t.inu,
        "mi_q2nu": lambda: -1j * t.q2 * t.nu, "ma": lambda: -t.a, "mq2": lambda: -t.q2, "nu": lambda: t.nu,
        "nu2": lambda: t.nu * t.nu, "nu3": lambda: t.nu * t.nu * t.nu, "qx": lambda: t.qx, "qy": lambda: t.qy,
        "q2_nu": lambda: t.q2 * t.inu, "q2": lambda: t.q2, "a": lambda: t.a,
        "mi_inu": lambda: np.where(t.nu > 0, -1j / np.where(t.nu > 0, t.nu, 1.0), 0.0),
    }
    return table[code]()


def to_grid(spec_box: np.ndarray, box: Box, gy: int, gx: int, scale: float) -> np.ndarray:
    """Inverse transform of a box spectrum onto a gy x gx grid (unnormalised * scale)."""
    g = np.zeros((gy, gx), dtype=np.complex128)
    rows = np.arange(box.my0, box.my0 + box.wy) % gy
    cols = np.arange(box.mx0, box.mx0 + box.wx) % gx
    g[np.ix_(rows, cols)] = spec_box * scale
    return np.fft.ifft2(g) * (gy * gx)


def from_grid(field: np.ndarray, box: Box, mask: np.ndarray, scale: float):
    """Forward transform to a box, masked; returns (spectrum, support max |x|)."""
    gy, gx = field.shape
    f = np.fft.fft2(field) * scale
    rows = np.arange(box.my0, box.my0 + box.wy) % gy
    cols = np.arange(box.mx0, box.mx0 + box.wx) % gx
    s = f[np.ix_(rows, cols)]
    s = np.where(mask.astype(bool), s, 0.0)
    return s, float(np.max(np.abs(s[mask.astype(bool)]), initial=0.0))


def floor(x: np.ndarray, level: float, top: float) -> np.ndarray:
    return np.where(np.abs(x) < level * top, 0.0, x)


def order3(plan: HostPlan, s_box: np.ndarray):
    """eta33+, psi33+ spectra on box3 (dimensionless), same steps as the kernels."""
    p, o3 = plan, plan.o3
    nn = p.nx * p.ny
    g2x, g2y = p.g2
    g3x, g3y = p.g3
    up2 = nn / (g2x * g2y)
    up3 = nn / (g3x * g3y)
    I, O = o3["inner_rank"], o3["outer_rank"]
    inner_codes = ["one", "nu", "nu2", "nu3", "qx_nu", "qy_nu", "qx", "qy", "q2_nu", "q2", "nu_qx", "nu_qy"]
    first_codes = ["one", "i_qx", "i_qy", "qx_nu", "qy_nu", "mi_nu", "nu_qx", "nu_qy", "mi_q2_nu", "mi_q2nu", "ma", "mq2"]
    pair_codes = [(0, "one"), (0, "i_qx"), (0, "i_qy"), (2, "i_qx"), (2, "i_qy"), (2, "a"), (2, "q2"), (3, "a"), (1, "one")]
    mask2 = p.mask2.astype(bool)
    mask3 = p.mask3.astype(bool)
    eta = np.zeros(p.box3.wy * p.box3.wx, dtype=complex).reshape(p.box3.wy, p.box3.wx)
    psi = np.zeros_like(eta)
    for o in range(O):
        pair = np.zeros((4,) + mask2.shape, dtype=complex)
        for i in range(I):
            d = s_box * o3["damp_inner"][o * I + i]
            v, vn, vn2, vn3, hx, hy, jx, jy, rad, q2f, njx, njy = (
                to_grid(d * mult(c, p.t1), p.box1, g2y, g2x, 1.0 / nn) for c in inner_codes)
            srcs = [2 * v * vn2 + vn * vn - hx * hx - hy * hy, 2 * v * rad + 2 * (hx * jx + hy * jy),
                    2 * v * vn3 + 4 * vn * vn2 - 2 * hx * jx - 2 * hy * jy,
                    2 * (vn * rad + v * q2f) + 2 * (jx * jx + hx * njx + jy * jy + hy * njy)]
            sp = []
            for s in srcs:
                x, top = from_grid(s, p.box2, p.mask2, up2)
                sp.append(floor(x, p.floor_level, top))
            sd, sk, sdt, skt = sp
            A = p.t2.a
            ch, sh, c = o3["inner_cosh"][i], o3["inner_sinhr"][i], o3["inner_coef"][i]
            pair[0] += c * (-A * sh * sd + ch * sk)
            pair[1] -= 1j * c * (-A * sh * sdt + ch * skt)
            pair[2] += 1j * c * (ch * sd - sh * sk)
            pair[3] += c * (ch * sdt - sh * skt)
        pair = np.where(mask2, pair, 0.0)
        pair[1] = floor(pair[1], p.floor_level, float(np.max(np.abs(pair[1][mask2]), initial=0.0)))
        d = s_box * o3["damp_outer"][o]
        f = [to_grid(d * mult(c, p.t1), p.box1, g3y, g3x, 1.0 / nn) for c in first_codes]
        q = [to_grid(pair[s] * mult(c, p.t2), p.box2, g3y, g3x, 1.0 / nn) for s, c in pair_codes]
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
            x, top = from_grid(s, p.box3, p.mask3, up3)
            sp.append(floor(x, p.floor_level, top))
        fk, fd, c0, ct = sp
        A = p.t3.a
        kp = fk + A * c0
        dp = fd - ct
        kp = floor(kp, p.floor_level, float(np.max(np.abs(kp[mask3]), initial=0.0)))
        dp = floor(dp, p.floor_level, float(np.max(np.abs(dp[mask3]), initial=0.0)))
        ch, sh, c = o3["outer_cosh"][o], o3["outer_sinhr"][o], o3["outer_coef"][o]
        eta += c * (A * sh * fd - 1j * ch * fk)
        psi += c * (-1j * ch * dp - sh * kp)
    return np.where(mask3, eta, 0.0), np.where(mask3, psi, 0.0)


def order2(plan: HostPlan, s_box: np.ndarray):
    """eta22+, psi22+ spectra on box2 (dimensionless)."""
    p, o2 = plan, plan.o2
    nn = p.nx * p.ny
    g2x, g2y = p.g2
    up2 = nn / (g2x * g2y)
    codes = ["one", "nu", "nu2", "qx_nu", "qy_nu", "q2_nu", "qx", "qy"]
    A = p.t2.a
    cand = np.zeros(p.mask2.shape, dtype=complex)
    for j in range(o2["eta_rank"]):
        d = s_box * o2["eta_damp"][j]
        v, vn, vn2, hx, hy, rad, jx, jy = (to_grid(d * mult(c, p.t1), p.box1, g2y, g2x, 1.0 / nn) for c in codes)
        sd, top_d = from_grid(2 * v * vn2 + vn * vn - hx * hx - hy * hy, p.box2, p.mask2, up2)
        sk, top_k = from_grid(2 * v * rad + 2 * (hx * jx + hy * jy), p.box2, p.mask2, up2)
        sd = floor(sd, p.floor_level, top_d)
        sk = floor(sk, p.floor_level, top_k)
        cand += o2["eta_coef"][j] * (-A * o2["eta_sinhr"][j] * sd + o2["eta_cosh"][j] * sk)
    res = np.zeros_like(cand)
    for b in range(2):
        for n in range(2):
            fields = []
            for ch in range(2):
                d = s_box * o2["psi_damp"][b, n, ch]
                fields.append([to_grid(d * mult(c, p.t1), p.box1, g2y, g2x, 1.0 / nn) for c in codes])
            resp = np.zeros_like(cand)
            for i in range(2):
                for k in range(2):
                    L, R = fields[i], fields[k]
                    sd, _ = from_grid(2 * L[0] * R[2] + L[1] * R[1] - L[3] * R[3] - L[4] * R[4], p.box2, p.mask2, up2)
                    sk, _ = from_grid(2 * L[0] * R[5] + 2 * (L[3] * R[6] + L[4] * R[7]), p.box2, p.mask2, up2)
                    resp += o2["psi_factor"][b, n, 2 * i + k] * (sd + o2["psi_sign"][b] * sk * p.t2.inu)
            res += 1j * o2["psi_coef"][b, n] * resp
    e = to_grid(s_box, p.box1, g2y, g2x, 1.0 / nn)
    pz = to_grid(s_box * mult("mi_nu", p.t1), p.box1, g2y, g2x, 1.0 / nn)
    cont, _ = from_grid(0.5 * e * pz, p.box2, p.mask2, up2)
    psi = np.where(p.mask2.astype(bool), res + cont, 0.0)
    return cand, psi


def synth(spec_box: np.ndarray, box: Box, ny: int, nx: int, scale: float) -> np.ndarray:
    return to_grid(spec_box, box, ny, nx, scale / (nx * ny))


def phase(plan: HostPlan, s0_box: np.ndarray, t: float) -> np.ndarray:
    return s0_box * np.exp(-1j * plan.omega * t)


def fields(plan: HostPlan, s0_box: np.ndarray, t: float) -> dict[str, np.ndarray]:
    """All components at one instant, scaled like the reference's *_plus outputs."""
    s = phase(plan, s0_box, t)
    h = plan.depth
    pot = h * math.sqrt(plan.gravity * h)
    out = {"eta11_plus": synth(s, plan.box1, plan.ny, plan.nx, h),
           "psi11_plus": synth(s * mult("mi_inu", plan.t1), plan.box1, plan.ny, plan.nx, pot)}
    if plan.order >= 2:
        e22, p22 = order2(plan, s)
        out["eta22_plus"] = synth(e22, plan.box2, plan.ny, plan.nx, h)
        out["psi22_plus"] = synth(p22, plan.box2, plan.ny, plan.nx, pot)
    if plan.order >= 3:
        e33, p33 = order3(plan, s)
        out["eta33_plus"] = synth(e33, plan.box3, plan.ny, plan.nx, h)
        out["psi33_plus"] = synth(p33, plan.box3, plan.ny, plan.nx, pot)
    return out
