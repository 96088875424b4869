"""FP64 NumPy restatement of the reference directional spatial GL graph.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Every function cites
the reference file and lines it restates.  Arrays are ``[Ny, Nx]`` (row = y),
complex128, with the ``numpy.fft`` conventions (``ifft2`` scaled by 1/N^2).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np

TINY = np.finfo(float).tiny
EPS = np.finfo(float).eps


# --------------------------------------------------------------------------
# quadrature and grids
# --------------------------------------------------------------------------
def gauss_laguerre_rule(rank: int) -> tuple[np.ndarray, np.ndarray]:
    """Golub--Welsch Gauss--Laguerre rule with the moment gate.

    Restates python/src/green_laplace/quadrature.py:11-31 (MATLAB twin
    src/internal/gl_no_stokes_eta33.m:638-646): Jacobi matrix with diagonal
    2i-1 and off-diagonal i, weights = first eigenvector component squared.
    """
    if not isinstance(rank, (int, np.integer)) or rank < 1:
        raise ValueError("rank must be a positive integer")
    rank = int(rank)
    i = np.arange(1, rank + 1, dtype=float)
    jac = np.diag(2.0 * i - 1.0)
    if rank > 1:
        off = np.arange(1, rank, dtype=float)
        jac += np.diag(off, 1) + np.diag(off, -1)
    x, vec = np.linalg.eigh(jac)
    w = vec[0, :] ** 2
    for d in range(2 * rank):
        exact = math.factorial(d)
        if abs(float(np.sum(w * x**d)) - exact) / max(1, exact) > 3.0e-11:
            raise ArithmeticError(f"Gauss--Laguerre moment gate failed at rank={rank}, d={d}")
    return x, w


def psi22_two_node_rule() -> tuple[np.ndarray, np.ndarray]:
    """Closed-form psi22 nodes 2-/+sqrt2 and weights (spatial.py:269-275)."""
    r2 = math.sqrt(2.0)
    x = np.array([2.0 - r2, 2.0 + r2])
    w = np.array([x[0] / (4.0 * (r2 - 1.0) ** 2), x[1] / (4.0 * (r2 + 1.0) ** 2)])
    return x, w


def wavenumber_grid(lx: float, ly: float, nx: int, ny: int):
    """FFT wavenumber meshgrid (python/src/green_laplace/_spectral.py:9-16)."""
    if lx <= 0 or ly <= 0 or nx < 2 or ny < 2 or nx % 2 or ny % 2:
        raise ValueError("positive domain lengths and positive even grid sizes are required")
    kx1 = 2.0 * np.pi * np.fft.fftfreq(nx, d=lx / nx)
    ky1 = 2.0 * np.pi * np.fft.fftfreq(ny, d=ly / ny)
    return np.meshgrid(kx1, ky1)


def analytic_spectrum(eta1: np.ndarray, depth: float, kx: np.ndarray, tol: float = 1e-11):
    """Strict-forward analytic spectrum of a real snapshot (_spectral.py:19-49)."""
    if eta1.ndim != 2 or eta1.shape != kx.shape or not np.all(np.isfinite(eta1)):
        raise ValueError("eta1 must be a finite two-dimensional grid")
    if depth <= 0:
        raise ValueError("depth must be positive")
    raw = np.fft.fft2(eta1 / depth)
    scale = max(float(np.max(np.abs(raw))), TINY)
    on_axis = np.abs(kx) <= 32.0 * EPS * max(1.0, float(np.max(np.abs(kx))))
    if np.max(np.abs(raw[on_axis]), initial=0.0) > tol * scale:
        raise ValueError("eta1 contains a non-negligible kx=0 component")
    fwd = kx > 0
    spec = np.zeros_like(raw, dtype=np.complex128)
    spec[fwd] = 2.0 * raw[fwd]
    back = np.real(depth * np.fft.ifft2(spec))
    err = float(np.linalg.norm(back - eta1) / max(np.linalg.norm(eta1), TINY))
    if err > tol:
        raise ValueError("eta1 is not consistent with the declared strict-forward x propagation")
    return spec, err


def floor_roundoff(spec: np.ndarray, support: np.ndarray) -> np.ndarray:
    """Zero support bins below 16*eps*log2(N^2)*max_support|x| (_spectral.py:52-64)."""
    out = spec.copy()
    active = np.abs(out[support])
    if active.size == 0:
        return out
    top = float(np.max(active))
    if top == 0.0:
        return out
    level = 16.0 * EPS * max(1.0, math.log2(out.size)) * top
    out[support & (np.abs(out) < level)] = 0.0
    return out


def _mode_grids(shape):
    ny, nx = shape
    return np.meshgrid(np.fft.fftfreq(nx) * nx, np.fft.fftfreq(ny) * ny)


def alias_guard(support: np.ndarray, order: int) -> None:
    """order*|m| < N/2 on the parent support (_spectral.py:67-77; _order3.py:290-296)."""
    ny, nx = support.shape
    mx, my = _mode_grids(support.shape)
    if not np.any(support):
        raise ValueError("eta1 has empty analytic support")
    if np.max(order * np.abs(mx[support])) >= nx / 2.0 or np.max(order * np.abs(my[support])) >= ny / 2.0:
        word = {2: "quadratic", 3: "cubic"}.get(order, f"order-{order}")
        raise ValueError(f"{word} output support reaches a Nyquist boundary")


def parent_support(spec: np.ndarray) -> np.ndarray:
    """|S| > 1e-13*max(1, max|S|) (_order3.py:33-34; spatial.py:163-164)."""
    return np.abs(spec) > 1.0e-13 * max(1.0, float(np.max(np.abs(spec))))


def sum_support(support: np.ndarray, n: int) -> np.ndarray:
    """n-fold Minkowski sum via real(ifft2(fft2(mask)^n)) > 0.5 (_order3.py:38-40)."""
    t = np.fft.fft2(support.astype(float))
    return np.real(np.fft.ifft2(t**n)) > 0.5


@dataclass
class Geometry:
    """Dimensionless wavenumber grids q=kh, nu=sqrt(q tanh q), A=q tanh q."""

    qx: np.ndarray
    qy: np.ndarray
    q: np.ndarray = field(init=False)
    nu: np.ndarray = field(init=False)
    a: np.ndarray = field(init=False)
    sa: np.ndarray = field(init=False)
    snu: np.ndarray = field(init=False)

    def __post_init__(self):
        self.q = np.hypot(self.qx, self.qy)
        self.a = self.q * np.tanh(self.q)
        self.nu = np.sqrt(self.a)
        self.sa = np.sqrt(self.a)
        self.snu = np.maximum(self.nu, TINY)


def sinh_ratio(sa: np.ndarray, tau: float) -> np.ndarray:
    """sinh(sqrt(A) tau)/sqrt(A) with the tau limit at A=0 (_order3.py:124-126)."""
    out = np.full_like(sa, tau)
    nz = sa > 0
    out[nz] = np.sinh(sa[nz] * tau) / sa[nz]
    return out


def peak_q_of(spec: np.ndarray, support: np.ndarray, q: np.ndarray) -> float:
    """q at the first argmax of |S|^2 over the support in row-major order (_order3.py:45-47)."""
    idx = np.flatnonzero(support)
    return float(q.flat[idx[int(np.argmax(np.abs(spec.flat[idx]) ** 2))]])


# --------------------------------------------------------------------------
# order 2
# --------------------------------------------------------------------------
def eta22(spec: np.ndarray, qx: np.ndarray, qy: np.ndarray, rank: int, peak_q: float | None = None):
    """Prescribed-rank GL eta22 (python/src/green_laplace/spatial.py:156-231;
    src/internal/green_laplace_eta22.m:1-79)."""
    g = Geometry(qx, qy)
    sup = parent_support(spec)
    if np.any(qx[sup] <= 0) or np.any(g.q[sup] <= 0):
        raise ValueError("strict-forward nonzero analytic eta1 support is required")
    alias_guard(sup, 2)
    out_sup = sum_support(sup, 2)
    if peak_q is None:
        peak_q = peak_q_of(spec, sup, g.q)
    disc = 4.0 * peak_q * math.tanh(peak_q) - 2.0 * peak_q * math.tanh(2.0 * peak_q)
    lam = math.sqrt(disc) if disc > 0 else float("nan")
    if not np.isfinite(lam) or lam <= 0:
        raise ValueError("invalid determinant-root Green--Laplace scale")
    x, w = gauss_laguerre_rule(rank)
    c = 0.5 * (float(np.min(g.nu[sup])) + float(np.max(g.sa[out_sup])) / 2.0)
    acc = np.zeros_like(spec)
    ifft2, fft2 = np.fft.ifft2, np.fft.fft2
    for xn, wn in zip(x, w):
        tau = xn / lam
        d = spec * np.exp(-tau * (g.nu - c))
        v = ifft2(d)
        vn = ifft2(d * g.nu)
        vn2 = ifft2(d * g.nu**2)
        hx = ifft2(d * qx / g.snu)
        hy = ifft2(d * qy / g.snu)
        rad = ifft2(d * g.q**2 / g.snu)
        jx = ifft2(d * qx)
        jy = ifft2(d * qy)
        sd = fft2(2.0 * v * vn2 + vn**2 - hx**2 - hy**2)
        sk = fft2(2.0 * v * rad + 2.0 * (hx * jx + hy * jy))
        sd[~out_sup] = 0.0
        sk[~out_sup] = 0.0
        sd = floor_roundoff(sd, out_sup)
        sk = floor_roundoff(sk, out_sup)
        coef = wn * math.exp(xn) / lam / 4.0 * math.exp(-2.0 * c * tau)
        acc += coef * (-g.a * sinh_ratio(g.sa, tau) * sd + np.cosh(g.sa * tau) * sk)
    audit = {
        "implementation": "prescribed-rank-green-laplace-eta22",
        "quadrature_rank": rank,
        "quadrature_exact_moment_degree": 2 * rank - 1,
        "peak_depth_wavenumber": peak_q,
        "fft_ifft_count": 10 * rank + 5,
        "pointwise_product_count": 7 * rank + 1,
        "interaction_enumeration": False,
    }
    return np.fft.ifft2(acc), audit


def radial_channels(support: np.ndarray, nu: np.ndarray, count: int = 2):
    """Radial balance channels on nu (spatial.py:322-346)."""
    active = nu[support]
    edges = np.linspace(float(np.min(active)), float(np.max(active)), count + 1)
    masks, shifts = [], []
    for ch in range(count):
        upper = (nu < edges[ch + 1]) if ch < count - 1 else (nu <= edges[ch + 1])
        m = support & (nu >= edges[ch]) & upper
        if np.any(m):
            vals = nu[m]
            masks.append(m)
            shifts.append(0.5 * (float(np.min(vals)) + float(np.max(vals))))
    ts = [np.fft.fft2(m.astype(float)) for m in masks]
    pairs = [[np.real(np.fft.ifft2(ts[i] * ts[j])) > 0.5 for j in range(len(masks))] for i in range(len(masks))]
    return masks, shifts, pairs


def psi22(spec: np.ndarray, qx: np.ndarray, qy: np.ndarray, allow_below: bool = False,
          peak_q: float | None = None):
    """Dual-branch radial2 GL psi22 (spatial.py:234-319, 349-414;
    src/internal/finite_depth_directional_dual_branch_green_laplace_psi22.m)."""
    g = Geometry(qx, qy)
    sup = parent_support(spec)
    if not np.any(sup) or np.any(qx[sup] <= 0) or np.any(g.q[sup] <= 0):
        raise ValueError("strict-forward nonzero analytic eta1 support is required")
    below = bool(np.any(g.q[sup] < 0.3))
    if below and not allow_below:
        raise ValueError("psi22 requires parent kh >= 0.3")
    below_mass = float(np.sum(np.abs(spec[g.q < 0.3]) ** 2) / max(np.sum(np.abs(spec) ** 2), TINY))
    alias_guard(sup, 2)
    out_sup = sum_support(sup, 2)
    masks, shifts, pairs = radial_channels(sup, g.nu, 2)
    if peak_q is None:
        peak_q = peak_q_of(spec, sup, g.q)
    pn = math.sqrt(peak_q * math.tanh(peak_q))
    pon = math.sqrt(2.0 * peak_q * math.tanh(2.0 * peak_q))
    scales = (2.0 * pn - pon, 2.0 * pn + pon)
    xs, ws = psi22_two_node_rule()
    ifft2, fft2 = np.fft.ifft2, np.fft.fft2
    res = np.zeros_like(spec)
    for branch, lam, sgn in (("slow", scales[0], -1.0), ("fast", scales[1], 1.0)):
        if lam <= 0 or not np.isfinite(lam):
            raise ValueError("invalid dual-branch Green--Laplace scale")
        for xn, wn in zip(xs, ws):
            tau = xn / lam
            coef = wn * math.exp(xn) / lam / 8.0
            banks = []
            for m, c in zip(masks, shifts):
                d = spec * m * np.exp(-tau * (g.nu - c))
                banks.append(dict(
                    v=ifft2(d), vn=ifft2(d * g.nu), vn2=ifft2(d * g.nu**2),
                    hx=ifft2(d * qx / g.snu), hy=ifft2(d * qy / g.snu),
                    rad=ifft2(d * g.q**2 / g.snu), jx=ifft2(d * qx), jy=ifft2(d * qy)))
            resp = np.zeros_like(spec)
            for i, L in enumerate(banks):
                for j, R in enumerate(banks):
                    sd = fft2(2.0 * L["v"] * R["vn2"] + L["vn"] * R["vn"] - L["hx"] * R["hx"] - L["hy"] * R["hy"])
                    sk = fft2(2.0 * L["v"] * R["rad"] + 2.0 * (L["hx"] * R["jx"] + L["hy"] * R["jy"]))
                    cs = shifts[i] + shifts[j]
                    if branch == "slow":
                        fac = np.exp((g.sa - cs) * tau)
                    else:
                        fac = np.exp(-(g.sa + cs) * tau)
                    pr = fac * (sd + sgn * sk / g.snu)
                    pr[~pairs[i][j]] = 0.0
                    resp += pr
            res += 1j * coef * resp
    e11 = ifft2(spec)
    p1z = -1j * ifft2(spec * g.nu)
    total = res + fft2(0.5 * e11 * p1z)
    total[~out_sup] = 0.0
    audit = {
        "implementation": "dual-branch-green-laplace-psi22",
        "quadrature_rank_total": 4,
        "quadrature_rank_per_branch": 2,
        "lambda_minus": float(scales[0]),
        "lambda_plus": float(scales[1]),
        "radial_balance_channel_count": len(masks),
        "below_model_parent_domain": below,
        "below_model_spectral_mass": below_mass,
        "interaction_enumeration": False,
    }
    return ifft2(total), audit


# --------------------------------------------------------------------------
# order 3
# --------------------------------------------------------------------------
def order3_multipliers(g: Geometry) -> dict[str, np.ndarray]:
    """Spectral multipliers of the order-3 graph (_order3.py:151-172)."""
    q2 = g.q**2
    return {
        "qx": g.qx, "qy": g.qy, "q2": q2, "nu": g.nu, "nu2": g.nu**2, "nu3": g.nu**3,
        "dx": 1j * g.qx, "dy": 1j * g.qy,
        "qx_nu": g.qx / g.snu, "qy_nu": g.qy / g.snu, "q2_nu": q2 / g.snu,
        "m_i_nu": -1j * g.nu, "nu_qx": g.nu * g.qx, "nu_qy": g.nu * g.qy,
        "m_i_q2_nu": -1j * q2 / g.snu, "m_i_q2nu": -1j * q2 * g.nu,
        "m_a": -g.a, "m_q2": -q2,
    }


def first_order_state(d: np.ndarray, m: dict) -> dict[str, np.ndarray]:
    """12 first-order surface fields of one damped spectrum (_order3.py:175-193)."""
    ifft2 = np.fft.ifft2
    s = {
        "eta": ifft2(d), "etax": ifft2(m["dx"] * d), "etay": ifft2(m["dy"] * d),
        "phix": ifft2(m["qx_nu"] * d), "phiy": ifft2(m["qy_nu"] * d), "phiz": ifft2(m["m_i_nu"] * d),
        "phixz": ifft2(m["nu_qx"] * d), "phiyz": ifft2(m["nu_qy"] * d),
        "phizz": ifft2(m["m_i_q2_nu"] * d), "phizzz": ifft2(m["m_i_q2nu"] * d),
        "phitz": ifft2(m["m_a"] * d), "phitzz": ifft2(m["m_q2"] * d),
    }
    s["eta_t"] = s["phiz"]
    s["phiz_t"] = s["phitz"]
    s["phizz_t"] = s["phitzz"]
    return s


def inner_pair_fields(d: np.ndarray, g: Geometry, m: dict, out_sup: np.ndarray,
                      x: np.ndarray, w: np.ndarray, lam2: float, c2: float) -> dict[str, np.ndarray]:
    """Inner GL pair state of one damped parent spectrum (_order3.py:196-272)."""
    ifft2, fft2 = np.fft.ifft2, np.fft.fft2
    e_hat = np.zeros_like(d)
    et_hat = np.zeros_like(d)
    p_hat = np.zeros_like(d)
    pt_hat = np.zeros_like(d)
    for xn, wn in zip(x, w):
        tau = xn / lam2
        dd = d * np.exp(-tau * (m["nu"] - c2))
        v = ifft2(dd)
        vn = ifft2(dd * m["nu"])
        vn2 = ifft2(dd * m["nu2"])
        vn3 = ifft2(dd * m["nu3"])
        hx = ifft2(dd * m["qx_nu"])
        hy = ifft2(dd * m["qy_nu"])
        jx = ifft2(dd * m["qx"])
        jy = ifft2(dd * m["qy"])
        rad = ifft2(dd * m["q2_nu"])
        q2f = ifft2(dd * m["q2"])
        njx = ifft2(dd * m["nu_qx"])
        njy = ifft2(dd * m["nu_qy"])
        srcs = [
            2.0 * v * vn2 + vn**2 - hx**2 - hy**2,
            2.0 * v * rad + 2.0 * (hx * jx + hy * jy),
            2.0 * v * vn3 + 4.0 * vn * vn2 - 2.0 * hx * jx - 2.0 * hy * jy,
            2.0 * (vn * rad + v * q2f) + 2.0 * (jx**2 + hx * njx + jy**2 + hy * njy),
        ]
        sd, sk, sdt, skt = (floor_roundoff(_masked(fft2(s), out_sup), out_sup) for s in srcs)
        sh = sinh_ratio(g.sa, tau)
        ch = np.cosh(g.sa * tau)
        coef = wn * math.exp(xn) / lam2 * math.exp(-2.0 * c2 * tau) / 4.0
        e_hat += coef * (-g.a * sh * sd + ch * sk)
        et_hat -= 1j * coef * (-g.a * sh * sdt + ch * skt)
        p_hat += 1j * coef * (ch * sd - sh * sk)
        pt_hat += coef * (ch * sdt - sh * skt)
    for arr in (e_hat, et_hat, p_hat, pt_hat):
        arr[~out_sup] = 0.0
    et_hat = floor_roundoff(et_hat, out_sup)
    pair = {
        "eta": ifft2(e_hat), "etax": ifft2(m["dx"] * e_hat), "etay": ifft2(m["dy"] * e_hat),
        "phix": ifft2(m["dx"] * p_hat), "phiy": ifft2(m["dy"] * p_hat),
        "phiz": ifft2(g.a * p_hat), "phizz": ifft2(m["q2"] * p_hat),
        "phitz": ifft2(g.a * pt_hat), "eta_t": ifft2(et_hat),
    }
    pair["phiz_t"] = pair["phitz"]
    return pair


def _masked(spec: np.ndarray, sup: np.ndarray) -> np.ndarray:
    spec[~sup] = 0.0
    return spec


def taylor_contact(f: dict, p: dict):
    """Surface Taylor contact and its time derivative (_order3.py:275-287)."""
    c = 0.5 * f["eta"] * p["phiz"] + 0.5 * p["eta"] * f["phiz"] + 0.125 * f["eta"] ** 2 * f["phizz"]
    ct = (0.5 * (f["eta_t"] * p["phiz"] + f["eta"] * p["phiz_t"])
          + 0.5 * (p["eta_t"] * f["phiz"] + p["eta"] * f["phiz_t"])
          + 0.25 * f["eta"] * f["eta_t"] * f["phizz"]
          + 0.125 * f["eta"] ** 2 * f["phizz_t"])
    return c, ct


def order3_forcing(f: dict, p: dict):
    """Cubic kinematic and dynamic forcing (_order3.py:76-102)."""
    k_pair = (f["phix"] * p["etax"] + f["phiy"] * p["etay"] + p["phix"] * f["etax"] + p["phiy"] * f["etay"]
              - f["eta"] * p["phizz"] - p["eta"] * f["phizz"])
    k_direct = f["eta"] * (f["phixz"] * f["etax"] + f["phiyz"] * f["etay"]) - 0.5 * f["eta"] ** 2 * f["phizzz"]
    d_pair = (f["eta"] * p["phitz"] + p["eta"] * f["phitz"] + f["phix"] * p["phix"] + f["phiy"] * p["phiy"]
              + f["phiz"] * p["phiz"])
    d_direct = 0.5 * f["eta"] ** 2 * f["phitzz"] + f["eta"] * (
        f["phix"] * f["phixz"] + f["phiy"] * f["phiyz"] + f["phiz"] * f["phizz"])
    return k_direct / 4.0 + k_pair / 2.0, d_direct / 4.0 + d_pair / 2.0


@dataclass
class Order3Scalars:
    support: np.ndarray
    pair_support: np.ndarray
    triple_support: np.ndarray
    inner_shift: float
    outer_shift: float
    peak_q: float
    lambda2: float
    lambda3: float


def order3_scalars(spec: np.ndarray, g: Geometry, peak_q: float | None = None) -> Order3Scalars:
    """Supports, balance shifts and scales of the order-3 graph (_order3.py:28-52)."""
    sup = parent_support(spec)
    if not np.any(sup) or np.any(g.qx[sup] <= 0) or np.any(g.q[sup] <= 0.5):
        raise ValueError("third-order reconstruction requires strict-forward parent kh > 0.5")
    alias_guard(sup, 3)
    s2 = sum_support(sup, 2)
    s3 = sum_support(sup, 3)
    nu_min = float(np.min(g.nu[sup]))
    c2 = 0.5 * (nu_min + float(np.max(g.sa[s2])) / 2.0)
    c3 = 0.5 * (nu_min + float(np.max(g.sa[s3])) / 3.0)
    if peak_q is None:
        peak_q = peak_q_of(spec, sup, g.q)
    pn = math.sqrt(peak_q * math.tanh(peak_q))
    l2 = 2.0 * pn - math.sqrt(2.0 * peak_q * math.tanh(2.0 * peak_q))
    l3 = 3.0 * pn - math.sqrt(3.0 * peak_q * math.tanh(3.0 * peak_q))
    if l2 <= 0 or l3 <= 0:
        raise ValueError("invalid third-order Green--Laplace scales")
    return Order3Scalars(sup, s2, s3, c2, c3, peak_q, l2, l3)


def order3(spec: np.ndarray, qx: np.ndarray, qy: np.ndarray, inner_rank: int = 4, outer_rank: int = 4,
           peak_q: float | None = None):
    """Nested no-Stokes GL order 3: eta33+, psi33+ (dimensionless).

    Restates python/src/green_laplace/_order3.py:18-148 (MATLAB
    src/internal/gl_no_stokes_eta33.m:1-274).
    """
    if inner_rank not in (4, 6, 8, 10, 12) or outer_rank not in (3, 4, 6, 8, 10):
        raise ValueError("unsupported inner or outer third-order GL rank")
    g = Geometry(qx, qy)
    sc = order3_scalars(spec, g, peak_q)
    xi, wi = gauss_laguerre_rule(inner_rank)
    xo, wo = gauss_laguerre_rule(outer_rank)
    m = order3_multipliers(g)
    s3 = sc.triple_support
    fft2 = np.fft.fft2
    eta_hat = np.zeros_like(spec)
    psi_hat = np.zeros_like(spec)
    for xn, wn in zip(xo, wo):
        tau = xn / sc.lambda3
        d = spec * np.exp(-tau * (g.nu - sc.outer_shift))
        f = first_order_state(d, m)
        p = inner_pair_fields(d, g, m, sc.pair_support, xi, wi, sc.lambda2, sc.inner_shift)
        fk, fd = order3_forcing(f, p)
        fk_hat = floor_roundoff(_masked(fft2(fk), s3), s3)
        fd_hat = floor_roundoff(_masked(fft2(fd), s3), s3)
        c, ct = taylor_contact(f, p)
        c_hat = floor_roundoff(_masked(fft2(c), s3), s3)
        ct_hat = floor_roundoff(_masked(fft2(ct), s3), s3)
        fk_psi = floor_roundoff(fk_hat + g.a * c_hat, s3)
        fd_psi = floor_roundoff(fd_hat - ct_hat, s3)
        sh = sinh_ratio(g.sa, tau)
        ch = np.cosh(g.sa * tau)
        coef = wn * math.exp(xn) / sc.lambda3 * math.exp(-3.0 * sc.outer_shift * tau)
        eta_hat += coef * (g.a * sh * fd_hat - 1j * ch * fk_hat)
        psi_hat += coef * (-1j * ch * fd_psi - sh * fk_psi)
    eta_hat[~s3] = 0.0
    psi_hat[~s3] = 0.0
    audit = {
        "implementation": "nested-no-stokes-green-laplace-order3",
        "inner_quadrature_rank": inner_rank,
        "outer_quadrature_rank": outer_rank,
        "inner_scale": float(sc.lambda2),
        "outer_scale": float(sc.lambda3),
        "inner_balance_shift": sc.inner_shift,
        "outer_balance_shift": sc.outer_shift,
        "interaction_enumeration": False,
        "core_fft_ifft_count": outer_rank * (12 + 16 * inner_rank + 9 + 4) + 2,
    }
    return np.fft.ifft2(eta_hat), np.fft.ifft2(psi_hat), audit


def order3_spectra(spec, qx, qy, inner_rank=4, outer_rank=4, peak_q=None):
    """Like ``order3`` but returns the masked eta/psi spectra before the final IFFT."""
    e, p, a = order3(spec, qx, qy, inner_rank, outer_rank, peak_q)
    return np.fft.fft2(e), np.fft.fft2(p), a


# --------------------------------------------------------------------------
# public APIs
# --------------------------------------------------------------------------
def linear_fields(spec: np.ndarray, nu: np.ndarray, depth: float, gravity: float):
    """eta11+ and psi11+ (spatial.py:74-79; gl_spectral_surface.m:26-31)."""
    e = depth * np.fft.ifft2(spec)
    ps = np.zeros_like(spec)
    nz = nu > 0
    ps[nz] = -1j * spec[nz] / nu[nz]
    return e, depth * math.sqrt(gravity * depth) * np.fft.ifft2(ps)


def reconstruct_directional_field(eta1, *, lx, ly, depth, gravity=9.81, time=0.0, order=2, eta22_rank=6,
                                  allow_below_model_parent_domain=False) -> dict[str, Any]:
    """Python API restatement (python/src/green_laplace/spatial.py:33-153).

    Returns a dict with keys x, y, eta, psi, components, audit.
    """
    fld = np.asarray(eta1, dtype=float)
    if fld.ndim != 2:
        raise ValueError("eta1 must be a two-dimensional array")
    if not np.isfinite(time) or depth <= 0 or gravity <= 0:
        raise ValueError("time must be finite; depth and gravity must be positive")
    if order not in (1, 2, 3):
        raise ValueError("order must be 1, 2, or 3")
    if not isinstance(eta22_rank, int) or eta22_rank < 1 or eta22_rank > 16:
        raise ValueError("eta22_rank must be an integer from 1 to 16")
    ny, nx = fld.shape
    kx, ky = wavenumber_grid(lx, ly, nx, ny)
    s0, in_err = analytic_spectrum(fld, depth, kx)
    qx, qy = depth * kx, depth * ky
    g = Geometry(qx, qy)
    spec = s0 * np.exp(-1j * math.sqrt(gravity / depth) * g.nu * time)
    return _assemble(spec, qx, qy, g, depth, gravity, order, eta22_rank, allow_below_model_parent_domain,
                     lx, ly, time, in_err, None)


def _assemble(spec, qx, qy, g, depth, gravity, order, rank, allow_below, lx, ly, time, in_err, peak_q):
    ny, nx = spec.shape
    e11, p11 = linear_fields(spec, g.nu, depth, gravity)
    pot = depth * math.sqrt(gravity * depth)
    comps: dict[str, np.ndarray] = {"eta11_plus": e11, "psi11_plus": p11, "eta11": e11.real, "psi11": p11.real}
    e_tot, p_tot = e11.copy(), p11.copy()
    audits: dict[str, Any] = {}
    if order >= 2:
        e22, a22 = eta22(spec, qx, qy, rank, peak_q)
        p22, ap22 = psi22(spec, qx, qy, allow_below, peak_q)
        comps.update(eta22_plus=depth * e22, eta22=(depth * e22).real, psi22_plus=pot * p22, psi22=(pot * p22).real)
        e_tot += depth * e22
        p_tot += pot * p22
        audits["eta22"] = a22
        audits["psi22"] = ap22
    if order >= 3:
        e33, p33, a33 = order3(spec, qx, qy, peak_q=peak_q)
        comps.update(eta33_plus=depth * e33, eta33=(depth * e33).real, psi33_plus=pot * p33, psi33=(pot * p33).real)
        e_tot += depth * e33
        p_tot += pot * p33
        audits["order3"] = a33
    xs = np.arange(nx, dtype=float) * lx / nx
    ys = np.arange(ny, dtype=float) * ly / ny
    x, y = np.meshgrid(xs, ys)
    comps.update(eta_plus=e_tot, psi_plus=p_tot, eta=e_tot.real, psi=p_tot.real)
    audit = {"order": order, "grid": [ny, nx], "domain": [ly, lx], "time": float(time),
             "input_reconstruction_relative": in_err, "component_audits": audits}
    return {"x": x, "y": y, "eta": e_tot.real, "psi": p_tot.real, "components": comps, "audit": audit}


def deposit(a, b, kx, ky, omega, h, lx, ly, nx, ny, t, grid_tolerance=1e-10):
    """Analytic spectrum of MF12-style components (src/gl_spectral_surface.m:123-149)."""
    dkx, dky = 2.0 * np.pi / lx, 2.0 * np.pi / ly
    spec = np.zeros((ny, nx), dtype=np.complex128)
    for i in range(len(a)):
        mx = int(np.round(kx[i] / dkx))
        my = int(np.round(ky[i] / dky))
        tol = grid_tolerance * max(1.0, math.hypot(kx[i], ky[i]))
        if abs(kx[i] - mx * dkx) > tol or abs(ky[i] - my * dky) > tol:
            raise ValueError("green_laplace:OffGridComponent")
        if abs(mx) >= nx / 2 or abs(my) >= ny / 2:
            raise ValueError("green_laplace:Nyquist")
        spec[my % ny, mx % nx] += nx * ny * complex(a[i], b[i]) / h * np.exp(-1j * omega[i] * t)
    return spec


def gl_spectral_surface(order, g, h, a, b, kx, ky, lx, ly, nx, ny, t, eta22_rank=6,
                        allow_below_model_parent_domain=False):
    """MATLAB two-stage API restatement (src/gl_spectral_coefficients.m:1-88 and
    src/gl_spectral_surface.m:1-102), with peak q from the largest a^2+b^2."""
    a, b, kx, ky = (np.asarray(v, dtype=float).ravel() for v in (a, b, kx, ky))
    kabs = np.hypot(kx, ky)
    omega = np.sqrt(g * kabs * np.tanh(h * kabs))
    peak = int(np.argmax(a**2 + b**2))
    peak_q = h * float(kabs[peak])
    kxg, kyg = wavenumber_grid(lx, ly, nx, ny)
    spec = deposit(a, b, kx, ky, omega, h, lx, ly, nx, ny, t)
    qx, qy = h * kxg, h * kyg
    return _assemble(spec, qx, qy, Geometry(qx, qy), h, g, order, eta22_rank, allow_below_model_parent_domain,
                     lx, ly, t, 0.0, peak_q)


# --------------------------------------------------------------------------
# order 4 (eta44, no Stokes corrections)
# --------------------------------------------------------------------------
def _suppress(x: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """Mask then roundoff floor (paper/order4/cpp/gl_eta44_fftw.cpp:157-171)."""
    y = np.where(mask, x, 0.0)
    top = float(np.max(np.abs(y[mask]), initial=0.0))
    if top:
        level = 16.0 * EPS * max(1.0, math.log2(x.size)) * top
        y[mask & (np.abs(y) < level)] = 0.0
    return y


class _O4:
    """Geometry, masks, shifts and scales of the order-4 graph (gl_eta44_fftw.cpp:76-155)."""

    def __init__(self, spec, qx, qy, peak, nodes, weights):
        self.g = Geometry(qx, qy)
        self.nodes, self.weights = nodes, weights
        sup = parent_support(spec)
        g = self.g
        if np.any(qx[sup] <= 0) or np.any(g.q[sup] <= 0):
            raise ValueError("support")
        self.m2, self.m3, self.m4 = (sum_support(sup, n) for n in (2, 3, 4))
        mn = float(np.min(g.nu[sup]))
        self.c2 = 0.5 * (mn + float(np.max(g.sa[self.m2])) / 2)
        self.c3 = 0.5 * (mn + float(np.max(g.sa[self.m3])) / 3)
        self.c4 = 0.5 * (mn + float(np.max(g.sa[self.m4])) / 4)
        pn = math.sqrt(peak * math.tanh(peak))
        self.l2 = 2 * pn - math.sqrt(2 * peak * math.tanh(2 * peak))
        self.l3 = 3 * pn - math.sqrt(3 * peak * math.tanh(3 * peak))
        self.l4 = 4 * pn - math.sqrt(4 * peak * math.tanh(4 * peak))

    def factors(self, tau):
        return sinh_ratio(self.g.sa, tau), np.cosh(self.g.sa * tau)


def _first4(G: _O4, s, order4=False):
    """First-order fields with time derivatives (gl_eta44_fftw.cpp:188-290)."""
    g, ifft2 = G.g, np.fft.ifft2
    q2 = g.q**2
    f = {
        "eta": ifft2(s), "ex": ifft2(1j * g.qx * s), "ey": ifft2(1j * g.qy * s),
        "px": ifft2(g.qx / g.snu * s), "py": ifft2(g.qy / g.snu * s), "pz": ifft2(-1j * g.nu * s),
        "pxz": ifft2(g.nu * g.qx * s), "pyz": ifft2(g.nu * g.qy * s), "pzz": ifft2(-1j * q2 / g.snu * s),
        "pzzz": ifft2(-1j * q2 * g.nu * s), "ptz": ifft2(-g.a * s), "ptzz": ifft2(-q2 * s),
        "pxz_t": ifft2(-1j * g.nu**2 * g.qx * s), "pyz_t": ifft2(-1j * g.nu**2 * g.qy * s),
        "pzzz_t": ifft2(-q2 * g.a * s), "ptz_t": ifft2(1j * g.a * g.nu * s),
    }
    if order4:
        phi = np.zeros_like(s)
        nz = g.nu > 0
        phi[nz] = -1j * s[nz] / g.nu[nz]
        f["pxzz"] = ifft2(1j * g.qx * q2 * phi)
        f["pyzz"] = ifft2(1j * g.qy * q2 * phi)
        f["pzzzz"] = ifft2(q2 * q2 * phi)
    return f


def _pair4(G: _O4, s, order4=False):
    """Nested pair state with first and second time derivatives (gl_eta44_fftw.cpp:299-529)."""
    g, ifft2, fft2 = G.g, np.fft.ifft2, np.fft.fft2
    q2 = g.q**2
    es, ets, ps, pts, ptts = (np.zeros_like(s) for _ in range(5))
    for x, w in zip(G.nodes, G.weights):
        t = x / G.l2
        d = s * np.exp(-t * (g.nu - G.c2))
        v, vn, vn2, vn3, vn4 = (ifft2(d * g.nu**p) for p in range(5))
        hx, hy = ifft2(d * g.qx / g.snu), ifft2(d * g.qy / g.snu)
        jx, jy = ifft2(d * g.qx), ifft2(d * g.qy)
        ro, q2f, q2n = ifft2(d * q2 / g.snu), ifft2(d * q2), ifft2(d * q2 * g.nu)
        njx, njy = ifft2(d * g.nu * g.qx), ifft2(d * g.nu * g.qy)
        n2jx, n2jy = ifft2(d * g.nu**2 * g.qx), ifft2(d * g.nu**2 * g.qy)
        sd = 2 * v * vn2 + vn * vn - hx * hx - hy * hy
        sk = 2 * v * ro + 2 * (hx * jx + hy * jy)
        sds = 2 * v * vn3 + 4 * vn * vn2 - 2 * hx * jx - 2 * hy * jy
        sks = 2 * (vn * ro + v * q2f) + 2 * (jx * jx + hx * njx + jy * jy + hy * njy)
        sdss = 2 * v * vn4 + 6 * vn * vn3 + 4 * vn2 * vn2 - 2 * (jx * jx + hx * njx + jy * jy + hy * njy)
        skss = 2 * (vn2 * ro + 2 * vn * q2f + v * q2n) + 6 * (jx * njx + jy * njy) + 2 * (hx * n2jx + hy * n2jy)
        sd, sk, sds, sks, sdss, skss = (_suppress(fft2(a), G.m2) for a in (sd, sk, sds, sks, sdss, skss))
        co = w * math.exp(x) / G.l2 * math.exp(-2 * G.c2 * t) / 4
        sh, ch = G.factors(t)
        es += co * (-g.a * sh * sd + ch * sk)
        ets -= 1j * co * (-g.a * sh * sds + ch * sks)
        ps += 1j * co * (ch * sd - sh * sk)
        pts += co * (ch * sds - sh * sks)
        ptts -= 1j * co * (ch * sdss - sh * skss)
    es, ets, ps, pts, ptts = (_suppress(a, G.m2) for a in (es, ets, ps, pts, ptts))
    p = {
        "eta": ifft2(es), "ex": ifft2(1j * g.qx * es), "ey": ifft2(1j * g.qy * es),
        "eta_t": ifft2(ets), "ex_t": ifft2(1j * g.qx * ets), "ey_t": ifft2(1j * g.qy * ets),
        "px": ifft2(1j * g.qx * ps), "py": ifft2(1j * g.qy * ps), "pz": ifft2(g.a * ps), "pzz": ifft2(q2 * ps),
        "ptz": ifft2(g.a * pts), "px_t": ifft2(1j * g.qx * pts), "py_t": ifft2(1j * g.qy * pts),
        "pzz_t": ifft2(q2 * pts), "ptz_t": ifft2(g.a * ptts),
    }
    if order4:
        p["pxz"] = ifft2(1j * g.qx * g.a * ps)
        p["pyz"] = ifft2(1j * g.qy * g.a * ps)
        p["pzzz"] = ifft2(q2 * g.a * ps)
    return p


def _third4(G: _O4, s):
    """Nested third-order state (gl_eta44_fftw.cpp:537-637)."""
    g, ifft2, fft2 = G.g, np.fft.ifft2, np.fft.fft2
    q2 = g.q**2
    es, ets, ps, pts = (np.zeros_like(s) for _ in range(4))
    for x, w in zip(G.nodes, G.weights):
        t = x / G.l3
        d = s * np.exp(-t * (g.nu - G.c3))
        f, p = _first4(G, d), _pair4(G, d)
        kp = f["px"] * p["ex"] + f["py"] * p["ey"] + p["px"] * f["ex"] + p["py"] * f["ey"] - f["eta"] * p["pzz"] \
            - p["eta"] * f["pzz"]
        kd = f["eta"] * (f["pxz"] * f["ex"] + f["pyz"] * f["ey"]) - 0.5 * f["eta"] ** 2 * f["pzzz"]
        dp = f["eta"] * p["ptz"] + p["eta"] * f["ptz"] + f["px"] * p["px"] + f["py"] * p["py"] + f["pz"] * p["pz"]
        dd = 0.5 * f["eta"] ** 2 * f["ptzz"] + f["eta"] * (f["px"] * f["pxz"] + f["py"] * f["pyz"] + f["pz"] * f["pzz"])
        kpt = (-f["ex"] * p["ex"] + f["px"] * p["ex_t"] - f["ey"] * p["ey"] + f["py"] * p["ey_t"]
               + p["px_t"] * f["ex"] + p["px"] * f["pxz"] + p["py_t"] * f["ey"] + p["py"] * f["pyz"]
               - f["pz"] * p["pzz"] - f["eta"] * p["pzz_t"] - p["eta_t"] * f["pzz"] - p["eta"] * f["ptzz"])
        kdt = (f["pz"] * (f["pxz"] * f["ex"] + f["pyz"] * f["ey"])
               + f["eta"] * (f["pxz_t"] * f["ex"] + f["pxz"] * f["pxz"] + f["pyz_t"] * f["ey"] + f["pyz"] * f["pyz"])
               - f["eta"] * f["pz"] * f["pzzz"] - 0.5 * f["eta"] ** 2 * f["pzzz_t"])
        dpt = (f["pz"] * p["ptz"] + f["eta"] * p["ptz_t"] + p["eta_t"] * f["ptz"] + p["eta"] * f["ptz_t"]
               - f["ex"] * p["px"] + f["px"] * p["px_t"] - f["ey"] * p["py"] + f["py"] * p["py_t"]
               + f["ptz"] * p["pz"] + f["pz"] * p["ptz"])
        ddt = (f["eta"] * f["pz"] * f["ptzz"] - 0.5 * f["eta"] ** 2 * f["pzzz"]
               + f["pz"] * (f["px"] * f["pxz"] + f["py"] * f["pyz"] + f["pz"] * f["pzz"])
               + f["eta"] * (-f["ex"] * f["pxz"] + f["px"] * f["pxz_t"] - f["ey"] * f["pyz"] + f["py"] * f["pyz_t"]
                             + f["ptz"] * f["pzz"] + f["pz"] * f["ptzz"]))
        fk, fd, fkt, fdt = (_suppress(fft2(a), G.m3) for a in (kd / 4 + kp / 2, dd / 4 + dp / 2,
                                                               kdt / 4 + kpt / 2, ddt / 4 + dpt / 2))
        co = w * math.exp(x) / G.l3 * math.exp(-3 * G.c3 * t)
        sh, ch = G.factors(t)
        es += co * (g.a * sh * fd - 1j * ch * fk)
        ets += co * (g.a * sh * fdt - 1j * ch * fkt)
        ps += co * (-1j * ch * fd - sh * fk)
        pts += co * (-1j * ch * fdt - sh * fkt)
    es, ps, pts = (_suppress(a, G.m3) for a in (es, ps, pts))
    return {"eta": ifft2(es), "ex": ifft2(1j * g.qx * es), "ey": ifft2(1j * g.qy * es),
            "px": ifft2(1j * g.qx * ps), "py": ifft2(1j * g.qy * ps), "pz": ifft2(g.a * ps),
            "pzz": ifft2(q2 * ps), "pzt": ifft2(g.a * pts)}


def eta44(spec: np.ndarray, qx: np.ndarray, qy: np.ndarray, peak_q: float, rank: int):
    """Pure-sum GL eta44+ (dimensionless), no Stokes corrections.

    Restates paper/order4/cpp/gl_eta44_fftw.cpp:639-752 (itself a restatement
    of src/internal/finite_depth_directional_green_laplace_eta44_three_kernels.m:57-219
    with finite_depth_directional_nested_green_laplace_eta33.m); one GL rank for
    every nesting level.
    """
    nodes, weights = gauss_laguerre_rule(rank)
    G = _O4(spec, qx, qy, peak_q, nodes, weights)
    g, fft2 = G.g, np.fft.fft2
    out = np.zeros_like(spec)
    for x, w in zip(nodes, weights):
        tm = x / G.l4
        d = spec * np.exp(-tm * (g.nu - G.c4))
        t = _third4(G, d)
        f = _first4(G, d, True)
        p = _pair4(G, d, True)
        k31 = f["px"] * t["ex"] + f["py"] * t["ey"] + t["px"] * f["ex"] + t["py"] * f["ey"] - f["eta"] * t["pzz"] \
            - t["eta"] * f["pzz"]
        d31 = f["eta"] * t["pzt"] + t["eta"] * f["ptz"] + f["px"] * t["px"] + f["py"] * t["py"] + f["pz"] * t["pz"]
        k22 = p["px"] * p["ex"] + p["py"] * p["ey"] - p["eta"] * p["pzz"]
        d22 = p["eta"] * p["ptz"] + 0.5 * (p["px"] ** 2 + p["py"] ** 2 + p["pz"] ** 2)
        fe = f["eta"]
        k211 = (fe * (f["pxz"] * p["ex"] + f["pyz"] * p["ey"]) + (fe * p["pxz"] + p["eta"] * f["pxz"]) * f["ex"]
                + (fe * p["pyz"] + p["eta"] * f["pyz"]) * f["ey"] - 0.5 * fe * fe * p["pzzz"]
                - fe * p["eta"] * f["pzzz"])
        d211 = (0.5 * fe * fe * p["pzz_t"] + fe * p["eta"] * f["ptzz"] + f["px"] * (fe * p["pxz"] + p["eta"] * f["pxz"])
                + f["py"] * (fe * p["pyz"] + p["eta"] * f["pyz"]) + f["pz"] * (fe * p["pzz"] + p["eta"] * f["pzz"])
                + p["px"] * (fe * f["pxz"]) + p["py"] * (fe * f["pyz"]) + p["pz"] * (fe * f["pzz"]))
        k1111 = 0.5 * fe * fe * (f["pxzz"] * f["ex"] + f["pyzz"] * f["ey"]) - fe**3 * f["pzzzz"] / 6
        d1111 = (fe**3 * f["pzzz_t"] / 6 + 0.5 * fe * fe * (f["px"] * f["pxzz"] + f["py"] * f["pyzz"] + f["pz"] * f["pzzz"])
                 + 0.5 * fe * fe * (f["pxz"] ** 2 + f["pyz"] ** 2 + f["pzz"] ** 2))
        fk = _suppress(fft2(k31 / 2 + k1111 / 8 + k211 / 4 + k22 / 2), G.m4)
        fd = _suppress(fft2(d31 / 2 + d1111 / 8 + d211 / 4 + d22 / 2), G.m4)
        co = w * math.exp(x) / G.l4 * math.exp(-4 * G.c4 * tm)
        sh, ch = G.factors(tm)
        out += co * (g.a * sh * fd - 1j * ch * fk)
    out = _suppress(out, G.m4)
    return np.fft.ifft2(out)
