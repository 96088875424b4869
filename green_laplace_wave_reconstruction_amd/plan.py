"""Host FP64 plan of a sea state: supports, boxes, GL rules and tables.

Everything here is t-independent setup that the reference recomputes on every
call (python/src/green_laplace/_order3.py:28-55, spatial.py:162-190 and
241-275) or caches per grid (src/internal/gl_no_stokes_eta33.m:535-621).  The
GPU kernels only consume these tables; no field, product or transform of the
hot path is computed here.

Spectral boxes and working grids
--------------------------------
The parent support P1 (|S| > 1e-13 max(1, max|S|)) lies in a rectangle of
signed modes ``box1``.  Pair (P2) and triple (P3) supports are the Minkowski
sums P1+P1 and P1+P1+P1 (computed like the reference, by thresholding
``ifft2(fft2(mask)^n)``); their bounding boxes have widths ``2w-1`` and
``3w-2``.  Physical fields are sampled on working grids ``g2`` (quadratic
products) and ``g3`` (cubic products) at least that wide, which makes the
pseudo-spectral products exact on every kept bin: bin ``(mx, my)`` of the
``ny x nx`` grid maps one-to-one to ``(my mod gy, mx mod gx)``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np

TINY = np.finfo(float).tiny
EPS = np.finfo(float).eps

#: FFT lengths the HIP kernels are instantiated for (2^a 3^b).
FFT_LENGTHS = (16, 24, 32, 48, 64, 96, 128, 144, 192, 256, 288, 384, 512, 576, 768, 1024, 1152, 1536, 2048)

PREC_FP64 = 0
PREC_FP32 = 1


def gauss_laguerre_rule(rank: int) -> tuple[np.ndarray, np.ndarray]:
    """Golub--Welsch Gauss--Laguerre rule with the reference moment gate.

    Same rule as python/src/green_laplace/quadrature.py:11-31: eigenpairs of
    the Jacobi matrix (diagonal 2i-1, off-diagonal i), weights from the first
    eigenvector component; moments sum(w x^d) = d! checked for d < 2 rank.
    """
    if not isinstance(rank, (int, np.integer)) or rank < 1:
        raise ValueError("rank must be a positive integer")
    rank = int(rank)
    diag = 2.0 * np.arange(1, rank + 1) - 1.0
    jac = np.diag(diag)
    if rank > 1:
        off = np.arange(1, rank, dtype=float)
        jac = jac + np.diag(off, 1) + np.diag(off, -1)
    nodes, vecs = np.linalg.eigh(jac)
    weights = vecs[0, :] ** 2
    for d in range(2 * rank):
        exact = math.factorial(d)
        if abs(float(np.sum(weights * nodes**d)) - exact) / max(1, exact) > 3.0e-11:
            raise ArithmeticError(f"Gauss--Laguerre moment check failed at rank={rank}, degree={d}")
    return nodes, weights


def psi22_rule() -> tuple[np.ndarray, np.ndarray]:
    """Two-node rule of the psi22 branches: nodes 2 -/+ sqrt(2) (spatial.py:269-275)."""
    r2 = math.sqrt(2.0)
    nodes = np.array([2.0 - r2, 2.0 + r2])
    weights = np.array([nodes[0] / (4.0 * (r2 - 1.0) ** 2), nodes[1] / (4.0 * (r2 + 1.0) ** 2)])
    return nodes, weights


def fft_length_at_least(n: int) -> int:
    for length in FFT_LENGTHS:
        if length >= n:
            return length
    raise ValueError(f"no supported FFT length >= {n} (largest is {FFT_LENGTHS[-1]})")


@dataclass(frozen=True)
class Box:
    """Signed modes mx0..mx0+wx-1, my0..my0+wy-1."""

    mx0: int
    my0: int
    wx: int
    wy: int

    @property
    def size(self) -> int:
        return self.wx * self.wy

    def index(self, ny: int, nx: int):
        """(rows, cols) index arrays into a full [ny, nx] grid."""
        my = np.arange(self.my0, self.my0 + self.wy) % ny
        mx = np.arange(self.mx0, self.mx0 + self.wx) % nx
        return np.ix_(my, mx)

    def take(self, full: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(full[self.index(*full.shape)])

    def put(self, values: np.ndarray, shape) -> np.ndarray:
        out = np.zeros(shape, dtype=values.dtype)
        out[self.index(*shape)] = values
        return out


def mode_grids(ny: int, nx: int):
    """Signed integer mode numbers (mx, my) of a [ny, nx] FFT grid.

    ``fftfreq(n) * n`` is not exact for n that are not powers of two
    (28/192*192 = 27.999999999999996), so the products are rounded.
    """
    mx = np.rint(np.fft.fftfreq(nx) * nx).astype(np.int64)
    my = np.rint(np.fft.fftfreq(ny) * ny).astype(np.int64)
    return np.meshgrid(mx, my)


def box_of(mask: np.ndarray) -> Box:
    ny, nx = mask.shape
    mx, my = mode_grids(ny, nx)
    mxs = mx[mask]
    mys = my[mask]
    return Box(int(mxs.min()), int(mys.min()), int(mxs.max() - mxs.min() + 1), int(mys.max() - mys.min() + 1))


def _smooth_length(n: int) -> int:
    """Smallest 2^a 3^b 5^c >= n (a fast pocketfft length)."""
    m = n
    while True:
        r = m
        for f in (2, 3, 5):
            while r % f == 0:
                r //= f
        if r == 1:
            return m
        m += 1


def minkowski_sum(masks: list) -> np.ndarray:
    """Minkowski sum of masks on their periodic [ny, nx] grid: the thresholded
    convolution count ``ifft2(prod fft2(mask)) > 0.5`` (_order3.py:38-40).

    When the sum of the masks' bounding boxes fits the grid, no bin of the sum
    wraps onto another, so the same count is formed on a grid of the boxes'
    size and placed back (a few thousand points instead of N^2: the host plan
    at 512^2 spent half its time in these transforms); otherwise on the full
    grid.
    """
    ny, nx = masks[0].shape
    if not all(np.any(m) for m in masks):
        return np.zeros((ny, nx), dtype=bool)
    boxes = [box_of(m) for m in masks]
    ex = sum(b.wx - 1 for b in boxes) + 1
    ey = sum(b.wy - 1 for b in boxes) + 1
    small = ex <= nx and ey <= ny
    if small:
        gy, gx = _smooth_length(ey), _smooth_length(ex)
    spectra = {}  # one transform per distinct mask
    acc = None
    for m, b in zip(masks, boxes):
        f = spectra.get(id(m))
        if f is None:
            if small:
                grid = np.zeros((gy, gx))
                grid[:b.wy, :b.wx] = b.take(m)
            else:
                grid = m.astype(float)
            f = spectra[id(m)] = np.fft.fft2(grid)
        acc = f if acc is None else acc * f
    count = np.real(np.fft.ifft2(acc)) > 0.5
    if not small:
        return count
    return Box(sum(b.mx0 for b in boxes), sum(b.my0 for b in boxes), ex, ey).put(count[:ey, :ex], (ny, nx))


def sum_support(mask: np.ndarray, n: int) -> np.ndarray:
    """n-fold Minkowski sum of a mask, thresholded convolution count (_order3.py:38-40)."""
    return minkowski_sum([mask] * n)


def _sinh_ratio(sa: np.ndarray, tau: float) -> np.ndarray:
    out = np.full_like(sa, tau)
    nz = sa > 0
    out[nz] = np.sinh(sa[nz] * tau) / sa[nz]
    return out


@dataclass
class BoxTables:
    qx: np.ndarray
    qy: np.ndarray
    nu: np.ndarray
    inu: np.ndarray
    q2: np.ndarray
    a: np.ndarray


def _tables(box: Box, qx_full: np.ndarray, qy_full: np.ndarray) -> BoxTables:
    qx = box.take(qx_full)
    qy = box.take(qy_full)
    q = np.hypot(qx, qy)
    a = q * np.tanh(q)
    nu = np.sqrt(q * np.tanh(q))
    return BoxTables(qx=qx, qy=qy, nu=nu, inu=1.0 / np.maximum(nu, TINY), q2=q**2, a=a)


@dataclass
class HostPlan:
    """All t-independent FP64 data of one sea state's GL graphs."""

    nx: int
    ny: int
    depth: float
    gravity: float
    order: int
    precision: int
    floor_level: float
    box1: Box
    box2: Box | None = None
    box3: Box | None = None
    g2: tuple[int, int] | None = None   # (gx, gy)
    g3: tuple[int, int] | None = None
    t1: BoxTables | None = None
    t2: BoxTables | None = None
    t3: BoxTables | None = None
    omega: np.ndarray | None = None
    mask2: np.ndarray | None = None
    mask3: np.ndarray | None = None
    peak_q: float = 0.0
    o2: dict[str, Any] = field(default_factory=dict)
    o3: dict[str, Any] = field(default_factory=dict)
    o4: dict[str, Any] = field(default_factory=dict)
    box4: Box | None = None
    g4: tuple[int, int] | None = None
    t4: BoxTables | None = None
    mask4: np.ndarray | None = None
    audits: dict[str, Any] = field(default_factory=dict)


def build_plan(spectrum: np.ndarray, qx: np.ndarray, qy: np.ndarray, *, depth: float, gravity: float,
               order: int, eta22_rank: int = 6, inner_rank: int = 4, outer_rank: int = 4,
               peak_q: float | None = None, allow_below_model_parent_domain: bool = False,
               precision: int = PREC_FP64) -> HostPlan:
    """Validate a t=0 analytic spectrum and build the FP64 plan for orders 1..order.

    ``spectrum`` is the full ``[ny, nx]`` analytic spectrum (reference
    convention: ``N^2 (a+ib)/h`` at strict-forward bins); ``qx = h kx`` and
    ``qy = h ky`` are the full dimensionless wavenumber grids.  Raises the
    reference's ``ValueError`` messages for unsupported inputs.
    """
    spectrum = np.asarray(spectrum, dtype=np.complex128)
    ny, nx = spectrum.shape
    q = np.hypot(qx, qy)
    mag = np.abs(spectrum)
    support = mag > 1.0e-13 * max(1.0, float(np.max(mag)))
    nu_full = np.sqrt(q * np.tanh(q))
    if peak_q is None and np.any(support):
        idx = np.flatnonzero(support)
        peak_q = float(q.flat[idx[int(np.argmax(mag.flat[idx] ** 2))]])
    plan = HostPlan(nx=nx, ny=ny, depth=float(depth), gravity=float(gravity), order=int(order),
                    precision=int(precision), floor_level=16.0 * EPS * max(1.0, math.log2(nx * ny)),
                    box1=box_of(support) if np.any(support) else Box(1, 0, 1, 1), peak_q=float(peak_q or 0.0))
    if order >= 2:
        _check_order2(spectrum, support, qx, q, allow_below_model_parent_domain)
    if order >= 3:
        _check_order3(support, qx, q, inner_rank, outer_rank)
    # Boxes: the parent box must also hold every bin the reference would carry
    # for the linear fields; bins outside the support box are below the
    # support threshold and are dropped (they only enter the reference graph
    # at <1e-13 relative size).
    plan.t1 = _tables(plan.box1, qx, qy)
    plan.omega = math.sqrt(gravity / depth) * plan.t1.nu
    if order >= 2:
        s2 = sum_support(support, 2)
        plan.box2 = box_of(s2)
        plan.g2 = (fft_length_at_least(max(2 * plan.box1.wx - 1, 16)),
                   fft_length_at_least(max(2 * plan.box1.wy - 1, 16)))
        plan.t2 = _tables(plan.box2, qx, qy)
        plan.mask2 = plan.box2.take(s2).astype(np.uint8)
        _order2_tables(plan, spectrum, support, s2, q, nu_full, eta22_rank, allow_below_model_parent_domain)
    if order >= 3:
        s3 = sum_support(support, 3)
        plan.box3 = box_of(s3)
        plan.g3 = (fft_length_at_least(max(3 * plan.box1.wx - 2, 16)),
                   fft_length_at_least(max(3 * plan.box1.wy - 2, 16)))
        plan.t3 = _tables(plan.box3, qx, qy)
        plan.mask3 = plan.box3.take(s3).astype(np.uint8)
        _order3_tables(plan, support, s2, s3, q, nu_full, inner_rank, outer_rank)
    return plan


def _alias_guard(support: np.ndarray, order: int) -> None:
    ny, nx = support.shape
    mx, my = mode_grids(ny, nx)
    if not np.any(support):
        raise ValueError("eta1 has empty analytic support")
    if np.max(order * np.abs(mx[support])) >= nx / 2.0 or np.max(order * np.abs(my[support])) >= ny / 2.0:
        word = "quadratic" if order == 2 else "cubic"
        raise ValueError(f"{word} output support reaches a Nyquist boundary")


def _check_order2(spectrum, support, qx, q, allow_below):
    if np.any(qx[support] <= 0) or np.any(q[support] <= 0):
        raise ValueError("strict-forward nonzero analytic eta1 support is required")
    _alias_guard(support, 2)
    if np.any(q[support] < 0.3) and not allow_below:
        raise ValueError("psi22 requires parent kh >= 0.3")


def _check_order3(support, qx, q, inner_rank, outer_rank):
    if inner_rank not in (4, 6, 8, 10, 12) or outer_rank not in (3, 4, 6, 8, 10):
        raise ValueError("unsupported inner or outer third-order GL rank")
    if not np.any(support) or np.any(qx[support] <= 0) or np.any(q[support] <= 0.5):
        raise ValueError("third-order reconstruction requires strict-forward parent kh > 0.5")
    _alias_guard(support, 3)


def _order2_tables(plan: HostPlan, spectrum, support, s2, q, nu_full, rank, allow_below):
    """eta22 (spatial.py:156-221) and psi22 (spatial.py:234-394) tables."""
    t1, t2 = plan.t1, plan.t2
    qp = plan.peak_q
    disc = (2.0 * math.sqrt(qp * math.tanh(qp))) ** 2 - 2.0 * qp * math.tanh(2.0 * qp)
    lam = math.sqrt(disc) if disc > 0 else float("nan")
    if not np.isfinite(lam) or lam <= 0:
        raise ValueError("invalid determinant-root Green--Laplace scale")
    nodes, weights = gauss_laguerre_rule(rank)
    sa_full = nu_full  # sqrt(A) = sqrt(q tanh q)
    c = 0.5 * (float(np.min(nu_full[support])) + float(np.max(sa_full[s2])) / 2.0)
    taus = nodes / lam
    sa2 = np.sqrt(t2.a)
    plan.o2["eta_rank"] = int(rank)
    plan.o2["eta_damp"] = np.stack([np.exp(-tau * (t1.nu - c)) for tau in taus])
    plan.o2["eta_cosh"] = np.stack([np.cosh(sa2 * tau) for tau in taus])
    plan.o2["eta_sinhr"] = np.stack([_sinh_ratio(sa2, tau) for tau in taus])
    plan.o2["eta_coef"] = np.array([w * math.exp(x) / lam / 4.0 * math.exp(-2.0 * c * (x / lam))
                                    for x, w in zip(nodes, weights)])
    # psi22 radial channels on nu over the support (spatial.py:322-346)
    active = nu_full[support]
    edges = np.linspace(float(np.min(active)), float(np.max(active)), 3)
    masks, shifts = [], []
    for ch in range(2):
        upper = (nu_full < edges[ch + 1]) if ch == 0 else (nu_full <= edges[ch + 1])
        m = support & (nu_full >= edges[ch]) & upper
        if np.any(m):
            masks.append(m)
            vals = nu_full[m]
            shifts.append(0.5 * (float(np.min(vals)) + float(np.max(vals))))
    pair_sup = [[minkowski_sum([masks[i], masks[j]]) for j in range(len(masks))] for i in range(len(masks))]
    pn = math.sqrt(qp * math.tanh(qp))
    pon = math.sqrt(2.0 * qp * math.tanh(2.0 * qp))
    scales = (2.0 * pn - pon, 2.0 * pn + pon)
    pnodes, pweights = psi22_rule()
    damp = np.zeros((2, 2, 2, plan.box1.wy, plan.box1.wx))
    fac = np.zeros((2, 2, 4, plan.box2.wy, plan.box2.wx))
    coef = np.zeros((2, 2))
    nu2 = t2.nu
    for b, lam_b in enumerate(scales):
        if lam_b <= 0 or not np.isfinite(lam_b):
            raise ValueError("invalid dual-branch Green--Laplace scale")
        for n, (x, w) in enumerate(zip(pnodes, pweights)):
            tau = x / lam_b
            coef[b, n] = w * math.exp(x) / lam_b / 8.0
            for ch, (m, cch) in enumerate(zip(masks, shifts)):
                damp[b, n, ch] = plan.box1.take(m).astype(float) * np.exp(-tau * (t1.nu - cch))
            for i in range(len(masks)):
                for j in range(len(masks)):
                    cs = shifts[i] + shifts[j]
                    f = np.exp((nu2 - cs) * tau) if b == 0 else np.exp(-(nu2 + cs) * tau)
                    fac[b, n, 2 * i + j] = f * plan.box2.take(pair_sup[i][j])
    plan.o2.update(psi_damp=damp, psi_factor=fac, psi_coef=coef, psi_sign=np.array([-1.0, 1.0]),
                   channels=2)
    below = bool(np.any(q[support] < 0.3))
    below_mass = float(np.sum(np.abs(spectrum[q < 0.3]) ** 2) / max(np.sum(np.abs(spectrum) ** 2), TINY))
    plan.audits["eta22"] = {
        "implementation": "prescribed-rank-green-laplace-eta22",
        "quadrature_rank": int(rank),
        "quadrature_exact_moment_degree": 2 * int(rank) - 1,
        "peak_depth_wavenumber": qp,
        "fft_ifft_count": 10 * int(rank) + 5,
        "pointwise_product_count": 7 * int(rank) + 1,
        "interaction_enumeration": False,
        "working_grid": list(plan.g2[::-1]),
    }
    plan.audits["psi22"] = {
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


def _order3_tables(plan: HostPlan, support, s2, s3, q, nu_full, inner_rank, outer_rank):
    """Nested GL order-3 tables (_order3.py:42-58, 124-135, 246-249)."""
    qp = plan.peak_q
    nu_min = float(np.min(nu_full[support]))
    c2 = 0.5 * (nu_min + float(np.max(nu_full[s2])) / 2.0)
    c3 = 0.5 * (nu_min + float(np.max(nu_full[s3])) / 3.0)
    pn = math.sqrt(qp * math.tanh(qp))
    l2 = 2.0 * pn - math.sqrt(2.0 * qp * math.tanh(2.0 * qp))
    l3 = 3.0 * pn - math.sqrt(3.0 * qp * math.tanh(3.0 * qp))
    if l2 <= 0 or l3 <= 0:
        raise ValueError("invalid third-order Green--Laplace scales")
    xi, wi = gauss_laguerre_rule(inner_rank)
    xo, wo = gauss_laguerre_rule(outer_rank)
    nu1 = plan.t1.nu
    sa2 = np.sqrt(plan.t2.a)
    sa3 = np.sqrt(plan.t3.a)
    tau_i = xi / l2
    tau_o = xo / l3
    damp_o = np.stack([np.exp(-to * (nu1 - c3)) for to in tau_o])
    damp_i = np.stack([np.exp(-ti * (nu1 - c2)) for ti in tau_i])
    plan.o3 = {
        "inner_rank": int(inner_rank),
        "outer_rank": int(outer_rank),
        "damp_outer": damp_o,
        "damp_inner": (damp_o[:, None] * damp_i[None, :]).reshape(len(xo) * len(xi), *nu1.shape),
        "inner_cosh": np.stack([np.cosh(sa2 * t) for t in tau_i]),
        "inner_sinhr": np.stack([_sinh_ratio(sa2, t) for t in tau_i]),
        "inner_coef": np.array([w * math.exp(x) / l2 * math.exp(-2.0 * c2 * (x / l2)) / 4.0 for x, w in zip(xi, wi)]),
        "outer_cosh": np.stack([np.cosh(sa3 * t) for t in tau_o]),
        "outer_sinhr": np.stack([_sinh_ratio(sa3, t) for t in tau_o]),
        "outer_coef": np.array([w * math.exp(x) / l3 * math.exp(-3.0 * c3 * (x / l3)) for x, w in zip(xo, wo)]),
        "inner_shift": c2,
        "outer_shift": c3,
        "lambda2": l2,
        "lambda3": l3,
    }
    plan.audits["order3"] = {
        "implementation": "nested-no-stokes-green-laplace-order3",
        "inner_quadrature_rank": int(inner_rank),
        "outer_quadrature_rank": int(outer_rank),
        "inner_scale": float(l2),
        "outer_scale": float(l3),
        "inner_balance_shift": c2,
        "outer_balance_shift": c3,
        "interaction_enumeration": False,
        "core_fft_ifft_count": int(outer_rank) * (12 + 16 * int(inner_rank) + 9 + 4) + 2,
        "working_grids": {"quadratic": list(plan.g2[::-1]), "cubic": list(plan.g3[::-1])},
    }


def build_order4_plan(spectrum: np.ndarray, qx: np.ndarray, qy: np.ndarray, *, peak_q: float, rank: int,
                      depth: float = 1.0, gravity: float = 1.0, precision: int = PREC_FP64) -> HostPlan:
    """FP64 plan of the pure-sum GL eta44 graph (one GL rank for every nesting level).

    Follows the setup of paper/order4/cpp/gl_eta44_fftw.cpp:76-155 and
    src/internal/finite_depth_directional_green_laplace_eta44_three_kernels.m:33-56:
    support, pair/triple/quartic supports, shifts c_n = (min nu + max sqrt(A)/n)/2,
    scales lambda_n = n nu_p - sqrt(n q_p tanh(n q_p)).
    """
    if rank not in (4, 6, 8, 10, 12):
        raise ValueError("Unsupported no-Stokes order-four Green--Laplace rank.")
    spectrum = np.asarray(spectrum, dtype=np.complex128)
    ny, nx = spectrum.shape
    q = np.hypot(qx, qy)
    mag = np.abs(spectrum)
    support = mag > 1.0e-13 * max(1.0, float(np.max(mag)))
    if not np.any(support) or np.any(qx[support] <= 0) or np.any(q[support] <= 0):
        raise ValueError("Strict-forward eta11 support with every parent q>0 is required.")
    mx, my = mode_grids(ny, nx)
    if np.max(4 * np.abs(mx[support])) >= nx / 2 or np.max(4 * np.abs(my[support])) >= ny / 2:
        raise ValueError("quartic output support reaches a Nyquist boundary")
    nu_full = np.sqrt(q * np.tanh(q))
    s2, s3, s4 = (sum_support(support, n) for n in (2, 3, 4))
    plan = HostPlan(nx=nx, ny=ny, depth=float(depth), gravity=float(gravity), order=1, precision=int(precision),
                    floor_level=16.0 * EPS * max(1.0, math.log2(nx * ny)), box1=box_of(support), peak_q=float(peak_q))
    b1 = plan.box1
    plan.box2, plan.box3, plan.box4 = box_of(s2), box_of(s3), box_of(s4)
    plan.g2 = (fft_length_at_least(max(2 * b1.wx - 1, 16)), fft_length_at_least(max(2 * b1.wy - 1, 16)))
    plan.g3 = (fft_length_at_least(max(3 * b1.wx - 2, 16)), fft_length_at_least(max(3 * b1.wy - 2, 16)))
    plan.g4 = (fft_length_at_least(max(4 * b1.wx - 3, 16)), fft_length_at_least(max(4 * b1.wy - 3, 16)))
    plan.t1 = _tables(b1, qx, qy)
    plan.t2 = _tables(plan.box2, qx, qy)
    plan.t3 = _tables(plan.box3, qx, qy)
    plan.t4 = _tables(plan.box4, qx, qy)
    plan.omega = math.sqrt(gravity / depth) * plan.t1.nu
    plan.mask2 = plan.box2.take(s2).astype(np.uint8)
    plan.mask3 = plan.box3.take(s3).astype(np.uint8)
    plan.mask4 = plan.box4.take(s4).astype(np.uint8)
    nu_min = float(np.min(nu_full[support]))
    c2 = 0.5 * (nu_min + float(np.max(nu_full[s2])) / 2)
    c3 = 0.5 * (nu_min + float(np.max(nu_full[s3])) / 3)
    c4 = 0.5 * (nu_min + float(np.max(nu_full[s4])) / 4)
    pn = math.sqrt(peak_q * math.tanh(peak_q))
    lam = {n: n * pn - math.sqrt(n * peak_q * math.tanh(n * peak_q)) for n in (2, 3, 4)}
    if not all(np.isfinite(v) and v > 0 for v in lam.values()):
        raise ValueError("Frozen order-four Green--Laplace tail scale is invalid.")
    x, w = gauss_laguerre_rule(rank)
    nu1 = plan.t1.nu
    d4 = np.stack([np.exp(-(xi / lam[4]) * (nu1 - c4)) for xi in x])
    d3 = np.stack([np.exp(-(xi / lam[3]) * (nu1 - c3)) for xi in x])
    d2 = np.stack([np.exp(-(xi / lam[2]) * (nu1 - c2)) for xi in x])
    J = len(x)
    dt = (d4[:, None] * d3[None, :]).reshape(J * J, *nu1.shape)
    sa = {n: np.sqrt(t.a) for n, t in ((2, plan.t2), (3, plan.t3), (4, plan.t4))}

    def factors(n):
        taus = x / lam[n]
        return (np.stack([np.cosh(sa[n] * t) for t in taus]), np.stack([_sinh_ratio(sa[n], t) for t in taus]))

    pc, ps = factors(2)
    tc, ts = factors(3)
    qc, qs = factors(4)
    cs = {2: c2, 3: c3, 4: c4}
    coef = {n: np.array([wi * math.exp(xi) / lam[n] * math.exp(-n * cs[n] * (xi / lam[n])) for xi, wi in zip(x, w)])
            for n in (2, 3, 4)}
    plan.o4 = {
        "rank": J,
        "damp_quartic": d4,
        "damp_third": dt,
        "damp_pair3": (dt[:, None] * d2[None, :]).reshape(J * J * J, *nu1.shape),
        "damp_pair4": (d4[:, None] * d2[None, :]).reshape(J * J, *nu1.shape),
        "pair_cosh": pc, "pair_sinhr": ps, "pair_coef": coef[2] / 4.0,
        "third_cosh": tc, "third_sinhr": ts, "third_coef": coef[3],
        "quartic_cosh": qc, "quartic_sinhr": qs, "quartic_coef": coef[4],
        "shifts": (c2, c3, c4), "scales": (lam[2], lam[3], lam[4]),
    }
    plan.audits["eta44"] = {
        "implementation": "pure-sum-green-laplace-eta44-no-stokes",
        "quadrature_rank": J,
        "outer_scale": lam[4], "outer_balance_shift": c4,
        "reference_fft_ifft_count": J * (J * (16 + J * 22 + 15 + 4) + 8 + 19 + J * 22 + 18 + 2) + 1,
        "working_grids": {"quadratic": list(plan.g2[::-1]), "cubic": list(plan.g3[::-1]),
                          "quartic": list(plan.g4[::-1])},
    }
    return plan
