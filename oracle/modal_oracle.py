"""FP64 NumPy restatement of the reference single-direction modal time path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Restates
python/src/green_laplace/modal_time.py:26-399 (eta11/eta22/eta33 on a
one-dimensional dimensionless-wavenumber modal grid, with the modal-points
ladder) and python/src/green_laplace/_eta20.py:13-186 (the difference-frequency
eta20 on a two-dimensional modal grid).  Time dependence is exp(-i omega t):
the record's analysis is ``fft(eta1)/N`` and the synthesis is ``fft`` of the
positive-bin coefficients.

Arrays of the modal solve are ``[P, W]``: P modal wavenumbers (axis 0, FFT
order after the shift) by W work frequencies (axis 1, bin ``base + w``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .gl_oracle import gauss_laguerre_rule

TINY = np.finfo(float).tiny
EPS = np.finfo(float).eps


def wavenumber_from_frequency(omega: np.ndarray, depth: float, gravity: float) -> np.ndarray:
    """Solve q tanh q = omega^2 h / g for q = k h >= 0 by safeguarded Newton.

    Same equation and acceptance test as python/src/green_laplace/dispersion.py:19-45
    (relative residual 2e-13); the iteration itself only has to converge.
    """
    target = np.asarray(omega, dtype=float) ** 2 * depth / gravity
    q = np.maximum(np.where(target < 1.0, np.sqrt(target), target), TINY)
    for _ in range(60):
        t = np.tanh(q)
        step = (q * t - target) / np.maximum(t + q * (1.0 - t * t), TINY)
        nxt = np.maximum(q - step, 0.5 * q)
        done = np.max(np.abs(nxt - q) / np.maximum(1.0, q)) < 5.0e-15
        q = nxt
        if done:
            break
    if not np.allclose(q * np.tanh(q), target, rtol=2.0e-13, atol=2.0e-15):
        raise ArithmeticError("dispersion solve did not converge")
    return q


@dataclass
class Band:
    """Selected positive-frequency band of a uniformly sampled record.

    modal_time.py:66-96: energy-fraction cut over bins 1..(N-1)//2, capped at
    ((N-1)//2)//order and at omega_max; amplitudes 2 conj(c_b), c = fft/N.
    """

    count: int
    dt: float
    bins: np.ndarray
    omega: np.ndarray
    amplitudes: np.ndarray
    q: np.ndarray
    peak_q: float
    eta11: np.ndarray
    retained: float


def select_band(values, times, *, depth, gravity, order, omega_max, energy_fraction) -> Band:
    values = np.asarray(values, dtype=float).ravel()
    times = np.asarray(times, dtype=float).ravel()
    n = values.size
    if n != times.size or n < 9:
        raise ValueError("eta1 and time must have the same length of at least nine")
    dt = float(np.mean(np.diff(times)))
    drift = float(np.max(np.abs(times - (times[0] + dt * np.arange(n)))))
    if dt <= 0 or drift > max(1e-10, 1024 * EPS * max(1.0, times[-1] - times[0])):
        raise ValueError("a uniformly increasing time record is required")
    half = (n - 1) // 2
    c = np.fft.fft(values) / n
    e = np.abs(c[1:half + 1]) ** 2
    total = float(e.sum())
    if total == 0:
        raise ValueError("eta1 has no positive-frequency content")
    top = 1 + min(int(np.searchsorted(np.cumsum(e) / total, energy_fraction)), half - 1)
    top = min(top, half // order)
    if omega_max is not None:
        top = min(top, int(np.floor(omega_max * n * dt / (2 * np.pi))))
    if top < 1:
        raise ValueError("the selected eta1 frequency band is empty")
    bins = np.arange(1, top + 1)
    omega = 2 * np.pi * bins / (n * dt)
    amp = 2 * np.conj(c[bins])
    spec = np.zeros(n, dtype=complex)
    spec[bins] = amp
    q = wavenumber_from_frequency(omega, depth, gravity)
    return Band(n, dt, bins, omega, amp, q, float(q[int(np.argmax(np.abs(amp)))]), np.real(np.fft.fft(spec)),
                float(e[:top].sum() / total))


# --------------------------------------------------------------------------
# modal grid (modal_time.py:344-381)
# --------------------------------------------------------------------------
def modal_axis(points: int, qmax: float, order: int, limit: float | None):
    """Symmetric axis [-P/2, P/2) dq reaching (order + 0.05) qmax unless a limit is given."""
    if limit is None:
        lim = (order + 0.05) * qmax / (1 - 2 / points)
    else:
        lim = float(limit)
        if lim <= order * qmax:
            raise ValueError("modal_limit must exceed the requested nonlinear output support")
    dq = 2 * lim / points
    return dq * np.arange(-points // 2, points // 2), dq


def bracket(value, axis, dq):
    """Linear-interpolation neighbours (lower index, upper index, weights)."""
    pos = (value - axis[0]) / dq
    lo = np.clip(np.floor(pos).astype(int), 0, axis.size - 2)
    wu = np.clip(pos - lo, 0.0, 1.0)
    return lo, lo + 1, 1.0 - wu, wu


def deposit_1d(amp, work, q, axis, dq, width):
    """Joint [P, width] spectrum (axis order) and the mass-weighted projection RMS."""
    lo, hi, wl, wu = bracket(q, axis, dq)
    joint = np.zeros((axis.size, width), dtype=complex)
    np.add.at(joint, (lo, work), amp * wl)
    np.add.at(joint, (hi, work), amp * wu)
    mass = np.abs(amp) ** 2
    err2 = np.concatenate(((axis[lo] - q) / q, (axis[hi] - q) / q)) ** 2
    wts = np.concatenate((mass * wl, mass * wu))
    return joint, float(np.sqrt(np.sum(wts * err2) / max(np.sum(wts), TINY)))


def to_field(spec, points):
    """(q, w) -> (x, t): inverse FFT over q scaled by P, forward FFT over w."""
    return np.fft.fft(np.fft.ifft(spec, axis=0) * points, axis=1)


def to_spectrum(fld, points):
    """(x, t) -> (q, w): the inverse of ``to_field``."""
    return np.fft.fft(np.fft.ifft(fld, axis=1), axis=0) / points


def node_sums(s, a, scale, nodes, weights):
    """GL sums S(s, a), C(s, a) on the support s > a (modal_time.py:383-399).

    With tau = x/scale and E(+-) = (w/scale) exp(x - (s -+ a) tau):
    S = sum (E+ - E-)/(2a), C = sum (E+ + E-)/2 for a > 0 and
    S = sum (w/scale) e^{x - s tau} tau, C = sum (w/scale) e^{x - s tau} at a = 0.
    ``s`` and ``a`` broadcast; zero where s <= a.
    """
    s, a = np.broadcast_arrays(np.asarray(s, float), np.asarray(a, float))
    ok = s > a
    pos = ok & (a > 0)
    zero = ok & (a == 0)
    S = np.zeros(s.shape)
    C = np.zeros(s.shape)
    ap = np.where(pos, a, 1.0)
    with np.errstate(over="ignore", invalid="ignore"):
        for x, w in zip(nodes, weights):
            tau = x / scale
            ep = w / scale * np.exp(x - (s - a) * tau)
            em = w / scale * np.exp(x - (s + a) * tau)
            e0 = w / scale * np.exp(x - s * tau)
            S += np.where(pos, (ep - em) / (2 * ap), np.where(zero, e0 * tau, 0.0))
            C += np.where(pos, (ep + em) / 2, np.where(zero, e0, 0.0))
    return S, C, ok


@dataclass
class ModalLevel:
    eta22: np.ndarray
    eta33: np.ndarray | None
    audit: dict[str, Any] = field(default_factory=dict)


def modal_level(band: Band, *, depth, gravity, rank, order, points, limit) -> ModalLevel:
    """One level of the modal ladder (modal_time.py:187-341)."""
    n = band.count
    base = int(band.bins.min())
    work = band.bins - base
    span = int(work.max())
    width = order * span + 1
    if width > n:
        raise ValueError("the nonlinear output band does not fit in the time record")
    axis, dq = modal_axis(points, float(np.max(np.abs(band.q))), order, limit)
    joint_ax, proj = deposit_1d(band.amplitudes, work, band.q, axis, dq, width)
    J = np.fft.ifftshift(joint_ax, axes=0)
    qs = np.fft.ifftshift(axis)[:, None]          # signed modal wavenumber, [P, 1]
    qa = np.abs(qs[:, 0])
    a = np.sqrt(qa * np.tanh(qa))                  # response frequency sqrt(|q| tanh|q|)
    aq = qa * np.tanh(qa)
    fstep = 2 * np.pi / (n * band.dt) * np.sqrt(depth / gravity)
    nu = ((base + np.arange(width)) * fstep)[None, :]
    nu_s = np.maximum(nu, TINY)
    F = lambda m: to_field(J * m, points)  # noqa: E731

    v, vn, vn2 = F(1.0), F(nu), F(nu**2)
    hx, jx, rad = F(qs / nu_s), F(qs), F(qs**2 / nu_s)
    sd_hat = to_spectrum(2 * v * vn2 + vn**2 - hx**2, points)
    sk_hat = to_spectrum(2 * v * rad + 2 * hx * jx, points)

    nodes, weights = gauss_laguerre_rule(rank)
    pq = band.peak_q
    nu_p = np.sqrt(pq * np.tanh(pq))

    def bin_sum(n_ord, lam, dh, kh, coef_d, coef_k, denom):
        """Coefficients at bins n_ord*base + w (w <= n_ord*span) and the invalid-energy fraction."""
        w_out = np.arange(n_ord * span + 1)
        s = ((n_ord * base + w_out) * fstep)[None, :]
        S, C, ok = node_sums(s, a[:, None], lam, nodes, weights)
        d, k = dh[:, w_out], kh[:, w_out]
        resp = np.where(ok, coef_d(S, C) * d + coef_k(S, C) * k, 0) / denom
        energy = np.abs(d) ** 2 + np.abs(k) ** 2
        coeff = np.zeros(n, dtype=complex)
        coeff[n_ord * base + w_out] = resp.sum(axis=0)
        return coeff, float(energy[~ok].sum()) / max(float(energy.sum()), TINY), S, C, ok

    lam2 = np.sqrt(4 * pq * np.tanh(pq) - 2 * pq * np.tanh(2 * pq))
    c2, inv2, *_ = bin_sum(2, lam2, sd_hat, sk_hat, lambda S, C: -aq[:, None] * S, lambda S, C: C, 4 * depth)
    eta22 = np.fft.fft(c2)

    eta33 = None
    inv3 = None
    if order == 3:
        lam_in = 2 * nu_p - np.sqrt(2 * pq * np.tanh(2 * pq))
        w2 = np.arange(2 * span + 1)
        s2 = ((2 * base + w2) * fstep)[None, :]
        S2, C2, ok2 = node_sums(s2, a[:, None], lam_in, nodes, weights)
        d2, k2 = sd_hat[:, w2], sk_hat[:, w2]
        e2 = np.zeros((points, width), dtype=complex)
        p2 = np.zeros_like(e2)
        e2[:, w2] = np.where(ok2, -aq[:, None] * S2 * d2 + C2 * k2, 0) / 4
        p2[:, w2] = np.where(ok2, 1j * (C2 * d2 - S2 * k2), 0) / 4
        nu2 = ((2 * base + np.arange(width)) * fstep)[None, :]
        A = aq[:, None]
        f_eta, f_etax, f_phix = v, F(1j * qs), hx
        f_phiz, f_phixz = F(-1j * nu), F(nu * qs)
        f_phizz, f_phizzz = F(-1j * qs**2 / nu_s), F(-1j * qs**2 * nu)
        f_phitz, f_phitzz = F(-nu**2), F(-qs**2)
        G = lambda m: to_field(m, points)  # noqa: E731
        g_eta, g_etax = G(e2), G(e2 * 1j * qs)
        g_phix, g_phiz = G(p2 * 1j * qs), G(p2 * A)
        g_phizz, g_phitz = G(p2 * qs**2), G(p2 * (-1j * nu2) * A)
        fk = ((f_eta * f_phixz * f_etax - 0.5 * f_eta**2 * f_phizzz) / 4
              + (f_phix * g_etax + g_phix * f_etax - f_eta * g_phizz - g_eta * f_phizz) / 2)
        fd = ((0.5 * f_eta**2 * f_phitzz + f_eta * (f_phix * f_phixz + f_phiz * f_phizz)) / 4
              + (f_eta * g_phitz + g_eta * f_phitz + f_phix * g_phix + f_phiz * g_phiz) / 2)
        fk_hat, fd_hat = to_spectrum(fk, points), to_spectrum(fd, points)
        lam3 = 3 * nu_p - np.sqrt(3 * pq * np.tanh(3 * pq))
        c3, inv3, *_ = bin_sum(3, lam3, fd_hat, fk_hat, lambda S, C: aq[:, None] * S,
                               lambda S, C: -1j * C, depth**2)
        eta33 = np.fft.fft(c3)

    audit = {
        "modal_points": points,
        "modal_limit": float(abs(axis[0])),
        "quadrature_rank": rank,
        "base_bin": base,
        "work_frequency_count": width,
        "wavevector_projection_rms_relative": proj,
        "eta22_invalid_source_energy_fraction": inv2,
        "eta33_invalid_source_energy_fraction": inv3,
    }
    return ModalLevel(eta22, eta33, audit)


# --------------------------------------------------------------------------
# eta20 (python/src/green_laplace/_eta20.py:13-186)
# --------------------------------------------------------------------------
def hermitian_project(c):
    """Average c[b] with conj(c[-b]); real DC and (even N) Nyquist (_eta20.py:175-185)."""
    out = np.array(c, dtype=complex)
    n = out.size
    b = np.arange(1, (n - 1) // 2 + 1)
    avg = (out[b] + np.conj(out[n - b])) / 2
    out[b] = avg
    out[n - b] = np.conj(avg)
    out[0] = out[0].real
    if n % 2 == 0:
        out[n // 2] = out[n // 2].real
    return out


def eta20(amp, omega, q, times, *, depth, gravity, rank=16, nx=128, ny=16):
    """Difference-frequency eta20 on the [ny, nx] modal grid; returns (eta20, audit)."""
    if rank not in (6, 12, 16):
        raise ValueError("eta20 rank must be 6, 12, or 16")
    if nx < 16 or ny < 16 or nx % 2 or ny % 2:
        raise ValueError("eta20 modal-grid sizes must be even and at least 16")
    n = times.size
    dt = float(np.mean(np.diff(times)))
    bins = np.rint(omega * n * dt / (2 * np.pi)).astype(int)
    if np.any(bins < 1) or not np.allclose(omega, 2 * np.pi * bins / (n * dt), rtol=1e-10, atol=1e-12):
        raise ValueError("eta20 input frequencies must lie on temporal DFT bins")
    base = int(bins.min())
    work = bins - base
    span = int(work.max())
    width = 2 * span + 1
    qmax = float(np.max(q))
    xax, dqx = modal_axis(nx, float(np.max(np.abs(q))), 2, None)
    yax, dqy = modal_axis(ny, 0.0, 2, 2.05 * qmax / (1 - 2 / ny))
    ix0, ix1, wx0, wx1 = bracket(q, xax, dqx)
    iy0, iy1, wy0, wy1 = bracket(np.zeros_like(q), yax, dqy)
    joint = np.zeros((ny, nx, width), dtype=complex)
    for ix, wx in ((ix0, wx0), (ix1, wx1)):
        for iy, wy in ((iy0, wy0), (iy1, wy1)):
            np.add.at(joint, (iy, ix, work), amp * wx * wy)
    J = np.fft.ifftshift(joint, axes=(0, 1))
    QX, QY = np.meshgrid(np.fft.ifftshift(xax), np.fft.ifftshift(yax))
    Q = np.hypot(QX, QY)
    a = np.sqrt(Q * np.tanh(Q))
    qx3, qy3 = QX[:, :, None], QY[:, :, None]
    fstep = 2 * np.pi / (n * dt) * np.sqrt(depth / gravity)
    nu = ((base + np.arange(width)) * fstep)[None, None, :]
    nu_s = np.maximum(nu, TINY)

    def fld(m):
        return np.fft.fft(np.fft.ifft2(J * m, axes=(0, 1)) * (nx * ny), axis=2)

    one, dno, nuf = fld(1.0), fld(nu**2), fld(nu)
    bx, by = fld(qx3 / nu_s), fld(qy3 / nu_s)
    qxf, qyf = fld(qx3), fld(qy3)
    rad = fld((qx3**2 + qy3**2) / nu_s)
    cj = np.conj
    fd = -dno * cj(one) - one * cj(dno) + nuf * cj(nuf) + bx * cj(bx) + by * cj(by)
    fk = 1j * (rad * cj(one) - bx * cj(qxf) - by * cj(qyf) + qxf * cj(bx) + qyf * cj(by) - one * cj(rad))
    spec = lambda f: np.fft.fft2(np.fft.ifft(f, axis=2), axes=(0, 1)) / (nx * ny)  # noqa: E731
    fd_hat, fk_hat = spec(fd), spec(fk)

    e = np.abs(amp) ** 2
    mean_q = float(np.sum(q * e) / e.sum())
    dq_rms = float(np.sqrt(2 * np.sum((q - mean_q) ** 2 * e) / e.sum()))
    audit = {"quadrature_rank": rank, "modal_grid": [ny, nx], "delta_q_rms_pair": dq_rms}
    if dq_rms <= 64 * EPS * max(1.0, qmax):
        return np.zeros(n), dict(audit, status="degenerate-zero-difference-output")
    nodes, weights = gauss_laguerre_rule(rank)
    coeff = np.zeros(n, dtype=complex)
    for idx in range(width):
        sb = idx if idx <= span else idx - width          # signed difference bin
        sigma = sb * fstep
        ok = (Q > 0) & (a > abs(sigma))
        md = np.zeros(Q.shape)
        mk = np.zeros(Q.shape, dtype=complex)
        with np.errstate(over="ignore", invalid="ignore"):
            for x, w in zip(nodes, weights):
                tau = x / dq_rms
                ep = w / dq_rms * np.exp(x - (a - sigma) * tau)
                em = w / dq_rms * np.exp(x - (a + sigma) * tau)
                md -= np.where(ok, a * (ep + em) / 4, 0)
                mk += np.where(ok, 1j * (ep - em) / 4, 0)
        r = np.where(ok, md * fd_hat[:, :, idx] + mk * fk_hat[:, :, idx], 0) / (2 * depth)
        coeff[sb % n] = r.sum()
    return np.real(np.fft.fft(hermitian_project(coeff))), audit


# --------------------------------------------------------------------------
# public entry (modal_time.py:26-184)
# --------------------------------------------------------------------------
@dataclass
class ModalResult:
    time: np.ndarray
    eta: np.ndarray
    components: dict
    audit: dict


def reconstruct_unidirectional_modal_timeseries(eta1, time, *, depth, gravity=9.81, omega_max=None,
                                                energy_fraction=0.999, quadrature_rank=8, order=3,
                                                modal_points=(2048, 4096, 8192, 16384, 32768, 65536),
                                                modal_limit=None, relative_tolerance=1.0e-3,
                                                include_eta20=True, eta20_rank=16, eta20_grid=(128, 16)):
    if depth <= 0 or gravity <= 0 or order not in (2, 3):
        raise ValueError("depth and gravity must be positive and order must be 2 or 3")
    ladder = (int(modal_points),) if np.isscalar(modal_points) else tuple(int(p) for p in modal_points)
    if not ladder or any(p < 16 or p % 2 for p in ladder) or any(b <= a for a, b in zip(ladder, ladder[1:])):
        raise ValueError("modal_points must be an increasing sequence of even integers")
    if relative_tolerance <= 0:
        raise ValueError("relative_tolerance must be positive")
    if order == 3 and quadrature_rank not in (4, 6, 8, 10):
        raise ValueError("third-order modal reconstruction supports rank 4, 6, 8, or 10")
    times = np.asarray(time, dtype=float).ravel()
    band = select_band(eta1, times, depth=depth, gravity=gravity, order=order, omega_max=omega_max,
                       energy_fraction=energy_fraction)
    prev = None
    levels = []
    converged = False
    lev = None
    for p in ladder:
        lev = modal_level(band, depth=depth, gravity=gravity, rank=quadrature_rank, order=order, points=p,
                          limit=modal_limit)
        cur = (lev.eta22,) if lev.eta33 is None else (lev.eta22, lev.eta33)
        changes = None
        if prev is not None:
            changes = [float(np.linalg.norm(c - o) / np.linalg.norm(c)) for c, o in zip(cur, prev)]
            converged = max(changes) <= relative_tolerance
        levels.append({"modal_points": p,
                       "eta22_change": None if changes is None else changes[0],
                       "eta33_change": None if changes is None or lev.eta33 is None else changes[1],
                       "wavevector_projection_rms_relative": lev.audit["wavevector_projection_rms_relative"]})
        prev = cur
        if converged:
            break
    comps = {"eta11": band.eta11, "eta22": lev.eta22.real, "eta22_analytic": lev.eta22}
    total = band.eta11 + lev.eta22.real
    if lev.eta33 is not None:
        comps["eta33"] = lev.eta33.real
        comps["eta33_analytic"] = lev.eta33
        total = total + lev.eta33.real
    audit = dict(lev.audit)
    if include_eta20:
        e20, a20 = eta20(band.amplitudes, band.omega, band.q, times, depth=depth, gravity=gravity,
                         rank=eta20_rank, nx=eta20_grid[0], ny=eta20_grid[1])
        comps["eta20"] = e20
        total = total + e20
        audit["eta20"] = a20
    vals = np.asarray(eta1, dtype=float).ravel()
    audit.update({"retained_positive_frequency_energy": band.retained,
                  "input_projection_relative": float(np.linalg.norm(band.eta11 - vals)
                                                     / max(np.linalg.norm(vals), TINY)),
                  "levels": levels, "converged": converged, "relative_tolerance": relative_tolerance,
                  "auxiliary_space_domain": False})
    return ModalResult(times.copy(), total, comps, audit)
