"""Plain-torch restatement of eta11 + eta22 + eta33 on the full grid, for autograd (test infrastructure).

Follows the FP64 oracle (oracle/gl_oracle.py: eta22, inner_pair_fields,
first_order_state, order3_forcing, order3) step for step with torch.fft, so
that torch autograd gives a reference gradient for the hand-written adjoint.
Supports, balance shifts, GL nodes and the peak wavenumber are computed once
(NumPy, from the plan's spectrum) and held fixed; every roundoff floor uses
the current forward magnitudes (detached), as in the adjoint.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from oracle import gl_oracle as O

EPS = float(np.finfo(float).eps)


class TorchGL:
    def __init__(self, spec0: np.ndarray, qx: np.ndarray, qy: np.ndarray, *, depth: float, order: int,
                 peak_q: float, eta22_rank: int = 6, inner_rank: int = 4, outer_rank: int = 4, device="cpu"):
        self.device = device
        t = lambda x: torch.as_tensor(x, device=device)  # noqa: E731
        g = O.Geometry(qx, qy)
        self.depth, self.order = depth, order
        self.ny, self.nx = spec0.shape
        self.level = 16.0 * EPS * max(1.0, math.log2(spec0.size))
        sup = O.parent_support(spec0)
        self.s2 = t(O.sum_support(sup, 2))
        self.s3 = t(O.sum_support(sup, 3))
        self.qx, self.qy, self.q = t(g.qx), t(g.qy), t(g.q)
        self.nu, self.sa, self.snu, self.a = t(g.nu), t(g.sa), t(g.snu), t(g.a)
        self.sa_np = g.sa
        # eta22 scalars (oracle eta22)
        disc = 4.0 * peak_q * math.tanh(peak_q) - 2.0 * peak_q * math.tanh(2.0 * peak_q)
        self.lam = math.sqrt(disc)
        self.c22 = 0.5 * (float(np.min(g.nu[sup])) + float(np.max(g.sa[O.sum_support(sup, 2)])) / 2.0)
        self.rule22 = O.gauss_laguerre_rule(eta22_rank)
        if order >= 3:
            self.sc3 = O.order3_scalars(spec0, g, peak_q)
            self.rule_in = O.gauss_laguerre_rule(inner_rank)
            self.rule_out = O.gauss_laguerre_rule(outer_rank)
            m = O.order3_multipliers(g)
            self.m = {k: t(v) for k, v in m.items()}

    def floor(self, x, sup):
        x = torch.where(sup, x, torch.zeros_like(x))
        top = torch.max(torch.abs(x.detach()))
        if float(top) == 0.0:
            return x
        keep = ~(sup & (torch.abs(x.detach()) < self.level * top))
        return torch.where(keep, x, torch.zeros_like(x))

    def ratio(self, tau):
        sa = torch.as_tensor(self.sa_np, device=self.device)
        out = torch.full_like(sa, tau)
        nz = sa > 0
        out[nz] = torch.sinh(sa[nz] * tau) / sa[nz]
        return out

    def eta22(self, spec):
        ifft2, fft2 = torch.fft.ifft2, torch.fft.fft2
        acc = torch.zeros_like(spec)
        x, w = self.rule22
        for xn, wn in zip(x, w):
            tau = xn / self.lam
            d = spec * torch.exp(-tau * (self.nu - self.c22))
            v, vn, vn2 = ifft2(d), ifft2(d * self.nu), ifft2(d * self.nu**2)
            hx, hy = ifft2(d * self.qx / self.snu), ifft2(d * self.qy / self.snu)
            rad = ifft2(d * self.q**2 / self.snu)
            jx, jy = ifft2(d * self.qx), ifft2(d * self.qy)
            sd = self.floor(fft2(2.0 * v * vn2 + vn**2 - hx**2 - hy**2), self.s2)
            sk = self.floor(fft2(2.0 * v * rad + 2.0 * (hx * jx + hy * jy)), self.s2)
            coef = wn * math.exp(xn) / self.lam / 4.0 * math.exp(-2.0 * self.c22 * tau)
            acc = acc + coef * (-self.a * self.ratio(tau) * sd + torch.cosh(self.sa * tau) * sk)
        return torch.fft.ifft2(acc)

    def eta33(self, spec):
        ifft2, fft2 = torch.fft.ifft2, torch.fft.fft2
        sc, m = self.sc3, self.m
        s2, s3 = self.s2, self.s3
        eta_hat = torch.zeros_like(spec)
        xo, wo = self.rule_out
        xi, wi = self.rule_in
        for xn, wn in zip(xo, wo):
            tau = xn / sc.lambda3
            d = spec * torch.exp(-tau * (self.nu - sc.outer_shift))
            f = {"eta": ifft2(d), "etax": ifft2(m["dx"] * d), "etay": ifft2(m["dy"] * d),
                 "phix": ifft2(m["qx_nu"] * d), "phiy": ifft2(m["qy_nu"] * d), "phiz": ifft2(m["m_i_nu"] * d),
                 "phixz": ifft2(m["nu_qx"] * d), "phiyz": ifft2(m["nu_qy"] * d),
                 "phizz": ifft2(m["m_i_q2_nu"] * d), "phizzz": ifft2(m["m_i_q2nu"] * d),
                 "phitz": ifft2(m["m_a"] * d), "phitzz": ifft2(m["m_q2"] * d)}
            e_hat = torch.zeros_like(spec)
            et_hat = torch.zeros_like(spec)
            p_hat = torch.zeros_like(spec)
            pt_hat = torch.zeros_like(spec)
            for xm, wm in zip(xi, wi):
                t2 = xm / sc.lambda2
                dd = d * torch.exp(-t2 * (m["nu"] - sc.inner_shift))
                v, vn, vn2, vn3 = ifft2(dd), ifft2(dd * m["nu"]), ifft2(dd * m["nu2"]), ifft2(dd * m["nu3"])
                hx, hy = ifft2(dd * m["qx_nu"]), ifft2(dd * m["qy_nu"])
                jx, jy = ifft2(dd * m["qx"]), ifft2(dd * m["qy"])
                rad, q2f = ifft2(dd * m["q2_nu"]), ifft2(dd * m["q2"])
                njx, njy = ifft2(dd * m["nu_qx"]), ifft2(dd * m["nu_qy"])
                srcs = [2.0 * v * vn2 + vn**2 - hx**2 - hy**2,
                        2.0 * v * rad + 2.0 * (hx * jx + hy * jy),
                        2.0 * v * vn3 + 4.0 * vn * vn2 - 2.0 * hx * jx - 2.0 * hy * jy,
                        2.0 * (vn * rad + v * q2f) + 2.0 * (jx**2 + hx * njx + jy**2 + hy * njy)]
                sd, sk, sdt, skt = (self.floor(fft2(z), s2) for z in srcs)
                sh, ch = self.ratio(t2), torch.cosh(self.sa * t2)
                coef = wm * math.exp(xm) / sc.lambda2 * math.exp(-2.0 * sc.inner_shift * t2) / 4.0
                e_hat = e_hat + coef * (-self.a * sh * sd + ch * sk)
                et_hat = et_hat - 1j * coef * (-self.a * sh * sdt + ch * skt)
                p_hat = p_hat + 1j * coef * (ch * sd - sh * sk)
                pt_hat = pt_hat + coef * (ch * sdt - sh * skt)
            zero = torch.zeros_like(spec)
            e_hat, et_hat, p_hat, pt_hat = (torch.where(s2, z, zero) for z in (e_hat, et_hat, p_hat, pt_hat))
            et_hat = self.floor(et_hat, s2)
            p = {"eta": ifft2(e_hat), "etax": ifft2(m["dx"] * e_hat), "etay": ifft2(m["dy"] * e_hat),
                 "phix": ifft2(m["dx"] * p_hat), "phiy": ifft2(m["dy"] * p_hat),
                 "phiz": ifft2(self.a * p_hat), "phizz": ifft2(m["q2"] * p_hat),
                 "phitz": ifft2(self.a * pt_hat), "eta_t": ifft2(et_hat)}
            k_pair = (f["phix"] * p["etax"] + f["phiy"] * p["etay"] + p["phix"] * f["etax"] + p["phiy"] * f["etay"]
                      - f["eta"] * p["phizz"] - p["eta"] * f["phizz"])
            k_direct = f["eta"] * (f["phixz"] * f["etax"] + f["phiyz"] * f["etay"]) - 0.5 * f["eta"] ** 2 * f["phizzz"]
            d_pair = (f["eta"] * p["phitz"] + p["eta"] * f["phitz"] + f["phix"] * p["phix"] + f["phiy"] * p["phiy"]
                      + f["phiz"] * p["phiz"])
            d_direct = 0.5 * f["eta"] ** 2 * f["phitzz"] + f["eta"] * (
                f["phix"] * f["phixz"] + f["phiy"] * f["phiyz"] + f["phiz"] * f["phizz"])
            fk_hat = self.floor(fft2(k_direct / 4.0 + k_pair / 2.0), s3)
            fd_hat = self.floor(fft2(d_direct / 4.0 + d_pair / 2.0), s3)
            sh, ch = self.ratio(tau), torch.cosh(self.sa * tau)
            coef = wn * math.exp(xn) / sc.lambda3 * math.exp(-3.0 * sc.outer_shift * tau)
            eta_hat = eta_hat + coef * (self.a * sh * fd_hat - 1j * ch * fk_hat)
        eta_hat = torch.where(s3, eta_hat, torch.zeros_like(eta_hat))
        return torch.fft.ifft2(eta_hat)

    def elevation(self, spec, orders=(1, 2, 3)):
        """Re(h (eta11+ + eta22+ + eta33+)) for a full-grid analytic spectrum (complex128 tensor)."""
        e = torch.zeros_like(spec)
        if 1 in orders:
            e = e + torch.fft.ifft2(spec)
        if 2 in orders:
            e = e + self.eta22(spec)
        if 3 in orders:
            e = e + self.eta33(spec)
        return self.depth * e.real


def deposit(a, b, mx, my, nx, ny, depth, device="cpu"):
    """Full-grid t=0 spectrum N^2 (a + i b) / h at bins (my, mx) (src/gl_spectral_surface.m:145-147)."""
    s = torch.zeros(ny * nx, dtype=torch.complex128, device=device)
    idx = torch.as_tensor((np.asarray(my) % ny) * nx + (np.asarray(mx) % nx), device=device)
    s = s.index_add(0, idx, torch.complex(a.to(torch.float64), b.to(torch.float64)) * (nx * ny / depth))
    return s.view(ny, nx)
