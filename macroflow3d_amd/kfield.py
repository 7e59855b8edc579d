"""Log-normal conductivity fields by FFT circulant embedding (synthetic input).

Y = ln K is a zero-mean stationary Gaussian field with exponential covariance
C(r) = sigma2 * exp(-r / corr_len), r the Euclidean distance in cell units.
Method (Dietrich & Newsam 1997): embed the covariance in a periodic grid of
m = next power of two >= 2n per axis (doubled while the circulant has negative
eigenvalues, at most three times, then clipped), take eigenvalues
lambda = FFT(c), draw complex standard normals xi, and return
Re FFT(sqrt(lambda / M) * xi) on the grid.

Draws are counter-based (splitmix64 finaliser, Box-Muller), so a field is a
pure function of (shape, sigma2, corr_len, seed) on any device. The per-
realisation seed rule base + r follows EnsembleRunner.cu:268.

This is input synthesis, not the solve: the reference generates K with a
randomized spectral method instead (src/physics/stochastic/stochastic.cu:56-197).
The FFTs here are torch.fft's; the draws are plain tensor arithmetic.
"""
from __future__ import annotations

import math

import torch

_MASK = (1 << 64) - 1
_GOLDEN = 0x9E3779B97F4A7C15
_C1 = 0xBF58476D1CE4E5B9
_C2 = 0x94D049BB133111EB


def _s64(u: int) -> int:
    u &= _MASK
    return u - (1 << 64) if u >= (1 << 63) else u


def _mix64_int(z: int) -> int:
    z &= _MASK
    z = ((z ^ (z >> 30)) * _C1) & _MASK
    z = ((z ^ (z >> 27)) * _C2) & _MASK
    return z ^ (z >> 31)


def _srl(t: torch.Tensor, s: int) -> torch.Tensor:
    """Logical right shift of int64 bit patterns."""
    return (t >> s) & ((1 << (64 - s)) - 1)


def _mix64(z: torch.Tensor) -> torch.Tensor:
    z = (z ^ _srl(z, 30)) * _s64(_C1)
    z = (z ^ _srl(z, 27)) * _s64(_C2)
    return z ^ _srl(z, 31)


def gauss_pairs(seed: int, k: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Standard normal pairs for counters k (int64 tensor)."""
    key = _s64(_mix64_int(seed + _GOLDEN))
    a = _mix64(key + (2 * k + 1) * _s64(_GOLDEN))
    b = _mix64(key + (2 * k + 2) * _s64(_GOLDEN))
    u1 = (_srl(a, 11).to(torch.float64) + 0.5) * 2.0 ** -53
    u2 = _srl(b, 11).to(torch.float64) * 2.0 ** -53
    rad = torch.sqrt(-2.0 * torch.log(u1))
    th = 6.283185307179586 * u2
    return rad * torch.cos(th), rad * torch.sin(th)


def _next_pow2(v: int) -> int:
    p = 1
    while p < v:
        p <<= 1
    return p


def _axis_dist(m: int, device) -> torch.Tensor:
    a = torch.arange(m, device=device, dtype=torch.float64)
    return torch.minimum(a, m - a)


def lognormal_lnK(nx: int, ny: int, nz: int, sigma2: float, corr_len: float, seed: int,
                  device="cpu", chunk: int = 1 << 24) -> tuple[torch.Tensor, dict]:
    """ln K on the grid as a flat float64 tensor (x fastest) plus embedding stats."""
    device = torch.device(device)
    mx, my, mz = _next_pow2(2 * nx), _next_pow2(2 * ny), _next_pow2(2 * nz)
    for attempt in range(4):
        dx = _axis_dist(mx, device)
        dy = _axis_dist(my, device)
        dz = _axis_dist(mz, device)
        r = torch.sqrt(dz.view(-1, 1, 1) ** 2 + dy.view(1, -1, 1) ** 2 + dx.view(1, 1, -1) ** 2)
        c = sigma2 * torch.exp(-r / corr_len)
        del r
        lam = torch.fft.fftn(c.to(torch.complex128)).real
        del c
        mn, mxv = float(lam.min()), float(lam.max())
        neg = float((-lam.clamp(max=0)).sum())
        tot = float(lam.abs().sum())
        if mn >= -1e-12 * mxv or attempt == 3:
            break
        mx, my, mz = 2 * mx, 2 * my, 2 * mz
    M = mx * my * mz
    scale = torch.sqrt(lam.clamp(min=0.0) / M).reshape(-1)
    del lam
    W = torch.empty(M, dtype=torch.complex128, device=device)
    for s in range(0, M, chunk):
        e = min(M, s + chunk)
        k = torch.arange(s, e, device=device, dtype=torch.int64)
        n1, n2 = gauss_pairs(seed, k)
        W[s:e] = torch.complex(scale[s:e] * n1, scale[s:e] * n2)
        del k, n1, n2
    del scale
    Z = torch.fft.fftn(W.view(mz, my, mx))
    del W
    Y = Z.real[:nz, :ny, :nx].contiguous().reshape(-1)
    stats = {"embedding": (mx, my, mz), "min_eig_ratio": mn / mxv, "clipped_mass": neg / tot,
             "sigma2": sigma2, "corr_len": corr_len, "seed": int(seed)}
    return Y, stats


def lognormal_K(nx: int, ny: int, nz: int, sigma2: float, corr_len: float, seed: int,
                device="cpu", K_mean: float = 1.0) -> tuple[torch.Tensor, dict]:
    """K = K_mean * exp(Y) (K_mean is the geometric mean, physics_config.hpp:23-40)."""
    Y, st = lognormal_lnK(nx, ny, nz, sigma2, corr_len, seed, device)
    return K_mean * torch.exp(Y), st


def empirical_covariance(Y: torch.Tensor, shape_zyx, lags=(0, 1, 2, 4, 8)) -> dict:
    """Mean over the three axes of the sample covariance at integer lags (a check)."""
    y = Y.reshape(shape_zyx)
    y = y - y.mean()
    out = {}
    for L in lags:
        acc = []
        for ax in range(3):
            n = y.shape[ax]
            if L >= n:
                continue
            a = y.narrow(ax, 0, n - L)
            b = y.narrow(ax, L, n - L)
            acc.append(float((a * b).mean()))
        out[L] = sum(acc) / len(acc) if acc else math.nan
    return out
