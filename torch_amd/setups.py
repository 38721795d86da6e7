"""Native initial-condition generators for the benchmark problems.

These replace the reference's per-cell Lua ``initialise(x, y, z, xs, ys, zs)``
set-up hook (src/Torch/Torch.cpp:192-228), which needs a Lua runtime.  Each
returns primitive variables Q as a float64 array of shape (7, nz, ny, nx) in
the reference's UID order (DEN, PRE, HII, ADV, VEL0, VEL1, VEL2;
src/Torch/Common.hpp:23-25), x fastest, in physical (cgs) units -- the same
values a Lua set-up would return before Torch::toCodeUnits.
"""
from __future__ import annotations

import numpy as np

DEN, PRE, HII, ADV, VEL = 0, 1, 2, 3, 4


def sod(n: int, ny: int | None = None, nz: int | None = None) -> np.ndarray:
    """Sod shock tube along x on the unit cube: rho=1, p=1 for x<0.5 and
    rho=0.125, p=0.1 for x>=0.5, zero velocity, HII = ADV = 0."""
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    x = (np.arange(n) + 0.5) / n
    Q = np.zeros((7, nz, ny, n))
    left = x < 0.5
    Q[DEN] = np.where(left, 1.0, 0.125)[None, None, :]
    Q[PRE] = np.where(left, 1.0, 0.1)[None, None, :]
    return Q


def gaussian_random_field(n: int, seed: int, kmin: float = 2.0, kmax: float = 32.0,
                          slope: float = -11.0 / 3.0, device: str = "cpu") -> np.ndarray:
    """Real Gaussian random field on an n^3 periodic grid with isotropic power
    spectrum P(k) ~ k^slope on integer wavenumbers kmin <= |k| <= kmax (zero
    elsewhere), normalised to zero mean and unit variance.

    Mode amplitudes are drawn with numpy's PCG64 from ``seed`` over the cube
    of wavenumbers [-kc, kc]^3 (kc = floor(kmax)) in C order, so the field is
    the same for any n >= 2*kc + 2; the inverse FFT runs in float64 with
    torch on ``device``.
    """
    import torch

    kc = int(np.floor(kmax))
    if n < 2 * kc + 2:
        raise ValueError(f"grid {n} too small for kmax {kmax}")
    rng = np.random.Generator(np.random.PCG64(seed))
    k1 = np.arange(-kc, kc + 1)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    amp = np.where((kk >= kmin) & (kk <= kmax), np.power(np.maximum(kk, 1e-30), slope / 2.0), 0.0)
    a = rng.standard_normal(kk.shape) + 1j * rng.standard_normal(kk.shape)
    coef = a * amp
    spec = np.zeros((n, n, n), dtype=np.complex128)
    idx = np.arange(-kc, kc + 1) % n
    spec[np.ix_(idx, idx, idx)] = coef
    t = torch.from_numpy(spec).to(device)
    field = torch.fft.ifftn(t).real  # real part of a Hermitian-symmetrised draw
    field = field - field.mean()
    field = field / torch.sqrt((field * field).mean())
    return field.cpu().numpy()


def turbulent(n: int, sigma: float, mach: float, gamma: float = 1.67, seed: int = 1001,
              device: str = "cpu", kmax: float = 32.0) -> np.ndarray:
    """Log-normal density rho = exp(sigma*g - sigma^2/2) (mean 1) from a
    Gaussian random field g, uniform pressure 1, and three independent
    velocity fields with the same spectrum scaled to rms Mach number
    sqrt(<|v|^2 / c_s^2>) = mach with c_s = sqrt(gamma p / rho).
    Seeds are seed, seed+1, seed+2, seed+3 for rho, vx, vy, vz.  Grids
    smaller than 66 cells use kmax = n/2 - 1."""
    kmax = min(kmax, n // 2 - 1)
    g = gaussian_random_field(n, seed, kmax=kmax, device=device)
    rho = np.exp(sigma * g - 0.5 * sigma * sigma)
    p = np.ones_like(rho)
    v = [gaussian_random_field(n, seed + 1 + i, kmax=kmax, device=device) for i in range(3)]
    cs2 = gamma * p / rho
    m2 = (v[0] ** 2 + v[1] ** 2 + v[2] ** 2) / cs2
    scale = mach / np.sqrt(m2.mean())
    Q = np.zeros((7, n, n, n))
    Q[DEN] = rho
    Q[PRE] = p
    for i in range(3):
        Q[VEL + i] = v[i] * scale
    return Q
