"""Seeded synthetic initial fields for the benchmark configurations.

Host-side numpy code (one-off work before the time loop). The arrays it
returns are handed to ``sim.state.init_statespect_from(...)`` the way
fluidsim's ``init_fields.type = "in_script"`` expects
(fluidsim/base/init_fields.py:435-443, fluidsim/base/state.py:352-383).

The host generators draw from ``numpy.random.Generator(PCG64(seed))``;
``ns3d_spectrum_vv_device`` (grids too large for the host, one-off set-up
on the GPU) draws from a seeded ``torch.Generator``.
"""

import os

import numpy as np
import scipy.fft as sfft


def _workers():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 1


def _k_signed(n, dk):
    idx = np.arange(n)
    return dk * np.where(idx <= n // 2, idx, idx - n).astype(np.float64)


def wavenumbers_2d(nx, ny, Lx, Ly):
    kx = (2 * np.pi / Lx) * np.arange(nx // 2 + 1, dtype=np.float64)
    ky = _k_signed(ny, 2 * np.pi / Ly)
    KX, KY = np.meshgrid(kx, ky)
    return KX, KY


def wavenumbers_3d(nx, ny, nz, Lx, Ly, Lz):
    kx = (2 * np.pi / Lx) * np.arange(nx // 2 + 1, dtype=np.float64)
    ky = _k_signed(ny, 2 * np.pi / Ly)
    kz = _k_signed(nz, 2 * np.pi / Lz)
    return np.meshgrid(kz, ky, kx, indexing="ij")  # Kz, Ky, Kx


def _sum_wavenumbers(A, nx):
    s = A[..., 0].sum()
    if nx % 2 == 0:
        s += 2 * A[..., 1:-1].sum() + A[..., -1].sum()
    else:
        s += 2 * A[..., 1:].sum()
    return float(np.real(s))


def _dealias_cubic(arrs, Ks, ns, dks, coef):
    if not coef:
        return
    mask = np.zeros(arrs[0].shape, dtype=bool)
    for K, n, dk in zip(Ks, ns, dks):
        mask |= np.abs(K) >= coef * dk * n / 2
    for a in arrs:
        a[mask] = 0.0


def ns2d_noise_rot_fft(
    nx, ny, Lx=2 * np.pi, Ly=2 * np.pi, seed=0, kmax=16.0, velo_max=1.0,
    coef_dealiasing=2.0 / 3,
):
    """Seeded band-limited 2D noise (vorticity, half spectrum (ny, nx//2+1)).

    i.i.d. complex normal (ux, uy) on the half spectrum, zero for |k| > kmax
    and at k = 0; projected perpendicular to k (as projection_perp in
    fluidsim/solvers/ns2d/init_fields.py:84); made Hermitian by an
    irfft2/rfft2 round trip; rescaled so that max sqrt(ux^2 + uy^2) = velo_max
    (init_fields.py:96-105); rot = i (kx uy - ky ux); dealiased.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    KX, KY = wavenumbers_2d(nx, ny, Lx, Ly)
    K2 = KX**2 + KY**2
    shapeK = KX.shape
    ux = rng.standard_normal(shapeK) + 1j * rng.standard_normal(shapeK)
    uy = rng.standard_normal(shapeK) + 1j * rng.standard_normal(shapeK)
    band = (K2 <= kmax**2) & (K2 > 0)
    ux[~band] = 0.0
    uy[~band] = 0.0
    K2nz = np.where(K2 == 0, 1e-14, K2)
    tmp = (KX * ux + KY * uy) / K2nz
    ux -= KX * tmp
    uy -= KY * tmp
    dks = (2 * np.pi / Lx, 2 * np.pi / Ly)
    _dealias_cubic([ux, uy], [KX, KY], [nx, ny], dks, coef_dealiasing)
    ux_x = np.fft.irfft2(ux, s=(ny, nx), norm="forward")
    uy_x = np.fft.irfft2(uy, s=(ny, nx), norm="forward")
    vmax = np.sqrt(ux_x**2 + uy_x**2).max()
    ux_x *= velo_max / vmax
    uy_x *= velo_max / vmax
    ux = np.fft.rfft2(ux_x, norm="forward")
    uy = np.fft.rfft2(uy_x, norm="forward")
    rot = 1j * (KX * uy - KY * ux)
    _dealias_cubic([rot], [KX, KY], [nx, ny], dks, coef_dealiasing)
    return rot


def ns3d_spectrum_vv_fft(
    nx, ny=None, nz=None, L=2 * np.pi, seed=0, k_peak=4.0, energy=1.0,
    coef_dealiasing=2.0 / 3, dtype=np.complex128,
):
    """Seeded isotropic solenoidal 3D velocity with E(k) ~ k^4 exp(-2 (k/k_peak)^2).

    Per-mode amplitude sqrt(E(k) / (4 pi k^2)) times i.i.d. complex normal
    phases, Leray-projected, dealiased, made Hermitian by an irfftn/rfftn round
    trip and normalised to sum_wavenumbers(|v|^2)/2 = energy.
    Returns (vx_fft, vy_fft, vz_fft), each (nz, ny, nx//2+1).
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    rng = np.random.Generator(np.random.PCG64(seed))
    Kz, Ky, Kx = wavenumbers_3d(nx, ny, nz, L, L, L)
    K2 = Kx**2 + Ky**2 + Kz**2
    K = np.sqrt(K2)
    shapeK = K.shape
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(
            K > 0,
            np.sqrt(K**4 * np.exp(-2 * (K / k_peak) ** 2) / (4 * np.pi * K2)),
            0.0,
        )
    del K
    vv = []
    for _ in range(3):
        a = rng.standard_normal(shapeK) + 1j * rng.standard_normal(shapeK)
        a *= amp
        vv.append(a)
    del amp
    K2nz = np.where(K2 == 0, 1e-14, K2)
    tmp = (Kx * vv[0] + Ky * vv[1] + Kz * vv[2]) / K2nz
    del K2nz
    vv[0] -= Kx * tmp
    vv[1] -= Ky * tmp
    vv[2] -= Kz * tmp
    del tmp
    dk = 2 * np.pi / L
    _dealias_cubic(vv, [Kx, Ky, Kz], [nx, ny, nz], [dk, dk, dk], coef_dealiasing)
    out = []
    w = _workers()
    for a in vv:
        x = sfft.irfftn(a, s=(nz, ny, nx), axes=(0, 1, 2), norm="forward", workers=w)
        out.append(sfft.rfftn(x, norm="forward", workers=w))
    del vv
    E = _sum_wavenumbers(
        0.5 * (np.abs(out[0]) ** 2 + np.abs(out[1]) ** 2 + np.abs(out[2]) ** 2), nx
    )
    fac = np.sqrt(energy / E)
    return tuple((fac * a).astype(dtype) for a in out)


def nu8_for_kmax_eta(n, L=2 * np.pi, epsilon=1.0, coef_dealiasing=2.0 / 3):
    """Order-8 hyperviscosity with k_max * eta_8 = 1.

    eta_8 = (nu_8^3 / epsilon)^(1/22) (fluidsim/base/preprocess/pseudo_spect.py
    :351), k_max = coef_dealiasing * (n/2) * 2pi/L, so
    nu_8 = k_max^(-22/3) * epsilon^(1/3).
    """
    kmax = coef_dealiasing * (n / 2) * (2 * np.pi / L)
    return kmax ** (-22.0 / 3.0) * epsilon ** (1.0 / 3.0)


def ns3d_spectrum_vv_device(oper, seed=0, k_peak=4.0, energy=1.0):
    """Device counterpart of ns3d_spectrum_vv_fft for grids too large to
    draw on the host (1024^3): the same spectrum E(k) ~ k^4 exp(-2 (k/k_peak)^2),
    per-mode amplitude sqrt(E(k) / (4 pi k^2)) = k exp(-(k/k_peak)^2) / sqrt(4 pi)
    times i.i.d. complex normal phases from a seeded ``torch.Generator`` on the
    GPU (a different stream of numbers than the host generator), projected
    with the engine's Leray projection, dealiased, made Hermitian by an
    ifft/fft round trip through the engine and normalised to unit energy.
    Returns three device tensors of the operator's complex dtype."""
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    kx, ky, kz = oper._kdev()
    real = oper.dtype_real
    K2 = kz * kz + ky * ky + kx * kx
    amp = torch.sqrt(K2) * torch.exp(-K2 / (k_peak * k_peak)) / np.sqrt(4 * np.pi)
    del K2
    shape = tuple(oper.shapeK)
    vv = []
    for _ in range(3):
        re = torch.randn(shape, generator=g, device="cuda", dtype=real)
        im = torch.randn(shape, generator=g, device="cuda", dtype=real)
        a = torch.complex(re, im)
        del re, im
        a.mul_(amp)
        vv.append(a)
    del amp
    oper.project_perpk3d(*vv)
    for i in range(3):
        oper.dealiasing(vv[i])
        vv[i] = oper.fft(oper.ifft(vv[i]))
    E = 0.5 * sum(oper.sum_wavenumbers(a.real * a.real + a.imag * a.imag) for a in vv)
    fac = float(np.sqrt(energy / E))
    for a in vv:
        a.mul_(fac)
    return vv
