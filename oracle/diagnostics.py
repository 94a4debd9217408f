"""FP64 restatement of fluidsim's spectra, spectral energy budgets and
spatial means for ns2d, ns3d and ns3d.strat.

Test infrastructure only (see ``oracle/__init__.py``).

Per-solver quantities follow the reference output modules (all under
fluidsim/solvers/):

  ns2d   output/spectra.py:16-28, output/spect_energy_budget.py:23-80,
         output/spatial_means.py:23-68
  ns3d   output/spectra.py:15-99, output/spect_energy_budget.py:106-190,
         output/spatial_means.py:23-77
  strat  strat/output/spectra.py:18-112,
         strat/output/spect_energy_budget.py:60-118,
         strat/output/spatial_means.py:27-106

The binning routines those modules call (compute_1dspectra,
compute_3dspectrum, compute_spectrum_kzkh, spectra1D_from_fft,
spectrum2D_from_fft) belong to fluidfft 0.4.x, which is not in the tree.
They are restated from its published algorithm:

* every mode of the half spectrum carries its Parseval weight (1 on the
  kx = 0 plane and on the kx = nx/2 plane for even nx, 2 elsewhere);
* 1D spectra bin |i_k| along one axis (negative ky, kz fold onto |k|) and
  divide by that axis' dk;
* shell spectra use floor bins k_j = j dk with a linear share
  c = (|k| - k_j) / dk between bins j and j + 1; modes with j >= nk - 1 all
  go to the last bin; 3D: dk = max(dkx, dky, dkz),
  nk = max(2, int(min_i(dk_i (n_i/2 + 1)) / dk)); 2D: dk = dkx,
  nk = nx/2 + 1 (khE = kxE);
* kz-kh spectra bin |i_kz| (nz/2 + 1 bins) times the shared kh bin
  (dkh = max(dkx, dky), nkh = max(2, int(min(dkx nkx, dky nky) / dkh))),
  divided by dkz dkh.

Parity with fluidfft itself is unpinned: neither fluidfft nor golden
spectra exist in the reference tree.  The restatement is pinned by a
brute-force loop over the full (both-sign) grid, the Parseval sum rule,
single-mode closed forms and, for the budgets, by the time derivative of
the spectra under the oracle's own time step
(tests/test_oracle_diagnostics.py).
"""

import numpy as np


def parseval_weights(oper):
    w = np.full(oper.shapeK, 2.0)
    w[..., 0] = 1.0
    if oper.nx % 2 == 0:
        w[..., -1] = 1.0
    return w


def iabs(n):
    """|signed index| along a y / z axis (the Nyquist index gives n/2)."""
    i = np.arange(n)
    return np.where(i <= n // 2, i, n - i)


class Bins:
    """Spectral bins of an oracle operator (see the module docstring)."""

    def __init__(self, oper):
        self.dim = len(oper.shapeK)
        self.nkx = oper.nx // 2 + 1
        self.nky = oper.ny // 2 + 1
        if self.dim == 3:
            self.nkz = oper.nz // 2 + 1
            self.dk = max(oper.deltakx, oper.deltaky, oper.deltakz)
            kmax = min(oper.deltakx * self.nkx, oper.deltaky * self.nky,
                       oper.deltakz * self.nkz)
            self.nk = max(2, int(kmax / self.dk))
            self.dkh = max(oper.deltakx, oper.deltaky)
            khmax = min(oper.deltakx * self.nkx, oper.deltaky * self.nky)
            self.nkh = max(2, int(khmax / self.dkh))
        else:
            self.nkz = 1
            self.dk = oper.deltakx
            self.nk = self.nkx
            self.dkh = oper.deltakx
            self.nkh = 1


def spectra_1d(q, oper):
    """(kx, ky[, kz]) spectra of the per-mode quantity q."""
    wq = parseval_weights(oper) * q
    if wq.ndim == 2:
        sx = wq.sum(axis=0)
        sy = np.bincount(iabs(oper.ny), wq.sum(axis=1), minlength=oper.ny // 2 + 1)
        return sx / oper.deltakx, sy / oper.deltaky
    sx = wq.sum(axis=(0, 1))
    sy = np.bincount(iabs(oper.ny), wq.sum(axis=(0, 2)), minlength=oper.ny // 2 + 1)
    sz = np.bincount(iabs(oper.nz), wq.sum(axis=(1, 2)), minlength=oper.nz // 2 + 1)
    return sx / oper.deltakx, sy / oper.deltaky, sz / oper.deltakz


def share(k, dk, n):
    """Floor bin and linear share; j >= n - 1 goes to the last bin."""
    j = (k / dk).astype(np.int64)
    last = j >= n - 1
    c = np.where(last, 0.0, (k - j * dk) / dk)
    return np.where(last, n - 1, j), c


def _shared_hist(wq, k, dk, n, offset=0, size=None):
    j, c = share(k, dk, n)
    size = n + 1 if size is None else size
    h = np.bincount((offset + j).ravel(), ((1.0 - c) * wq).ravel(), minlength=size)
    h += np.bincount((offset + j + 1).ravel(), (c * wq).ravel(), minlength=size)
    return h


def shell_spectrum(q, oper):
    """compute_3dspectrum / spectrum2D_from_fft."""
    b = Bins(oper)
    wq = parseval_weights(oper) * q
    return _shared_hist(wq, np.sqrt(oper.K2), b.dk, b.nk)[: b.nk] / b.dk


def spectrum_kzkh(q, oper):
    """compute_spectrum_kzkh: (nkz, nkh) array."""
    b = Bins(oper)
    wq = parseval_weights(oper) * q
    kh = np.sqrt(oper.Kx**2 + oper.Ky**2)
    jz = np.broadcast_to(iabs(oper.nz)[:, None, None], wq.shape)
    stride = b.nkh + 1
    h = _shared_hist(wq, kh, b.dkh, b.nkh, offset=jz * stride, size=b.nkz * stride)
    return h.reshape(b.nkz, stride)[:, : b.nkh] / (oper.deltakz * b.dkh)


# --------------------------------------------------------------------- ns2d
def energy_2d(oper, rot_fft):
    """compute_energy_fft (ns2d/output/__init__.py:110-114)."""
    ux, uy = oper.vecfft_from_rotfft(rot_fft)
    return (np.abs(ux) ** 2 + np.abs(uy) ** 2) / 2.0


def ns2d_spectra(oper, rot_fft):
    """SpectraNS2D.compute (ns2d/output/spectra.py:16-28)."""
    e = energy_2d(oper, rot_fft)
    kx, ky = spectra_1d(e, oper)
    return ({"spectrum1Dkx_E": kx, "spectrum1Dky_E": ky},
            {"spectrum2D_E": shell_spectrum(e, oper)})


def ns2d_budget_fields(oper, rot_fft, beta=0.0):
    """Frot, Fx, Fy of SpectralEnergyBudgetNS2D.compute
    (ns2d/output/spect_energy_budget.py:27-57)."""
    ifft, fft = oper.ifft, oper.fft
    ux_fft, uy_fft = oper.vecfft_from_rotfft(rot_fft)
    ux, uy = ifft(ux_fft), ifft(uy_fft)
    px_rot, py_rot = (ifft(a) for a in oper.gradfft_from_fft(rot_fft))
    px_ux, py_ux = (ifft(a) for a in oper.gradfft_from_fft(ux_fft))
    px_uy, py_uy = (ifft(a) for a in oper.gradfft_from_fft(uy_fft))
    Frot_fft = fft(-ux * px_rot - uy * (py_rot + beta))
    Fx_fft = fft(-ux * px_ux - uy * py_ux)
    Fy_fft = fft(-ux * px_uy - uy * py_uy)
    oper.dealiasing(Frot_fft, Fx_fft, Fy_fft)
    return Frot_fft, Fx_fft, Fy_fft


def ns2d_budget(oper, rot_fft, beta=0.0):
    """SpectralEnergyBudgetNS2D.compute (spect_energy_budget.py:23-80)."""
    Frot_fft, Fx_fft, Fy_fft = ns2d_budget_fields(oper, rot_fft, beta)
    ux_fft, uy_fft = oper.vecfft_from_rotfft(rot_fft)
    transferZ = np.real(rot_fft.conj() * Frot_fft + rot_fft * Frot_fft.conj()) / 2.0
    transferE = np.real(ux_fft.conj() * Fx_fft + ux_fft * Fx_fft.conj()
                        + uy_fft.conj() * Fy_fft + uy_fft * Fy_fft.conj()) / 2.0
    return {"transfer2D_E": shell_spectrum(transferE, oper),
            "transfer2D_Z": shell_spectrum(transferZ, oper)}


def ns2d_means(oper, rot_fft, f_d, f_d_hypo, Frot_fft=None, deltat=None):
    """SpatialMeansNS2D._save_one_time (ns2d/output/spatial_means.py:23-68)."""
    sw = oper.sum_wavenumbers
    e = energy_2d(oper, rot_fft)
    z = np.abs(rot_fft) ** 2 / 2.0
    out = {"E": sw(e), "Z": sw(z), "epsK": sw(f_d * 2 * e), "epsK_hypo": sw(f_d_hypo * 2 * e),
           "epsZ": sw(f_d * 2 * z), "epsZ_hypo": sw(f_d_hypo * 2 * z)}
    if Frot_fft is not None:
        Fx, Fy = oper.vecfft_from_rotfft(Frot_fft)
        ux, uy = oper.vecfft_from_rotfft(rot_fft)
        out["PZ1"] = sw(np.real(rot_fft.conj() * Frot_fft))
        out["PZ2"] = deltat / 2 * sw(np.abs(Frot_fft) ** 2)
        out["PK1"] = sw(np.real(ux.conj() * Fx + ux * Fx.conj()
                                + uy.conj() * Fy + uy * Fy.conj()) / 2)
        out["PK2"] = sw((np.abs(Fx) ** 2 + np.abs(Fy) ** 2) * deltat / 2)
    return out


# --------------------------------------------------------------------- 3D
def _add_1d(d, name, q, oper):
    sx, sy, sz = spectra_1d(q, oper)
    d[f"spectra_{name}_kx"], d[f"spectra_{name}_ky"], d[f"spectra_{name}_kz"] = sx, sy, sz


def ns3d_spectra(oper, vx, vy, vz, b=None, N=None):
    """SpectraNS3D.compute (ns3d/output/spectra.py:15-99) and, with b,
    SpectraNS3DStrat.compute (strat/output/spectra.py:18-112); the kzkh part
    is returned whatever kzkh_periodicity says."""
    ex, ey, ez = (0.5 * np.abs(a) ** 2 for a in (vx, vy, vz))
    urx, ury, udx, udy = oper.urudfft_from_vxvyfft(vx, vy)
    Khr = 0.5 * (np.abs(urx) ** 2 + np.abs(ury) ** 2)
    Khd = 0.5 * (np.abs(udx) ** 2 + np.abs(udy) ** 2)
    d1 = {}
    for name, q in (("vx", ex), ("vy", ey), ("vz", ez)):
        _add_1d(d1, name, q, oper)
    for ax in ("kx", "ky", "kz"):
        d1[f"spectra_E_{ax}"] = (d1[f"spectra_vx_{ax}"] + d1[f"spectra_vy_{ax}"]
                                 + d1[f"spectra_vz_{ax}"])
    quantities = []
    if b is not None:
        A = 0.5 / N**2 * np.abs(b) ** 2
        quantities.append(("A", A))
    quantities += [("Khr", Khr), ("Khd", Khd)]
    for name, q in quantities:
        _add_1d(d1, name, q, oper)
    d3 = {f"spectra_{n}": shell_spectrum(q, oper) for n, q in
          (("vx", ex), ("vy", ey), ("vz", ez))}
    d3["spectra_E"] = d3["spectra_vx"] + d3["spectra_vy"] + d3["spectra_vz"]
    for name, q in quantities:
        d3[f"spectra_{name}"] = shell_spectrum(q, oper)
    if b is None:
        kzkh = {"K": spectrum_kzkh(ex + ey + ez, oper), "Khr": spectrum_kzkh(Khr, oper),
                "Khd": spectrum_kzkh(Khd, oper)}
    else:
        kzkh = {"A": spectrum_kzkh(A, oper), "Khr": spectrum_kzkh(Khr, oper),
                "Khd": spectrum_kzkh(Khd, oper), "Kz": spectrum_kzkh(ez, oper)}
    return d1, d3, kzkh


def ns3d_budget_terms(oper, vx, vy, vz, projector):
    """P fft(v x omega) (ns3d/output/spect_energy_budget.py:128-152)."""
    ifft, fft = oper.ifft, oper.fft
    ox, oy, oz = (ifft(a) for a in oper.rotfft_from_vecfft(vx, vy, vz))
    px, py, pz = ifft(vx), ifft(vy), ifft(vz)
    fx, fy, fz = oper.vector_product(px, py, pz, ox, oy, oz)
    fx_fft, fy_fft, fz_fft = fft(fx), fft(fy), fft(fz)
    projector(fx_fft, fy_fft, fz_fft)
    return fx_fft, fy_fft, fz_fft


def _budget_add(results, name, q, oper):
    results[name] = spectrum_kzkh(q, oper)
    sx, sy, sz = spectra_1d(q, oper)
    results[name + "_kx"], results[name + "_ky"], results[name + "_kz"] = sx, sy, sz


def ns3d_budget(oper, vx, vy, vz, f_d, projector=None, b=None, N=None):
    """SpectralEnergyBudgetNS3D.compute (ns3d/output/spect_energy_budget.py
    :106-190) and, with b, SpectralEnergyBudgetNS3DStrat.compute
    (strat/output/spect_energy_budget.py:60-118)."""
    if projector is None:
        projector = oper.project_perpk3d
    fx, fy, fz = ns3d_budget_terms(oper, vx, vy, vz, projector)
    res = {}
    _budget_add(res, "transfer_Kh",
                np.real(vx.conj() * fx) + np.real(vy.conj() * fy), oper)
    _budget_add(res, "transfer_Kz", np.real(vz.conj() * fz), oper)
    urx, ury, _, _ = oper.urudfft_from_vxvyfft(vx, vy)
    _budget_add(res, "transfer_Khr",
                np.real(urx.conj() * fx) + np.real(ury.conj() * fy), oper)
    _budget_add(res, "diss_Kh", f_d * (np.abs(vx) ** 2 + np.abs(vy) ** 2), oper)
    _budget_add(res, "diss_Kz", f_d * np.abs(vz) ** 2, oper)
    if b is not None:
        _budget_add(res, "diss_A", 1 / N**2 * f_d * np.abs(b) ** 2, oper)
        _budget_add(res, "conv_K2A", -np.real(vz.conj() * b), oper)
        ifft = oper.ifft
        fb = -1 / N**2 * oper.div_vb_fft_from_vb(ifft(vx), ifft(vy), ifft(vz), ifft(b))
        _budget_add(res, "transfer_A", np.real(b.conj() * fb), oper)
    return res


def ns3d_means(oper, vx, vy, vz, f_d, f_d_hypo, nu_4=0.0, nu_8=0.0, f=None, deltat=None):
    """SpatialMeansNS3D._save_one_time (ns3d/output/spatial_means.py:23-77)."""
    sw = oper.sum_wavenumbers
    ex, ey, ez = (0.5 * np.abs(a) ** 2 for a in (vx, vy, vz))
    e = ex + ey + ez
    out = {"Ex": sw(ex), "Ey": sw(ey), "Ez": sw(ez)}
    out["E"] = out["Ex"] + out["Ey"] + out["Ez"]
    out["epsK"] = sw(f_d * 2 * e)
    out["epsK_hypo"] = sw(f_d_hypo * 2 * e)
    if nu_4 > 0:
        out["epsK4"] = sw(nu_4 * oper.K4 * 2 * e)
    if nu_8 > 0:
        out["epsK8"] = sw(nu_8 * oper.K8 * 2 * e)
    if f is not None:
        fx, fy, fz = f
        out["PK1"] = sw(np.real(vx.conj() * fx + vy.conj() * fy + vz.conj() * fz))
        out["PK2"] = sw((np.abs(fx) ** 2 + np.abs(fy) ** 2 + np.abs(fz) ** 2) * deltat / 2)
    return out


def strat_means(oper, vx, vy, vz, b, N, f_d, f_d_hypo, nu_4=0.0, nu_8=0.0, f=None,
                fb=None, deltat=None):
    """SpatialMeansNS3DStrat._save_one_time (strat/output/spatial_means.py
    :27-106)."""
    sw = oper.sum_wavenumbers
    urx, ury, udx, udy = oper.urudfft_from_vxvyfft(vx, vy)
    nrj_A = 0.5 / N**2 * np.abs(b) ** 2
    nrj_Kz = 0.5 * np.abs(vz) ** 2
    nrj_Khr = 0.5 * (np.abs(urx) ** 2 + np.abs(ury) ** 2)
    nrj_Khd = 0.5 * (np.abs(udx) ** 2 + np.abs(udy) ** 2)
    eK = nrj_Kz + nrj_Khr + nrj_Khd
    shear = oper.Kx**2 + oper.Ky**2 == 0.0
    out = {"EA": sw(nrj_A), "EAs": sw(nrj_A * shear), "EKz": sw(nrj_Kz),
           "EKhs": sw(nrj_Khr * shear), "EKhd": sw(nrj_Khd)}
    out["EKhr"] = sw(nrj_Khr) - out["EKhs"]
    out["E"] = out["EA"] + out["EKz"] + out["EKhr"] + out["EKhd"] + out["EKhs"]
    out["epsK"] = sw(f_d * 2 * eK)
    out["epsK_hypo"] = sw(f_d_hypo * 2 * eK)
    out["epsA"] = sw(f_d * 2 * nrj_A)
    out["epsA_hypo"] = sw(f_d_hypo * 2 * nrj_A)
    if nu_4 > 0:
        out["epsK4"] = sw(nu_4 * oper.K4 * 2 * eK)
        out["epsA4"] = sw(nu_4 * oper.K4 * 2 * nrj_A)
    if nu_8 > 0:
        out["epsK8"] = sw(nu_8 * oper.K8 * 2 * eK)
        out["epsA8"] = sw(nu_8 * oper.K8 * 2 * nrj_A)
    if f is not None:
        fx, fy, fz = f
        out["PK1"] = sw(np.real(vx.conj() * fx + vy.conj() * fy + vz.conj() * fz))
        out["PK2"] = sw(np.abs(fx) ** 2 + np.abs(fy) ** 2 + np.abs(fz) ** 2) * deltat / 2
        out["PA1"] = sw(np.real(b.conj() * fb)) / N**2
        out["PA2"] = sw(np.abs(fb) ** 2) * deltat / 2 / N**2
    return out
