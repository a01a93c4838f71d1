"""Synthetic FFT-domain AVO inversion problems for the tests (numpy, CPU).

The recipe follows the plain CPU-runnable case: constant background ln(Vp, Vs, rho) =
ln(3000, 1500, 2.3); parameter std (0.10, 0.15, 0.05) with correlations (0.7, 0.4, 0.3); a
spherical lateral variogram and an exponential temporal correlation; a seeded Gaussian elastic
field drawn from that prior; seismic = Aki-Richards reflectivity of the field convolved with a
Ricker wavelet, plus white noise at a given S/N.  Everything here is test input generation;
the per-k_z tables come from the oracle.
"""
import math

import numpy as np

import oracle as O

LN_BG = (math.log(3000.0), math.log(1500.0), math.log(2.3))


def sigma0_from(std, corr):
    s = np.asarray(std, dtype=np.float64)
    c01, c02, c12 = corr
    C = np.array([[1, c01, c02], [c01, 1, c12], [c02, c12, 1]], dtype=np.float64)
    return C * np.outer(s, s)


def ricker_fft_order(nzp, dt_ms, peak_hz, scale=1.0):
    """Wavelet sampled at z = dt*k (k < nzp/2+1) or dt*(k-nzp): the FFT-order layout that
    Wavelet::resample produces (wavelet.cpp:424-475), here sampled analytically."""
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, k, k - nzp) * dt_ms
    return np.array([scale * O.lib().orc_ricker(float(tt), float(peak_hz)) for tt in t], dtype=np.float32)


def make_problem(nx, ny, nz, nxp=None, nyp=None, nzp=None, angles_deg=(0.0, 15.0, 30.0), dx=25.0, dy=25.0,
                 dt=4.0, peak_hz=25.0, std=(0.10, 0.15, 0.05), corr=(0.7, 0.4, 0.3), lateral_range=1000.0,
                 temporal_range=40.0, sn=5.0, wnc=0.1, vsvp=0.5, low_cut=5.0, high_cut=55.0, seed=0,
                 wavelet_scales=None, sn_per_angle=None):
    rng = np.random.default_rng(seed)
    # default padding: half a lateral range in x/y, 200 ms in z (commondata.cpp:9621-9682,
    # multiintervalgrid.cpp:885-912)
    if nxp is None:
        nxp = O.lib().orc_padding_size(nx, min(1.0, max(0.5 * lateral_range, dx) / (nx * dx)))
    if nyp is None:
        nyp = O.lib().orc_padding_size(ny, min(1.0, max(0.5 * lateral_range, dy) / (ny * dy)))
    if nzp is None:
        nzp = O.lib().orc_padding_size(nz, min(1.0, 200.0 / (nz * dt)))
    nt = len(angles_deg)
    cnxp = nxp // 2 + 1
    # reflection matrix and wavelets
    A = np.zeros((nt, 3))
    for l, a in enumerate(angles_deg):
        row = np.zeros(3)
        O.lib().orc_aki_richards(math.radians(a), vsvp, O._ptr(row, O.ctypes.c_double))
        A[l] = row
    scales = wavelet_scales if wavelet_scales is not None else [1.0] * nt
    wav = np.stack([ricker_fft_order(nzp, dt, peak_hz, s) for s in scales])
    K, e1, e2, e3, inv = O.kz_tables(nz, nzp, dt, wav, A, low_cut, high_cut, 1.0)
    # prior correlation factors
    cxy = O.corr_xy(nxp, nyp, dx, dy, lateral_range)
    n_corr_t = nzp // 2 + 1 if nzp % 2 == 0 else nzp // 2
    corr_t = np.exp(-3.0 * np.float32(dt) * np.arange(n_corr_t + 1) / np.float32(temporal_range))
    low_int_cut = int(math.floor(low_cut * (nzp * 0.001 * dt)))
    circ_t = O.circ_corr_t(corr_t, low_int_cut, nzp)
    rxy = np.fft.fft2(cxy.astype(np.float64)).real[:, :cnxp].astype(np.float32)  # COVARIANCE: unnormalised
    rt = np.fft.fft(circ_t.astype(np.float64)).real.astype(np.float32)
    S0 = sigma0_from(std, corr)
    # seeded Gaussian elastic field from the prior: m = mu + F^-1( sqrt(r(k)) L0 z )
    r = np.abs(rxy[None, :, :].astype(np.float64) * rt[:, None, None].astype(np.float64))
    L0 = np.linalg.cholesky(S0)
    zs = rng.standard_normal((3, nzp, nyp, nxp))
    zhat = np.stack([np.fft.rfftn(z, axes=(0, 1, 2)) / math.sqrt(nxp * nyp * nzp) for z in zs])
    mhat = np.einsum("ab,bzyx->azyx", L0, zhat) * np.sqrt(r)[None]
    field = np.stack([np.fft.irfftn(mh, s=(nzp, nyp, nxp), axes=(0, 1, 2)) * math.sqrt(nxp * nyp * nzp)
                      for mh in mhat])
    mean_real = np.stack([np.full((nzp, nyp, nxp), v, dtype=np.float64) for v in LN_BG])
    m_true = mean_real + field
    # seismic d_l = F^-1(sum_j K_lj(kz) m_hat_j) + white noise, unitary transforms
    mth = np.stack([np.fft.rfftn(m, axes=(0, 1, 2)) / math.sqrt(nxp * nyp * nzp) for m in m_true])
    dhat = np.einsum("zlj,jzyx->lzyx", K.astype(np.complex128), mth)
    sig = np.stack([np.fft.irfftn(dh, s=(nzp, nyp, nxp), axes=(0, 1, 2)) * math.sqrt(nxp * nyp * nzp)
                    for dh in dhat])
    sns = sn_per_angle if sn_per_angle is not None else [sn] * nt
    seis = np.empty_like(sig)
    s0e = np.zeros((nt, nt))
    sig_e = np.zeros(nt)
    for l in range(nt):
        vs = np.mean(sig[l, :nz, :ny, :nx] ** 2)
        ve = vs / (sns[l] - 1.0)  # S/N = (|s|^2 + |e|^2) / |e|^2
        seis[l] = sig[l] + rng.standard_normal(sig[l].shape) * math.sqrt(ve)
        sig_e[l] = math.sqrt(ve)
    for l in range(nt):
        for m in range(nt):
            s0e[l, m] = sig_e[l] * sig_e[m] * math.exp(-3.0 * abs(angles_deg[l] - angles_deg[m]) / 10.0)
    # padding region outside the simbox keeps the field (the reference fills it smoothly)
    return dict(nx=nx, ny=ny, nz=nz, nxp=nxp, nyp=nyp, nzp=nzp, cnxp=cnxp, ntheta=nt, dt=dt,
                K=K, e1=e1, e2=e2, e3=e3, invert=inv, sigma0e=s0e, wnc=float(wnc), prior_var0=S0,
                prior_mode=1, rxy=rxy, rt=rt, lam2d=rxy, lam_grid=None, A=A, wavelets=wav,
                mean_real=mean_real.astype(np.float32), seis_real=seis.astype(np.float32),
                cov_xy=cxy, circ_t=circ_t, m_true=m_true.astype(np.float32))


def prior_cov_grids(p):
    """The six real prior covariance grids of the reference (fillInParamCorr x priorVar0,
    seismicparametersholder.cpp:334-348), shape (6, nzp, nyp, nxp)."""
    S0 = p["prior_var0"]
    base = p["circ_t"][:, None, None].astype(np.float32) * p["cov_xy"][None, :, :].astype(np.float32)
    order = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    return np.stack([base * np.float32(S0[a, b]) for a, b in order])


def err_corr_grid(p):
    """errCorr grid of fillInErrCorr (fftgrid.cpp:777-828): corrXY in the k = 0 plane."""
    g = np.zeros((p["nzp"], p["nyp"], p["nxp"]), dtype=np.float32)
    g[0] = p["cov_xy"]
    return g


def forward_spectra(p, dtype=np.float64):
    """Reference forward transforms of the path inputs: mean (3), seismic (nt) as DATA/PARAMETER."""
    mean = np.stack([O.grid_fft(g, O.PARAMETER, dtype) for g in p["mean_real"]])
    data = np.stack([O.grid_fft(g, O.DATA, dtype) for g in p["seis_real"]])
    return mean, data
