"""Seeded synthetic inversion problems, generated on the device (for benchmarks and serving runs).

A problem follows the reference's synthetic set-up: a prior of background ln(Vp, Vs, rho) with a
3x3 parameter covariance, a spherical lateral and an exponential temporal correlation
(SeismicParametersHolder::InitializeCorrelations, seismicparametersholder.cpp:235-352); an
elastic field drawn from that prior (AVOInversion::simulate, avoinversion.cpp:1355-1540, with
the prior in place of the posterior); seismic = Aki-Richards reflectivity convolved with a
Ricker wavelet per angle (the reference's forward model K m) plus white noise at a stated S/N
((|s|^2 + |e|^2) / |e|^2), and the angular error covariance of
ModelAVODynamic::SetupErrorCorrelation (modelavodynamic.cpp:623-651).

Everything large lives in device grids; the host holds only the per-plane tables.
"""
import math

import numpy as np

from . import (COVARIANCE, DATA, PARAMETER, PREC_MIXED, Arena, FFTGrid, forward_model, invert, posterior,
               simulate)
from . import model as M

BACKGROUND = (math.log(3000.0), math.log(1500.0), math.log(2.3))


def spec(nx, ny, nz, angles, peak_hz, seed, nxp=None, nyp=None, nzp=None, dx=25.0, dy=25.0, dt=4.0,
         std=(0.10, 0.15, 0.05), corr=(0.7, 0.4, 0.3), lateral_range=1000.0, temporal_range=40.0, wnc=0.1,
         vsvp=0.5, low_cut=5.0, high_cut=55.0, realisations=0, wavelet_scale=(0.8, 1.2), sn=(3.0, 8.0),
         ormsby_hz=None, out_sets=None):
    """One problem.  Unpadded sizes are padded by the reference's rule when nxp/nyp/nzp are not
    given: x/y by half the lateral range, z by a 200 ms wavelet length
    (CommonData::EstimateXYPaddingSizes, commondata.cpp:9621-9682; EstimateZPaddingSize,
    multiintervalgrid.cpp:885-912), each rounded up to a 2^a 3^b 5^c 7^d length.  ormsby_hz: corner
    frequencies of an Ormsby wavelet instead of the Ricker; out_sets: realisation grid sets kept
    (realisations beyond are streamed through them; default one set per realisation)."""
    if nxp is None:
        nxp = M.padding_size(nx, min(1.0, max(0.5 * lateral_range, dx) / (nx * dx)))
    if nyp is None:
        nyp = M.padding_size(ny, min(1.0, max(0.5 * lateral_range, dy) / (ny * dy)))
    if nzp is None:
        nzp = M.padding_size(nz, min(1.0, 200.0 / (nz * dt)))
    return dict(nx=nx, ny=ny, nz=nz, nxp=nxp, nyp=nyp, nzp=nzp, dx=dx, dy=dy, dt=dt, angles=tuple(angles),
                peak_hz=peak_hz, std=tuple(std), corr=tuple(corr), lateral_range=lateral_range,
                temporal_range=temporal_range, wnc=wnc, vsvp=vsvp, low_cut=low_cut, high_cut=high_cut, seed=seed,
                realisations=realisations, wavelet_scale=wavelet_scale, sn=sn, ormsby_hz=ormsby_hz,
                out_sets=out_sets)


def single_cube_spec(realisations=2):
    """The production single cube: 512 x 512 x 256 padded to 576 x 576 x 320, dx = dy = 25 m,
    dt = 4 ms, 5 angles 0-40 deg, a 30 Hz Ricker wavelet scaled U(0.8, 1.2) per angle, S/N U(3, 8)
    per angle, the default prior (spherical 1000 m lateral, exponential 40 ms temporal)."""
    return spec(512, 512, 256, (0.0, 10.0, 20.0, 30.0, 40.0), 30.0, seed=20261202, nxp=576, nyp=576, nzp=320,
                realisations=realisations)


def large_cube_spec(realisations=20):
    """The large cube with simulation: 1024 x 1024 x 320 padded to 1152 x 1152 x 384, 6 angles
    0-50 deg, an Ormsby 5-10-60-80 Hz wavelet, 1500 m lateral range, 20 posterior realisations
    (coloured in pairs), every one kept resident: 84 grids of 2.04 GB, 171 GB of HBM."""
    return spec(1024, 1024, 320, tuple(float(a) for a in range(0, 60, 10)), 25.0, seed=20261205, nxp=1152,
                nyp=1152, nzp=384, lateral_range=1500.0, ormsby_hz=(5.0, 10.0, 60.0, 80.0),
                realisations=realisations)


def cells(s):
    """Unpadded cells of one problem (the unit of the cells/s throughput)."""
    return s["nx"] * s["ny"] * s["nz"]


def grid_bytes(s):
    """Device bytes of one grid of the problem (in-place R2C layout)."""
    return 2 * (s["nxp"] // 2 + 1) * s["nyp"] * s["nzp"] * 4


def out_sets(s):
    """Realisation grid sets a problem holds (at least one, used to draw the synthetic truth)."""
    k = s.get("out_sets")
    return max(1, s["realisations"] if k is None else k)


def ngrids(s):
    """Grids a problem holds: 3 + n_theta inputs, 3 + n_theta work spectra, 6 covariance grids and
    3 grids per realisation set."""
    return 2 * (3 + len(s["angles"])) + 6 + 3 * out_sets(s)


def footprint(s):
    """Device bytes a problem holds (its arena)."""
    return Arena.grid_bytes(s["nxp"], s["nyp"], s["nzp"]) * ngrids(s)


def log_uniform_int(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def uneven_mix(n=64, seed=20261203):
    """Independent inversions of uneven sizes: grid sizes log-uniform from 32 x 32 x 64 to
    256 x 256 x 192, 3-6 angles, wavelet peak U(20, 40) Hz, ln-elastic std U(0.05, 0.2),
    lateral range U(300, 2000) m, all from one seed."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx, ny = log_uniform_int(rng, 32, 256), log_uniform_int(rng, 32, 256)
        nz = log_uniform_int(rng, 64, 192)
        nt = int(rng.integers(3, 7))
        angles = tuple(float(a) for a in np.linspace(0.0, 10.0 * (nt - 1), nt))
        std = tuple(float(v) for v in rng.uniform(0.05, 0.2, 3))
        out.append(spec(nx, ny, nz, angles, float(rng.uniform(20.0, 40.0)), seed=int(rng.integers(1 << 30)),
                        std=std, lateral_range=float(rng.uniform(300.0, 2000.0))))
    return out


def serving_requests(n, seed=20261204):
    """Requests of a serving run: grid sizes log-uniform from 64 x 64 x 64 to 384 x 384 x 256,
    3-8 angles (0, 15, 30 deg for three, evenly up to 45 deg for more), 0-10 posterior
    realisations, a 25 Hz Ricker wavelet, S/N 5 and the default prior, each from its own seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx, ny = log_uniform_int(rng, 64, 384), log_uniform_int(rng, 64, 384)
        nz = log_uniform_int(rng, 64, 256)
        nt = int(rng.integers(3, 9))
        angles = tuple(float(a) for a in np.linspace(0.0, min(15.0 * (nt - 1), 45.0), nt))
        out.append(spec(nx, ny, nz, angles, 25.0, seed=int(rng.integers(1 << 30)),
                        realisations=int(rng.integers(0, 11)), wavelet_scale=(1.0, 1.0), sn=(5.0, 5.0)))
    return out


class Problem:
    """Device-resident inputs and work grids of one synthetic inversion, with the plain
    per-request call sequence (step) and the batch descriptor (item)."""

    def __init__(self, ctx, s, arena=True):
        """arena: carve every grid from one allocation (one allocation and one release per
        problem) instead of one allocation per grid."""
        self.ctx, self.s = ctx, s
        self.arena = Arena(ctx, footprint(s)) if arena else None
        rng = np.random.default_rng(s["seed"])
        nx, ny, nz, nxp, nyp, nzp = (s[k] for k in ("nx", "ny", "nz", "nxp", "nyp", "nzp"))
        self.nt = nt = len(s["angles"])
        self.R = s["realisations"]
        scales = rng.uniform(*s["wavelet_scale"], nt)
        sn = rng.uniform(*s["sn"], nt)
        tab = M.inversion_tables(nz, nzp, s["dt"], s["angles"], s["peak_hz"], scales, s["vsvp"], s["low_cut"],
                                 s["high_cut"], s.get("ormsby_hz"))
        cxy = M.spherical_corr_xy(nxp, nyp, s["dx"], s["dy"], s["lateral_range"])
        self.rxy = M.lateral_spectrum(cxy)
        self.rt = M.temporal_spectrum(M.circ_corr_t(s["dt"], s["temporal_range"], nzp, s["low_cut"]))
        S0 = M.sigma0(s["std"], s["corr"])
        mk = lambda ct: FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, ct, arena=self.arena)
        self.mean_in = [mk(PARAMETER) for _ in range(3)]
        self.seis_in = [mk(DATA) for _ in range(nt)]
        self.mean_w = [mk(PARAMETER) for _ in range(3)]
        self.data_w = [mk(DATA) for _ in range(nt)]
        self.cov = [mk(COVARIANCE) for _ in range(6)]
        self.out = [[mk(PARAMETER) for _ in range(3)] for _ in range(out_sets(s))]
        for g, v in zip(self.mean_in, BACKGROUND):
            g.fillInConstant(v)
        # prior covariance spectra: a posterior with no inverted plane writes the prior
        for g, src in zip(self.mean_w, self.mean_in):
            g.fft_from(src)
        for g, src in zip(self.data_w, self.seis_in):
            g.fft_from(src)
        base = dict(K=tab["K"], e1=tab["e1"], e2=tab["e2"], e3=tab["e3"], sigma0e=np.eye(nt), wnc=s["wnc"],
                    prior_var0=S0)
        posterior(ctx, dict(base, invert=np.zeros(nzp, np.int32)), self.mean_w, self.data_w, self.cov,
                  rxy=self.rxy, rt=self.rt, lam2d=self.rxy, want_residual=False)
        # elastic field from the prior, forward-modelled seismic, white noise at S/N
        truth = self.out[0]
        simulate(ctx, self.cov, self.mean_in, truth, seed=s["seed"], realisation=1000)
        for g, src in zip(self.mean_w, truth):
            g.fft_from(src)
        forward_model(ctx, tab["K"], self.mean_w, self.data_w)
        sig_e = np.zeros(nt)
        for l in range(nt):
            self.data_w[l].setType(DATA)
            self.data_w[l].invFFTInPlace()
            var_s = self.data_w[l].sum_squares() / (nx * ny * nz)
            sig_e[l] = math.sqrt(max(var_s, 1e-30) / (sn[l] - 1.0))
            self.seis_in[l].add(self.data_w[l])
            self.seis_in[l].add_white_noise(float(sig_e[l]), s["seed"], 500 + l)
        self.tables = dict(base, invert=tab["invert"], sigma0e=M.angular_error_covariance(sig_e, s["angles"]))
        self.n_inverted_planes = int(np.sum(tab["invert"] != 0))
        ctx.sync()

    def step(self, realisations=None, stats=False, precision=PREC_MIXED):
        """One request: cg_invert (forward transforms, posterior, realisations, inverse transforms
        of the means; the split-z passes where the grid has them).  Realisations beyond the kept
        output sets are streamed through them afterwards, with the real-domain means.  Returns
        (inverted, failures, FP64 re-solves) when stats (a synchronisation), zeros otherwise."""
        R = self.R if realisations is None else realisations
        if R <= len(self.out):  # the realisations inside cg_invert (one output set each)
            st = invert(self.ctx, self.item(realisations=R), precision=precision, cond_limit=1e6, stats=stats)
        else:  # more realisations than output sets: streamed through them afterwards
            st = invert(self.ctx, self.item(realisations=0), precision=precision, cond_limit=1e6, stats=stats)
            self.realisations(R)
        return st if stats else (0, 0, 0)

    def step_dropin(self, realisations=None, stats=False, precision=PREC_MIXED):
        """The same request as the reference's call sequence, one drop-in call each: forward FFTs,
        posterior, realisations, inverse FFTs of the means.  The realisations run while the means
        are still spectra, which the colouring adds."""
        self.forward()
        st = self.posterior(stats=stats, precision=precision)
        self.realisations(realisations)
        self.inverse_means()
        return st

    def forward(self):
        """Forward FFTs of the resident real-domain inputs into the work spectra."""
        for g, src in zip(self.mean_w, self.mean_in):
            g.fft_from(src)
        for g, src in zip(self.data_w, self.seis_in):
            g.fft_from(src)

    def posterior(self, stats=False, precision=PREC_MIXED):
        """Per-wavenumber posterior on the work spectra; returns (inverted, failures, FP64 re-solves)
        when stats (a synchronisation), zeros otherwise."""
        return posterior(self.ctx, self.tables, self.mean_w, self.data_w, self.cov, rxy=self.rxy, rt=self.rt,
                         lam2d=self.rxy, want_residual=False, precision=precision, cond_limit=1e6, stats=stats)

    def realisations(self, realisations=None):
        """Posterior realisations 0..R-1 from the posterior means (spectral or real-domain, as
        the means are at the time); realisation r ends in
        self.out[r % len(self.out)].  Realisations beyond the kept sets are streamed through them;
        with two or more sets, consecutive realisations are coloured in pairs (one read of the
        covariance per pair)."""
        R = self.R if realisations is None else realisations
        n, r = len(self.out), 0
        while r < R:
            if r + 1 < R and n > 1:
                simulate(self.ctx, self.cov, self.mean_w, self.out[r % n], seed=self.s["seed"], realisation=r,
                         next_out=self.out[(r + 1) % n])
                r += 2
            else:
                simulate(self.ctx, self.cov, self.mean_w, self.out[r % n], seed=self.s["seed"], realisation=r)
                r += 1

    def inverse_means(self):
        for g in self.mean_w:
            g.invFFTInPlace()

    def item(self, realisations=None):
        """Batch descriptor: the same inputs, work grids and outputs as step()."""
        R = self.R if realisations is None else realisations
        if R > len(self.out):
            raise ValueError("a batch needs one realisation grid set per realisation")
        return dict(tables=self.tables, rxy=self.rxy, rt=self.rt, lam2d=self.rxy, seed=self.s["seed"],
                    mean_in=self.mean_in, seis_in=self.seis_in, mean=self.mean_w, data=self.data_w, cov=self.cov,
                    out=self.out[:R])

    def grids(self):
        return (self.mean_in + self.seis_in + self.mean_w + self.data_w + self.cov +
                [g for r in self.out for g in r])

    def close(self):
        for g in self.grids():
            g.close()
        if self.arena is not None:
            self.arena.close()
            self.arena = None
