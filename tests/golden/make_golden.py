"""Writes the golden fixtures of tests/test_golden.py from the reference's own compiled code.

Run in a container that has the CRAVA tree (oracle/_ref/libcrava_ref.so built by
`make -C oracle ref`):  python tests/golden/make_golden.py
Every fixture holds inputs and the outputs the reference code produced for them:
  fftw_*.npz      FFTW 2.1.2 rfftw3d (FFTGrid::fftInPlace / invFFTInPlace scaling)
  posterior.npz   the per-wavenumber loop driven over the reference's lib_matr
  rng.npz         RandomGen::rnorm01 stream (seed 23665) and fillInComplexNoise layout
  colour.npz      lib_matrCholCpx(3) + lib_matrProdCholVec colouring
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
import problem as P  # noqa: E402


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    rng = np.random.default_rng(20261202)
    for shape in [(10, 6, 8), (5, 7, 9), (14, 10, 12), (30, 12, 16)]:
        nzp, nyp, nxp = shape
        x = rng.standard_normal(shape).astype(np.float32)
        cnxp = nxp // 2 + 1
        z = (rng.standard_normal((nzp, nyp, cnxp)) + 1j * rng.standard_normal((nzp, nyp, cnxp))).astype(np.complex64)
        np.savez_compressed(os.path.join(HERE, "fftw_%dx%dx%d.npz" % (nxp, nyp, nzp)),
                            x=x, fwd_data=O.ref_grid_fft(x, O.DATA), fwd_cov=O.ref_grid_fft(x, O.COVARIANCE),
                            z=z, inv_param=O.ref_grid_ifft(z, nxp, O.PARAMETER),
                            inv_cov=O.ref_grid_ifft(z, nxp, O.COVARIANCE))
    p = P.make_problem(12, 10, 16, seed=77, angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0), peak_hz=30.0)
    mean, data = P.forward_spectra(p, np.float32)
    covs = np.stack([O.grid_fft(g, O.COVARIANCE, np.float32) for g in P.prior_cov_grids(p)])
    lam = O.grid_fft(P.err_corr_grid(p), O.COVARIANCE, np.float32)
    q = dict(p, prior_mode=0, lam_grid=lam)
    m, d, c, flags, nf = O.ref_posterior(q, mean, data, covs)
    np.savez_compressed(os.path.join(HERE, "posterior.npz"), nxp=p["nxp"], nyp=p["nyp"], nzp=p["nzp"],
                        K=p["K"], e1=p["e1"], e2=p["e2"], e3=p["e3"], invert=p["invert"], sigma0e=p["sigma0e"],
                        wnc=p["wnc"], prior_var0=p["prior_var0"], mean=mean, data=data, cov=covs, lam=lam,
                        out_mean=m, out_res=d, out_cov=c, flags=flags, nfail=nf)
    O.ref().ref_rng_seed(23665)
    normals = np.zeros(4000)
    O.ref().ref_rnorm01_n(O._ptr(normals, O.ctypes.c_double), len(normals))
    O.ref().ref_rng_seed(777)
    noise = np.zeros((8, 6, 6), dtype=np.complex64)
    O.ref().ref_fill_complex_noise(10, 6, 8, O._ptr(noise.view(np.float32), O.ctypes.c_float))
    np.savez_compressed(os.path.join(HERE, "rng.npz"), normals_23665=normals, noise_10x6x8_seed777=noise)
    seed = (rng.standard_normal((3,) + c.shape[1:]) + 1j * rng.standard_normal((3,) + c.shape[1:])).astype(np.complex64)
    zc, nfc = O.ref_colour(c, seed)
    np.savez_compressed(os.path.join(HERE, "colour.npz"), cov=c, seed=seed, out=zc, nfail=nfc)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
