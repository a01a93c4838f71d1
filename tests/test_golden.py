"""The oracle reproduces the fixtures written from the reference's own compiled code
(FFTW 2.1.2, lib_matr, RandomGen; see tests/golden/make_golden.py).  Runs without the
reference tree, so the oracle stays pinned wherever the tests run."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(G, "fftw_*.npz"))))
def test_fft_matches_fftw_fixture(path):
    f = np.load(path)
    x, z = f["x"], f["z"]
    nxp = x.shape[2]
    assert rel_l2(O.grid_fft(x, O.DATA), f["fwd_data"]) < 5e-7
    assert rel_l2(O.grid_fft(x, O.COVARIANCE), f["fwd_cov"]) < 5e-7
    # FFTW's hc2r ignores Im of the DC and Nyquist bins of this non-Hermitian input
    assert rel_l2(O.grid_ifft(z, nxp, O.PARAMETER), f["inv_param"]) < 5e-7
    assert rel_l2(O.grid_ifft(z, nxp, O.COVARIANCE), f["inv_cov"]) < 5e-7


def _spec_from_fixture(f):
    return dict(nxp=int(f["nxp"]), nyp=int(f["nyp"]), nzp=int(f["nzp"]), ntheta=f["K"].shape[1], K=f["K"],
                e1=f["e1"], e2=f["e2"], e3=f["e3"], invert=f["invert"], sigma0e=f["sigma0e"], wnc=float(f["wnc"]),
                prior_var0=f["prior_var0"], prior_mode=0, rxy=None, rt=None, lam2d=None, lam_grid=f["lam"])


def test_posterior_fp32_is_bit_identical_to_reference_lib_matr():
    f = np.load(os.path.join(G, "posterior.npz"))
    p = _spec_from_fixture(f)
    m, d, c, flags, nf = O.posterior(p, f["mean"], f["data"], f["cov"], "f32")
    assert nf == int(f["nfail"])
    assert np.array_equal(flags, f["flags"])
    assert np.array_equal(m, f["out_mean"])
    assert np.array_equal(d, f["out_res"])
    assert np.array_equal(c, f["out_cov"])


def test_posterior_fp64_within_fp32_rounding_of_reference():
    f = np.load(os.path.join(G, "posterior.npz"))
    p = _spec_from_fixture(f)
    m, d, c, _, _ = O.posterior(p, f["mean"].astype(np.complex128), f["data"].astype(np.complex128),
                                f["cov"].astype(np.complex128), "f64")
    assert rel_l2(m - f["mean"], f["out_mean"] - f["mean"]) < 1e-4
    assert rel_l2(c, f["out_cov"]) < 1e-5


def test_rng_stream_is_bit_identical_to_reference():
    f = np.load(os.path.join(G, "rng.npz"))
    st = np.array([23665], dtype=np.uint32)
    out = np.zeros(len(f["normals_23665"]))
    O.lib().orc_rnorm01_n(O._ptr(st, O.ctypes.c_uint32), O._ptr(out, O.ctypes.c_double), len(out))
    assert np.array_equal(out, f["normals_23665"])
    st = np.array([777], dtype=np.uint32)
    assert np.array_equal(O.fill_complex_noise(10, 6, 8, st), f["noise_10x6x8_seed777"])


def test_colouring_is_bit_identical_to_reference():
    f = np.load(os.path.join(G, "colour.npz"))
    z, nf = O.colour(f["cov"], f["seed"], "f32")
    assert nf == int(f["nfail"])
    assert np.array_equal(z, f["out"])
