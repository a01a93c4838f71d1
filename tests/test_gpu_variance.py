"""Data variance of the seismic stacks on the device: ModelAVODynamic::ComputeDataVariance
(modelavodynamic.cpp:575-620) bit for bit against the oracle's sequential sweep, and the fixed-order
sum of squares (the same value on every call)."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(13, 9, 7, 16, 12, 10), (100, 90, 80, 120, 100, 96),
                                   (512, 512, 256, 576, 576, 320)])
def test_data_variance_matches_reference_order(ctx, shape):
    nx, ny, nz, nxp, nyp, nzp = shape
    print("data variance", shape, flush=True)
    rng = np.random.default_rng(nx)
    rv = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), dtype=np.float32)
    rv[:, :, :nxp] = (rng.standard_normal((nzp, nyp, nxp)) * rng.uniform(0.5, 2e3)).astype(np.float32)
    box = rv[:nz, :ny, :nx]
    box[rng.random(box.shape) < 0.05] = 0.0  # dead samples are left out
    g = C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, C.DATA)
    g.upload_real(rv)
    v, n = g.data_variance()
    ov, on = O.data_variance(rv, nx, ny, nz, nxp)
    assert n == on and v == ov
    s1 = g.sum_squares()
    s2 = g.sum_squares()
    assert s1 == s2 and abs(s1 - float(np.sum(box.astype(np.float64) ** 2))) <= 1e-12 * s1
    g.close()


def test_all_zero_stack_has_zero_variance(ctx):
    g = C.FFTGrid(ctx, 10, 8, 6, 12, 10, 8, C.DATA)
    g.fillInConstant(0.0)
    assert g.data_variance() == (0.0, 0)
    g.close()
