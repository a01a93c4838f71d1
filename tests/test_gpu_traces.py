"""CommonData::FillInData's seismic branch on the device (cg_grid_fill_from_traces): SEG-Y
traces onto the padded grid -- guard-zone taper, ten-fold resampling through the spectrum, the
cell interpolation -- against the oracle on the same traces and columns."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O

pytestmark = pytest.mark.gpu


def _case(seed, nx, ny, nz, nxp, nyp, nzp, nmin, nmax, missing=0.1):
    rng = np.random.default_rng(seed)
    ntr = nx * ny
    traces = []
    for _ in range(ntr):
        n = int(rng.integers(nmin, nmax + 1))
        t = np.arange(n)
        tr = np.sin(2 * np.pi * t * rng.uniform(0.02, 0.2)) + 0.3 * rng.standard_normal(n)
        traces.append(tr.astype(np.float32))
    z0_data = rng.uniform(0.0, 40.0, ntr).astype(np.float32)
    rnxp = 2 * (nxp // 2 + 1)
    # columns as FillInData resolves them for a simbox over the traces: the fill number of the
    # padding, some columns without a trace (outside the data or dead)
    ct = np.full((nyp, rnxp), -1, dtype=np.int32)
    for j in range(nyp):
        for i in range(nxp):
            ri, rj = O.fill_number(i, nx, nxp), O.fill_number(j, ny, nyp)
            if rng.uniform() >= missing:
                ct[j, i] = rj * nx + ri
    z0_grid = (60.0 + rng.uniform(-8.0, 8.0, (nyp, rnxp))).astype(np.float32)
    dz_grid = rng.uniform(3.0, 5.0, (nyp, rnxp)).astype(np.float32)
    return traces, z0_data, ct, z0_grid, dz_grid


@pytest.mark.parametrize("dims,lengths,sl", [((20, 12, 30, 24, 16, 40), (60, 80), 20.0),
                                             ((9, 7, 25, 12, 10, 30), (41, 45), 12.0),      # nt = 45, odd
                                             ((6, 5, 60, 8, 6, 72), (900, 1000), 100.0)])   # nt = 1000
def test_fill_from_traces_matches_oracle(ctx, dims, lengths, sl):
    nx, ny, nz, nxp, nyp, nzp = dims
    print("fill from traces", dims, lengths, sl, flush=True)
    traces, z0d, ct, z0g, dzg = _case(7, nx, ny, nz, nxp, nyp, nzp, *lengths)
    dz_data = 4.0
    want, nt = O.fill_from_traces(traces, z0d, dz_data, ct, z0g, dzg, nz, nzp, sl)
    g = C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, C.DATA)
    g.fillInData(traces, z0d, dz_data, ct, z0g, dzg, sl)
    got = g.download_real(padded=True)  # every column of the in-place rows, as FillInData fills them
    err = float(np.max(np.abs(got - want))) / float(np.max(np.abs(want)))
    print("nt %d: max |device - oracle| / max |oracle| %.2e" % (nt, err), flush=True)
    assert err < 2e-6
    assert np.all(got[:, ct < 0] == 0.0)
    g.close()


def test_fill_from_traces_refuses_short_traces(ctx):
    g = C.FFTGrid(ctx, 6, 5, 10, 8, 6, 12, C.DATA)
    traces = [np.ones(8, np.float32)]
    ct = np.zeros((6, 10), np.int32)
    z = np.zeros((6, 10), np.float32)
    with pytest.raises(C.CravaError):
        g.fillInData(traces, [0.0], 4.0, ct, z, z + 4.0, 40.0)  # a 10-sample taper on 8 samples
    g.close()
