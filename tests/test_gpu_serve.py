"""Continuous batching on the GPU: requests that arrive together are batched over concurrent
lanes, and each result equals the result of the same request served alone."""
import numpy as np
import pytest

from crava_amd import serve
from crava_amd import synthetic as S

pytestmark = pytest.mark.gpu

# Small requests of different sizes, angle counts and realisation counts (odd and even nxp).
SPECS = [S.spec(20, 16, 24, (0.0, 15.0, 30.0), 25.0, seed=101, realisations=2),
         S.spec(15, 21, 20, (0.0, 10.0, 20.0, 30.0), 30.0, seed=102, realisations=0),
         S.spec(24, 12, 30, (0.0, 15.0, 30.0, 45.0, 20.0), 25.0, seed=103, realisations=1),
         S.spec(18, 18, 18, (0.0, 15.0, 30.0), 35.0, seed=104, realisations=3),
         S.spec(30, 10, 16, (0.0, 9.0, 18.0, 27.0, 36.0, 45.0), 25.0, seed=105, realisations=1)]

# Inputs are generated on the device from each request's seed (the S/N scaling's reduction runs
# in a fixed order) and a batch computes every inversion with the same kernels and descriptors,
# so each result is the same alone and in the mix bit for bit.


def _serve(specs, lanes, max_batch):
    ex = serve.GpuExecutor(lanes=lanes)
    # all requests are ingested before the first dispatch, so the lanes receive mixed batches
    srv = serve.Server(ex, hbm_budget=8e9, lanes=lanes, max_batch=max_batch, ingest_per_step=len(specs),
                       asynchronous_ingest=False)
    out = {}

    def on_done(r):
        p = r.work
        out[r.spec["seed"]] = (np.stack([g.download_real() for g in p.mean_w]),
                               np.stack([g.download_complex() for g in p.cov]),
                               [np.stack([g.download_real() for g in o]) for o in p.out[:p.R]])

    for s in specs:
        srv.submit(s)
    srv.drain(on_done)
    ex.close()
    return out, srv


def test_results_alone_and_in_the_mix_agree():
    mix, srv = _serve(SPECS, lanes=2, max_batch=3)
    assert len(mix) == len(SPECS) and srv.in_use == 0
    assert max(srv.batch_sizes) > 1
    for s in SPECS:
        alone, _ = _serve([s], lanes=1, max_batch=1)
        (ma, ca, ra), (mm, cm, rm) = alone[s["seed"]], mix[s["seed"]]
        assert np.array_equal(mm, ma) and np.array_equal(cm, ca)
        assert len(ra) == len(rm) == s["realisations"]
        for x, y in zip(rm, ra):
            assert np.array_equal(x, y)
