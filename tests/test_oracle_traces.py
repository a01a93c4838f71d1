"""The oracle of CommonData::FillInData's seismic branch (SEG-Y traces onto the padded grid):
the resampling against the reference's own FFTW 2.1.2 calls and against exact band-limited
interpolation, the guard-zone taper and the cell interpolation against properties that fix
their indices."""
import math

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("n_data,nt", [(7, 8), (11, 12), (13, 15), (96, 96), (250, 252), (301, 315)])
def test_resample_trace_matches_the_reference_fftw(n_data, nt):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    import ctypes
    rng = np.random.default_rng(n_data)
    data = rng.standard_normal(n_data).astype(np.float32)
    rmt = 2 * (10 * nt // 2 + 1)
    ref = np.zeros(rmt, dtype=np.float32)
    R = O.ref()
    R.ref_resample_trace.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    R.ref_resample_trace(n_data, data.ctypes.data, nt, ref.ctypes.data)
    got = O.resample_trace(data, nt)
    assert got.shape == (rmt,)
    # float FFTW against the double restatement
    assert np.max(np.abs(got - ref)) < 2e-6 * max(1.0, np.max(np.abs(ref)))
    assert np.all(ref[10 * nt:] == 0.0) and np.all(got[10 * nt:] == 0.0)


def test_resample_trace_interpolates_band_limited_signals_exactly():
    # a periodic signal below the Nyquist frequency of nt samples is reproduced at every fine
    # sample l / 10; the Nyquist bin of an even nt enters the fine spectrum twice (as in the
    # reference), so the test signal keeps below it
    nt = 40
    n = np.arange(nt)
    freqs, amps, phases = [0, 3, 7, 19], [0.5, 1.0, -0.3, 0.25], [0.0, 0.4, 1.1, -0.7]
    x = sum(a * np.cos(2 * np.pi * f * n / nt + ph) for f, a, ph in zip(freqs, amps, phases))
    fine = O.resample_trace(x, nt)
    l = np.arange(10 * nt)
    want = sum(a * np.cos(2 * np.pi * f * l / (10 * nt) + ph) for f, a, ph in zip(freqs, amps, phases))
    assert np.max(np.abs(fine[:10 * nt] - want)) < 1e-12


def test_guard_zone_taper_starts_at_zero_ends_at_zero_and_meets_one():
    dz, sl = 4.0, 40.0
    ns = int(math.floor(sl / dz))
    t = O.smooth_trace_in_guard_zone(np.ones(60, np.float32), dz, sl)
    assert t[0] == 0.0 and abs(t[-1]) < 1e-30  # sin^2(0) and cos^2(pi/2) (of a double pi/2)
    assert np.all(t[ns:60 - ns] == 1.0)  # untouched between the two zones
    assert np.all(np.diff(t[:ns + 1]) > 0) and np.all(np.diff(t[60 - ns - 1:]) < 0)
    # complementary zones: the start factor sin^2(pi/2 k/ns) at sample k and the end factor
    # cos^2(pi/2 (q + 1)/ns) at the end zone's sample q meet at q = k - 1, where they sum to one;
    # an index off by one on either side breaks it
    s = t[:ns].astype(np.float64)
    e = t[60 - ns:].astype(np.float64)
    assert np.allclose(s[1:] + e[:-1], 1.0, atol=1e-6)


def test_cell_interpolation_of_a_linear_fine_trace_and_its_edges():
    nz, nzp, n_fine = 20, 32, 400
    fine = (0.5 + 0.01 * np.arange(n_fine)).astype(np.float32)
    z0_grid, dz_grid, z0_data, dz_fine = 13.0, 4.0, 5.0, 0.4
    g = O.interpolate_grid_values(fine, z0_grid, dz_grid, z0_data, dz_fine, nz, nzp)
    for k in range(nzp):
        refk = k if k < (nz + nzp) // 2 else k - nzp
        dl = (z0_grid - z0_data + refk * dz_grid) / dz_fine
        if dl < -1:
            assert g[k] == 0.0  # above the trace
        elif dl < 0:
            assert g[k] == fine[0]  # the first fine sample where only l2 = 0 is inside
        else:
            assert abs(g[k] - (0.5 + 0.01 * dl)) < 2e-5, k  # linear data is reproduced
    # cells below the last fine sample are zero
    g2 = O.interpolate_grid_values(fine[:50], z0_grid, dz_grid, z0_data, dz_fine, nz, nzp)
    assert g2[10] == 0.0


def test_fill_from_traces_columns():
    tr = [np.sin(np.arange(40) * 0.3), np.cos(np.arange(35) * 0.2)]
    ct = np.array([[0, 1, -1, 0], [1, 1, 0, -1]])
    z0g = np.full((2, 4), 30.0, np.float32)
    dzg = np.full((2, 4), 4.0, np.float32)
    g, nt = O.fill_from_traces(tr, np.array([10.0, 12.0], np.float32), 4.0, ct, z0g, dzg, 12, 16, 20.0)
    assert nt == 40 and g.shape == (16, 2, 4)
    assert np.all(g[:, 0, 2] == 0) and np.all(g[:, 1, 3] == 0)  # columns without a trace
    assert np.array_equal(g[:, 0, 0], g[:, 0, 3]) and np.array_equal(g[:, 0, 1], g[:, 1, 0])
