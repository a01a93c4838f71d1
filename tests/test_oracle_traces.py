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


# ---------------------------------------------------------------- the branch for other grids
def _param_fill(traces, z0d, dz, cols, z0g, dzg, nz, nzp):
    ct = np.asarray(cols, dtype=np.int32)[None, :]
    z0g = np.asarray(z0g, dtype=np.float32)[None, :]
    dzg = np.asarray(dzg, dtype=np.float32)[None, :]
    return O.fill_from_traces(traces, np.asarray(z0d, np.float32), dz, ct, z0g, dzg, nz, nzp, 0.0, parameter=True)


def test_parameter_fill_of_a_straight_line_is_the_line():
    # a trace on a line leaves nothing after the detrending: the column is the line at the cell
    # times, its first value above the trace and its last below it
    a, b, n, dz = 2500.0, 1.25, 60, 4.0
    tr = (a + b * np.arange(n)).astype(np.float32)
    nz = nzp = 100
    z0d, z0g, dzg = 1000.0, 1002.0 - 40.0, 2.9
    g, _ = _param_fill([tr], [z0d], dz, [0], [z0g], [dzg], nz, nzp)
    col = g[:, 0, 0].astype(np.float64)
    t = z0g + np.arange(nzp) * dzg
    inside = (t >= z0d) & (t <= z0d + (n - 1) * dz)
    want = a + b * (t - z0d) / dz
    assert np.max(np.abs(col[inside] - want[inside])) < 2e-3  # float trend, values ~2500
    assert np.all(col[t < z0d - dz] == np.float32(a)) and np.all(col[t > z0d + (n - 1) * dz + dz] == tr[-1])


def test_parameter_fill_reproduces_the_samples_at_their_times():
    # odd nt (45): the ten-fold resampling of the detrended trace is exact at the coarse samples,
    # so cells at the sample times return the trace; dz 5 ms makes the fine step 0.5 ms exact
    rng = np.random.default_rng(3)
    n, dz = 45, 5.0
    tr = (1800.0 + 3.0 * np.arange(n) + 40.0 * rng.standard_normal(n)).astype(np.float32)
    nz, nzp = 40, 64
    g, nt = _param_fill([tr], [700.0], dz, [0], [700.0], [dz], nz, nzp)
    assert nt == 45
    col = g[:, 0, 0]
    ks = np.array([k for k in range(nzp) if 0 <= O.z_simbox_index(k, nz, nzp) < n])
    assert np.max(np.abs(col[ks] - tr[[O.z_simbox_index(k, nz, nzp) for k in ks]])) < 2e-6 * 2000
    # the padding's negative cell indices lie above the trace: the first sample exactly
    neg = [k for k in range(nzp) if O.z_simbox_index(k, nz, nzp) < -1]
    assert len(neg) > 0 and np.all(col[neg] == tr[0])


def test_parameter_fill_trims_zeros_when_the_first_sample_is_zero():
    rng = np.random.default_rng(4)
    x = (1000.0 + 50.0 * rng.standard_normal(30)).astype(np.float32)
    dz, z0 = 4.0, np.float32(812.0)
    lead = np.concatenate([np.zeros(3, np.float32), x[:10], np.zeros(2, np.float32), x[10:]])
    z0s = z0
    for _ in range(3):  # the reference adds dz once per leading zero
        z0s = np.float32(z0s + np.float32(dz))
    inner = np.concatenate([x[:10], np.zeros(2, np.float32), x[10:]])  # first sample nonzero: zeros kept
    traces = [lead, x, inner, np.zeros(12, np.float32)]
    nz = nzp = 48
    cols = [0, 1, 2, 3, -2, -1]
    z0g = [820.0, 820.0, 820.0, 820.0, 820.0, 820.0]
    dzg = [3.3] * 6
    g, _ = _param_fill(traces, [z0, z0s, z0, z0], dz, cols, z0g, dzg, nz, nzp)
    c = g[:, 0, :]
    assert np.array_equal(c[:, 0], c[:, 1])          # all zeros dropped, z0 moved three samples down
    assert not np.array_equal(c[:, 2], c[:, 1])      # interior zeros stay when the first sample is not zero
    assert np.all(c[:, 3] == np.float32(O.RMISSING))  # no nonzero sample: dead
    assert np.all(c[:, 4] == np.float32(O.RMISSING)) and np.all(c[:, 5] == 0.0)
