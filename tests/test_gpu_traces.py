"""CommonData::FillInData on the device (cg_grid_fill_from_traces): SEG-Y traces onto the grid --
the seismic branch's guard-zone taper, the other grids' zero trimming and trend, the ten-fold
resampling through the spectrum, the cell interpolation -- against the oracle on the same traces
and columns."""
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


def test_segy_file_onto_the_grid(ctx, tmp_path):
    """A SEG-Y file written from a grid (FFTGrid::writeSegyFile on the device), read back and
    resampled onto a grid over the same box (CommonData::FillInData): the device grid equals the
    oracle's columns from the same traces."""
    import os
    from crava_amd import seismic as SE
    rng = np.random.default_rng(11)
    nx, ny, nz, nxp, nyp, nzp = 10, 8, 50, 12, 10, 64
    src = C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, C.DATA)
    kk = np.arange(nzp, dtype=np.float32)[:, None, None]
    field = (np.sin(0.25 * kk + rng.uniform(0, 3, (1, nyp, 2 * (nxp // 2 + 1))))).astype(np.float32)
    src.upload_real(field)
    top = 1000.0 + rng.uniform(0.0, 30.0, (ny, nx))
    geo = dict(x0=4000.0, y0=7000.0, dx=25.0, dy=25.0, angle=0.2, il0=100.2, xl0=200.3,
               il_step_x=np.cos(0.2) / 25.0, il_step_y=np.sin(0.2) / 25.0, xl_step_x=-np.sin(0.2) / 25.0,
               xl_step_y=np.cos(0.2) / 25.0, dz=4.0, z0=900.0, zmax=1400.0, top=top, rel_thick=np.ones((ny, nx)))
    path = os.path.join(str(tmp_path), "in.segy")
    src.writeSegyFile(path, geo)
    transform = dict(in_line0=geo["il0"], cross_line0=geo["xl0"], x0=geo["x0"], y0=geo["y0"],
                     il_step_x=geo["il_step_x"], il_step_y=geo["il_step_y"], xl_step_x=geo["xl_step_x"],
                     xl_step_y=geo["xl_step_y"])
    simbox = dict(x0=geo["x0"], y0=geo["y0"], dx=25.0, dy=25.0, angle=0.2, top=top, bot=top + nz * 4.0)
    dst = C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, C.DATA)
    tr, ct = SE.fill_in_data_from_segy(dst, path, transform, simbox, 20.0)
    got = dst.download_real(padded=True)
    ct2, z0g, dzg, z0d = SE.resolve_columns(tr, SE.segy_lattice(tr), transform, simbox, nx, ny, nz, nxp, nyp)
    want, _ = O.fill_from_traces(list(tr["samples"]), z0d, tr["dz"], ct2, z0g, dzg, nz, nzp, 20.0)
    err = float(np.max(np.abs(got - want))) / float(np.max(np.abs(want)))
    print("SEG-Y onto the grid: %d traces of %d samples, %d columns filled; max |device - oracle| / max %.2e"
          % (len(tr["il"]), tr["samples"].shape[1], int((ct >= 0).sum()), err), flush=True)
    assert err < 2e-6 and int((ct >= 0).sum()) == nxp * nyp
    src.close()
    dst.close()


def _param_case(seed, nx, ny, nxp, nyp, nmin, nmax):
    """Background-like traces (a level, a trend, noise) with leading zeros on some, interior zeros
    on others, a few all zero; columns with dead traces (-2) and, in the padding only, columns
    outside the data (-1)."""
    rng = np.random.default_rng(seed)
    ntr = nx * ny
    traces = []
    for q in range(ntr):
        n = int(rng.integers(nmin, nmax + 1))
        tr = (2500.0 + rng.uniform(-2, 4) * np.arange(n) + 30.0 * rng.standard_normal(n)).astype(np.float32)
        kind = q % 7
        if kind == 1:
            tr[:int(rng.integers(1, n // 3))] = 0.0   # zeros above a top
        elif kind == 2:
            tr[0] = 0.0
            tr[n // 2:n // 2 + 3] = 0.0               # every zero goes when the first one is zero
        elif kind == 3:
            tr[n // 3:n // 3 + 2] = 0.0               # kept: the first sample is not zero
        elif kind == 4 and q % 2 == 0:
            tr[:] = 0.0                               # nothing left: a dead trace
        traces.append(tr)
    z0_data = rng.uniform(0.0, 40.0, ntr).astype(np.float32)
    rnxp = 2 * (nxp // 2 + 1)
    ct = np.full((nyp, rnxp), -1, dtype=np.int32)
    for j in range(nyp):
        for i in range(nxp):
            ri, rj = O.fill_number(i, nx, nxp), O.fill_number(j, ny, nyp)
            u = rng.uniform()
            ct[j, i] = -2 if u < 0.05 else (-1 if (u < 0.1 and (i >= nx or j >= ny)) else rj * nx + ri)
    z0_grid = (60.0 + rng.uniform(-8.0, 8.0, (nyp, rnxp))).astype(np.float32)
    dz_grid = rng.uniform(3.0, 5.0, (nyp, rnxp)).astype(np.float32)
    return traces, z0_data, ct, z0_grid, dz_grid


@pytest.mark.parametrize("dims,lengths", [((20, 12, 30, 24, 16, 40), (60, 80)),
                                          ((10, 7, 25, 10, 7, 25), (41, 45)),   # no padding, as the reference's
                                          ((6, 5, 60, 8, 6, 72), (900, 1000))])
def test_parameter_fill_from_traces_matches_oracle(ctx, dims, lengths):
    """FillInData's branch for other grids (background and parameter cubes): zero trimming, the
    trend removed before and added after the resampling, RMISSING in dead columns."""
    nx, ny, nz, nxp, nyp, nzp = dims
    print("parameter fill from traces", dims, lengths, flush=True)
    traces, z0d, ct, z0g, dzg = _param_case(8, nx, ny, nxp, nyp, *lengths)
    dz_data = 4.0
    want, nt = O.fill_from_traces(traces, z0d, dz_data, ct, z0g, dzg, nz, nzp, 0.0, parameter=True)
    g = C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, C.PARAMETER)
    g.fillInData(traces, z0d, dz_data, ct, z0g, dzg, 0.0, C.PARAMETER)
    got = g.download_real(padded=True)
    dead = want == np.float32(O.RMISSING)
    assert np.array_equal(got == np.float32(O.RMISSING), dead)
    err = float(np.max(np.abs(got[~dead] - want[~dead]))) / float(np.max(np.abs(want[~dead])))
    print("nt %d: %d dead columns; max |device - oracle| / max |oracle| %.2e"
          % (nt, int(dead[0].sum()), err), flush=True)
    assert err < 2e-6 and dead[0].sum() > 0
    assert np.all(got[:, ct == -1] == 0.0)
    g.close()


def test_parameter_fill_refuses_an_uncovered_box(ctx):
    # a box column outside the data: the reference stops and reports the grid as not covering it
    g = C.FFTGrid(ctx, 6, 5, 10, 6, 5, 10, C.PARAMETER)
    traces = [np.arange(1, 21, dtype=np.float32)]
    ct = np.zeros((5, 8), np.int32)
    ct[2, 3] = -1
    z = np.zeros((5, 8), np.float32)
    with pytest.raises(C.CravaError):
        g.fillInData(traces, [0.0], 4.0, ct, z, z + 4.0, 0.0, C.PARAMETER)
    ct[2, 3] = -2  # a dead trace is allowed: RMISSING
    g.fillInData(traces, [0.0], 4.0, ct, z, z + 4.0, 0.0, C.PARAMETER)
    got = g.download_real(padded=True)
    assert np.all(got[:, 2, 3] == np.float32(O.RMISSING)) and np.all(got[:, 2, 2] != np.float32(O.RMISSING))
    g.close()


def test_background_segy_file_onto_the_box(ctx, tmp_path):
    """A background cube written to SEG-Y (zeros above and below the box, as FFTGrid::writeSegyFile
    writes them), read back, cut to the volume (SegY::ReadAllTraces with the box and twice a guard
    zone of padding) and resampled onto the unpadded box as CommonData's background reading does
    (FillInData with a PARAMETER grid, is_seismic false): the device grid equals the oracle's
    columns, and stays within a few per cent of the cube it came from."""
    import os
    from crava_amd import seismic as SE
    rng = np.random.default_rng(12)
    nx, ny, nz = 10, 8, 50
    kk = np.arange(nz, dtype=np.float32)[:, None, None]
    field = (2500.0 + 6.0 * kk + 20.0 * np.sin(0.15 * kk + rng.uniform(0, 3, (1, ny, nx)))).astype(np.float32)
    src = C.FFTGrid(ctx, nx, ny, nz, nx, ny, nz, C.PARAMETER)
    rnx = 2 * (nx // 2 + 1)
    full = np.zeros((nz, ny, rnx), np.float32)
    full[:, :, :nx] = field
    src.upload_real(full)
    top = 1000.0 + rng.uniform(0.0, 30.0, (ny, nx))
    geo = dict(x0=4000.0, y0=7000.0, dx=25.0, dy=25.0, angle=0.2, il0=100.2, xl0=200.3,
               il_step_x=np.cos(0.2) / 25.0, il_step_y=np.sin(0.2) / 25.0, xl_step_x=-np.sin(0.2) / 25.0,
               xl_step_y=np.cos(0.2) / 25.0, dz=4.0, z0=900.0, zmax=1400.0, top=top, rel_thick=np.ones((ny, nx)))
    path = os.path.join(str(tmp_path), "bg.segy")
    src.writeSegyFile(path, geo)
    tr = SE.read_segy(path)
    assert np.any(tr["samples"][:, 0] == 0.0)
    transform = dict(in_line0=geo["il0"], cross_line0=geo["xl0"], x0=geo["x0"], y0=geo["y0"],
                     il_step_x=geo["il_step_x"], il_step_y=geo["il_step_y"], xl_step_x=geo["xl_step_x"],
                     xl_step_y=geo["xl_step_y"])
    simbox = dict(x0=geo["x0"], y0=geo["y0"], dx=25.0, dy=25.0, angle=0.2, top=top, bot=top + nz * 4.0)
    # the volume's surfaces at each trace: the box cell the trace was written for
    lat = SE.segy_lattice(tr)
    cell = {}
    cr, sr = np.cos(0.2), np.sin(0.2)
    for j in range(ny):
        for i in range(nx):
            rx, ry = (i + 0.5) * 25.0, (j + 0.5) * 25.0
            xf = float(np.float32(rx * cr - ry * sr + geo["x0"]))
            yf = float(np.float32(rx * sr + ry * cr + geo["y0"]))
            il = int(0.5 + geo["il0"] + (xf - geo["x0"]) * geo["il_step_x"] + (yf - geo["y0"]) * geo["il_step_y"])
            xl = int(0.5 + geo["xl0"] + (xf - geo["x0"]) * geo["xl_step_x"] + (yf - geo["y0"]) * geo["xl_step_y"])
            cell[(il, xl)] = (i, j)
    ij = [cell[(int(a), int(b))] for a, b in zip(tr["il"], tr["xl"])]
    z_top = np.array([top[j, i] for i, j in ij])
    window = dict(z_top=z_top, z_bot=z_top + nz * 4.0, z_pad=2 * 20.0)
    dst = C.FFTGrid(ctx, nx, ny, nz, nx, ny, nz, C.PARAMETER)
    tr2, ct = SE.fill_in_data_from_segy(dst, path, transform, simbox, 0.0, grid_type=C.PARAMETER, window=window)
    got = dst.download_real(padded=True)
    ct2, z0g, dzg, z0d = SE.resolve_columns(tr2, lat, transform, simbox, nx, ny, nz, nx, ny, C.PARAMETER)
    samples, z0k, ctk = SE.kept_traces(tr2, ct2, z0d)
    want, nt = O.fill_from_traces(samples, z0k, tr2["dz"], ctk, z0g, dzg, nz, nz, 0.0, parameter=True)
    err = float(np.max(np.abs(got - want))) / float(np.max(np.abs(want)))
    box = got[:, :, :nx]
    dev = float(np.max(np.abs(box - field))) / float(np.max(np.abs(field)))
    print("background SEG-Y onto the box: %d traces cut to %d..%d samples (nt %d); max |device - oracle| / max "
          "%.2e; max |grid - written cube| / max %.2e"
          % (len(samples), min(len(s) for s in samples), max(len(s) for s in samples), nt, err, dev), flush=True)
    assert err < 2e-6 and int((ct >= 0).sum()) == ny * nx and dev < 0.02
    src.close()
    dst.close()


@pytest.mark.parametrize("grid_type", [C.DATA, C.PARAMETER])
def test_storm_grid_onto_the_box(ctx, tmp_path, grid_type):
    """FillInData with a STORM grid as the source (is_storm): a STORM cube on its own rotated,
    coarser lattice and its own time span resampled onto a simbox inside it -- seismic (tapered)
    and parameter grids (zeros trimmed at the top of some columns without moving their start,
    the trend, RMISSING for a box cell with a missing top); the device grid equals the oracle's
    columns from the same STORM columns."""
    import os
    from crava_amd import seismic as SE
    rng = np.random.default_rng(13 + grid_type)
    nx, ny, nz, nxp, nyp, nzp = 12, 10, 40, 16, 12, 48
    ni, nj, nk = 9, 8, 70
    kk = np.arange(nk, dtype=np.float32)[:, None, None]
    if grid_type == C.DATA:
        vals = np.sin(0.3 * kk + rng.uniform(0, 3, (1, nj, ni))) + 0.2 * rng.standard_normal((nk, nj, ni))
    else:
        vals = 2400.0 + 5.0 * kk + 30.0 * rng.standard_normal((nk, nj, ni))
        vals[:rng.integers(1, 6), 2, 3] = 0.0
        vals[:4, 5, 1] = 0.0
    vals = vals.astype(np.float32)
    path = os.path.join(str(tmp_path), "in.storm")
    ang = 25.0
    head = ("storm_petro_binary\n0 2 -999.000000\nFFTGrid\n%f %f %f %f %s %s 0.0 0.0\n%f %f\n\n%d %d %d\n"
            % (3900.0, 420.0, 6950.0, 380.0, 950, 1300, 350.0, ang, ni, nj, nk))
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(np.ascontiguousarray(vals, dtype=">f4").tobytes())
        f.write(b"0\n")
    top = 1000.0 + rng.uniform(0.0, 30.0, (ny, nx))
    top[4, 5] = O.RMISSING
    simbox = dict(x0=4000.0, y0=7000.0, dx=20.0, dy=20.0, angle=np.radians(28.0), top=top, bot=top + nz * 4.0)
    g = C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, grid_type)
    st, ct = SE.fill_in_data_from_storm(g, path, simbox, 20.0, grid_type)
    got = g.download_real(padded=True)
    traces, z0d, dzd, ct2, z0g, dzg = SE.storm_columns(st, simbox, nx, ny, nz, nxp, nyp, grid_type)
    want, nt = O.fill_from_traces(traces, z0d, dzd, ct2, z0g, dzg, nz, nzp, 20.0,
                                  parameter=grid_type == C.PARAMETER, trim_shift=False)
    miss = want == np.float32(O.RMISSING)
    assert np.array_equal(got == np.float32(O.RMISSING), miss)
    err = float(np.max(np.abs(got[~miss] - want[~miss]))) / float(np.max(np.abs(want[~miss])))
    print("STORM %dx%dx%d onto %dx%dx%d (%s): %d STORM columns, %d outside, %d dead; max |device - oracle| / max %.2e"
          % (ni, nj, nk, nxp, nyp, nzp, "DATA" if grid_type == C.DATA else "PARAMETER", len(traces),
             int((ct == -1).sum()), int((ct == -2).sum()), err), flush=True)
    assert err < 2e-6 and len(traces) > 20 and int((ct == -2).sum()) > 0
    if grid_type == C.PARAMETER:
        assert not np.any(ct[:ny, :nx] == -1)  # the STORM volume covers the box
    g.close()
