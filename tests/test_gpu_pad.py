"""The path's inputs on the device: padding fills (cg_grid_fill_padded,
cg_grid_interpolate_seismic, cg_grid_extrapolate_seismic), prior and error correlation grids
(cg_grid_fill_param_corr, cg_grid_fill_err_corr) and CRAVA binary grid files
(cg_grid_write_crava, cg_grid_read_crava) vs the oracle, bit for bit."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _grid(ctx, dims, ct=C.PARAMETER):
    nx, ny, nz, nxp, nyp, nzp = dims
    return C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, ct)


def _device_cube(ctx, cube):
    """A scratch grid whose raw memory holds the cube (nz*ny*nx floats) at its start."""
    nz, ny, nx = cube.shape
    s = C.FFTGrid(ctx, nx, ny, nz, nx, ny, nz, C.PARAMETER)
    flat = np.zeros(nz * ny * s.rnxp, dtype=np.float32)
    flat[:cube.size] = cube.ravel()
    s.upload_real(flat.reshape(nz, ny, s.rnxp))
    return s


# odd and even nxp, no padding on an axis, one trace wide, one sample deep
SHAPES = [(7, 5, 6, 12, 8, 10), (9, 4, 5, 9, 6, 8), (6, 6, 3, 10, 6, 3), (1, 3, 4, 4, 5, 7), (33, 1, 20, 45, 1, 24),
          (20, 17, 1, 25, 18, 2)]


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("where", ["host", "device"])
def test_fill_padded_matches_oracle(ctx, dims, where):
    nx, ny, nz, nxp, nyp, nzp = dims
    print("fill_padded", dims, where, flush=True)
    cube = np.random.default_rng(nx * 31 + nz).normal(size=(nz, ny, nx)).astype(np.float32)
    g = _grid(ctx, dims, C.DATA)
    g.fillInConstant(7.0)
    if where == "host":
        g.fill_padded(cube)
    else:
        s = _device_cube(ctx, cube)
        g.fill_padded(s.device_ptr())
        ctx.sync()
        s.close()
    want = O.fill_padded(cube, nxp, nyp, nzp)
    got = g.download_real(padded=True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert g.getIsTransformed() == 0
    g.close()


def test_fill_padded_single_cube_size(ctx):
    """The single-cube workload's background cube, 512 x 512 x 256 into 576 x 576 x 320."""
    dims = (512, 512, 256, 576, 576, 320)
    print("fill_padded", dims, flush=True)
    cube = np.random.default_rng(5).normal(size=(256, 512, 512)).astype(np.float32)
    g = _grid(ctx, dims)
    g.fill_padded(cube)
    got = g.download_real(padded=True)
    g.close()
    want = O.fill_padded(cube, 576, 576, 320)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _bad_traces(rv, nx, ny, frac, seed, block=None):
    rng = np.random.default_rng(seed)
    bad = rng.random((ny, nx)) < frac
    if block:
        i0, i1, j0, j1 = block
        bad[j0:j1, i0:i1] = True
    rv[:, :ny, :nx][:, bad] = 0.0
    return int(bad.sum())


def _seismic(dims, seed):
    nx, ny, nz, nxp, nyp, nzp = dims
    rng = np.random.default_rng(seed)
    rv = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), dtype=np.float32)
    rv[:, :, :nxp] = rng.normal(size=(nzp, nyp, nxp)).astype(np.float32)
    return rv


@pytest.mark.parametrize("dims,frac,block", [((40, 30, 50, 48, 36, 64), 0.15, (5, 12, 7, 10)),
                                             ((6, 5, 4, 8, 7, 6), 0.3, None),
                                             ((33, 1, 20, 45, 1, 24), 0.2, None),
                                             ((64, 48, 40, 80, 60, 50), 0.05, (0, 64, 0, 3)),
                                             ((25, 20, 12, 32, 27, 16), 0.6, (20, 25, 0, 20))])
def test_interpolate_seismic_matches_oracle(ctx, dims, frac, block):
    """Random bad traces plus blocks of them (a band along an edge, a band of columns): the
    device's interpolation and extrapolation equal the oracle's sweep bit for bit."""
    nx, ny, nz, nxp, nyp, nzp = dims
    print("interpolate_seismic", dims, frac, block, flush=True)
    rv = _seismic(dims, 11)
    nbad_set = _bad_traces(rv, nx, ny, frac, 12, block)
    want, nbad, rect = O.interpolate_seismic(rv, nx, ny, nz, nxp, 0.01)
    assert nbad >= nbad_set
    g = _grid(ctx, dims, C.DATA)
    g.upload_real(rv)
    got_nbad, got_rect = g.interpolateSeismic(0.01)
    got = g.download_real(padded=True)
    g.close()
    assert (got_nbad, got_rect) == (nbad, rect)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_extrapolate_seismic_inner_rectangle(ctx):
    dims = (30, 24, 20, 40, 32, 27)
    nx, ny, nz, nxp, nyp, nzp = dims
    rv = _seismic(dims, 21)
    rect = (4, 25, 3, 20)
    want = O.extrapolate_seismic(rv, nx, ny, nz, nxp, rect)
    g = _grid(ctx, dims, C.DATA)
    g.upload_real(rv)
    g.extrapolateSeismic(*rect)
    got = g.download_real(padded=True)
    g.close()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_padding_fill_errors(ctx):
    dims = (6, 5, 4, 8, 7, 6)
    g = _grid(ctx, dims, C.DATA)
    g.upload_real(np.zeros((6, 7, 10), dtype=np.float32))
    with pytest.raises(C.CravaError, match="no trace above"):
        g.interpolateSeismic(0.01)  # every trace has zero energy
    for rect in ((3, 2, 0, 4), (0, 6, 0, 4), (-1, 3, 0, 4), (0, 5, 2, 1)):
        with pytest.raises(C.CravaError, match="rectangle"):
            g.extrapolateSeismic(*rect)
    g.setType(C.PARAMETER)
    with pytest.raises(C.CravaError, match="DATA"):
        g.interpolateSeismic(0.01)
    with pytest.raises(ValueError):
        g.fill_padded(np.zeros((4, 5, 5), dtype=np.float32))
    g.close()


GEO = dict(x0=451000.5, y0=6780000.25, dx=25.0, dy=12.5, nx=30, ny=24, il0=1001.0, xl0=2002.0, il_step_x=0.5,
           il_step_y=-0.25, xl_step_x=0.125, xl_step_y=1.5, angle=0.3)


@pytest.mark.parametrize("dims", [(30, 24, 20, 40, 32, 27), (9, 7, 5, 9, 7, 5)])
def test_crava_file_matches_oracle(ctx, tmp_path, dims):
    """writeCravaFile from the device gives the oracle's bytes; readCravaFile of the oracle's file
    restores the grid bit for bit and returns the geometry."""
    nx, ny, nz, nxp, nyp, nzp = dims
    print("crava file", dims, flush=True)
    rv = _seismic(dims, 31)
    g = _grid(ctx, dims)
    g.upload_real(rv)
    g.writeCravaFile(str(tmp_path / "dev"), GEO)
    O.write_crava(str(tmp_path / "orc.crava"), rv, GEO)
    assert open(tmp_path / "dev.crava", "rb").read() == open(tmp_path / "orc.crava", "rb").read()
    h = _grid(ctx, dims)
    geo = h.readCravaFile(str(tmp_path / "orc.crava"))
    assert geo == GEO
    got = h.download_real(padded=True)
    assert np.array_equal(got.view(np.uint32), rv.view(np.uint32))
    g.close()
    h.close()


def test_crava_file_errors(ctx, tmp_path):
    dims = (9, 7, 5, 9, 7, 5)
    rv = _seismic(dims, 32)
    O.write_crava(str(tmp_path / "a.crava"), rv, GEO)
    other = _grid(ctx, (9, 7, 5, 10, 7, 5))
    with pytest.raises(C.CravaError, match="Grid dimension is wrong"):
        other.readCravaFile(str(tmp_path / "a.crava"))
    g = _grid(ctx, dims)
    with pytest.raises(C.CravaError, match="cannot open"):
        g.readCravaFile(str(tmp_path / "missing.crava"))
    raw = open(tmp_path / "a.crava", "rb").read()
    open(tmp_path / "short.crava", "wb").write(raw[:-4])
    with pytest.raises(C.CravaError, match="ends inside the grid"):
        g.readCravaFile(str(tmp_path / "short.crava"))
    open(tmp_path / "head.crava", "wb").write(raw[:60])
    with pytest.raises(C.CravaError, match="ends inside the header"):
        g.readCravaFile(str(tmp_path / "head.crava"))
    other.close()
    g.close()


@pytest.mark.parametrize("dims,gi,gj,scale", [((9, 6, 10), 0.0, 0.0, 1.0), ((9, 6, 10), 2.0, 0.0, 0.7),
                                              ((16, 12, 20), 0.25, -0.5, 3.0), ((15, 7, 9), -1.75, 0.3, 1.0),
                                              ((576, 576, 320), 0.1, 0.05, 0.01)])
def test_correlation_fills_match_oracle(ctx, dims, gi, gj, scale):
    """fillInParamCorr (+ multiplyByScalar) and fillInErrCorr with no, integer, fractional and
    negative gradients, bit for bit; the last case at the single-cube padded size."""
    nxp, nyp, nzp = dims
    print("correlation fills", dims, gi, gj, flush=True)
    rng = np.random.default_rng(nxp + nzp)
    cxy = rng.random((nyp, nxp)).astype(np.float32)
    ct = rng.random(nzp).astype(np.float32)
    g = C.FFTGrid(ctx, nxp, nyp, nzp, nxp, nyp, nzp, C.COVARIANCE)
    g.fillInParamCorr(cxy, ct, gi, gj, scale)
    got = g.download_real(padded=True)
    want = O.fill_param_corr(cxy, ct, gi, gj, scale)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    g.fillInErrCorr(cxy, gi, gj)
    got = g.download_real(padded=True)
    want = O.fill_err_corr(cxy, nzp, gi, gj)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    g.close()


@pytest.mark.parametrize("dims,wavelet", [((20, 16, 30, 24, 18, 40), "ricker"), ((13, 9, 25, 15, 10, 32), "random"),
                                          ((64, 48, 100, 72, 54, 128), "ricker")])
def test_synthetic_seismic_matches_oracle(ctx, dims, wavelet):
    """ComputeSyntSeismic on the device (trace taper, 3D transforms, the backward-difference and
    wavelet operator, inverse transforms) against the oracle's trace-by-trace computation."""
    from crava_amd import model as M
    nx, ny, nz, nxp, nyp, nzp = dims
    print("synthetic seismic", dims, wavelet, flush=True)
    rng = np.random.default_rng(nz)
    vp, vs, rho = (rng.normal(size=(nz, ny, nx)).astype(np.float32) for _ in range(3))
    angles = (0.0, 15.0, 30.0)
    A = M.aki_richards(angles, 0.5)
    if wavelet == "ricker":
        w = np.stack([M.wavelet_fft_order(nzp, 4.0, 30.0, scale=1.0 + 0.1 * l) for l in range(3)])
    else:
        w = rng.normal(size=(3, nzp)).astype(np.float32)
    got = M.synthetic_seismic(ctx, vp, vs, rho, w, A, nxp, nyp, nzp)
    want = O.synthetic_seismic(vp, vs, rho, w, A, nzp)
    for g, o in zip(got, want):
        assert rel_l2(g, o) < 1e-5


def test_real_accessors(ctx):
    """getRealValue / getRealTrace (getOrigin and createPostCov00 read through them): box values,
    RMISSING outside the box, and the padding part of a padded trace reads RMISSING as well."""
    dims = (7, 5, 6, 10, 8, 9)
    nx, ny, nz, nxp, nyp, nzp = dims
    rv = _seismic(dims, 41)
    g = _grid(ctx, dims, C.COVARIANCE)
    g.upload_real(rv)
    assert g.getRealValue(0, 0, 0) == rv[0, 0, 0]
    assert g.getRealValue(3, 2, 4) == rv[4, 2, 3]
    assert g.getRealValue(8, 2, 4) == -99999.0 and g.getRealValue(8, 2, 4, extSimbox=True) == rv[4, 2, 8]
    assert g.getRealValue(-1, 0, 0, extSimbox=True) == -99999.0
    assert np.array_equal(g.getRealTrace(0, 0), rv[:nz, 0, 0])
    assert np.array_equal(g.getRealTrace(4, 3), rv[:nz, 3, 4])
    t = g.getRealTrace(4, 3, add_padding=True)
    assert len(t) == nzp and np.array_equal(t[:nz], rv[:nz, 3, 4]) and np.all(t[nz:] == -99999.0)
    assert np.all(g.getRealTrace(nx, 0) == -99999.0)
    g.close()


STORM_GEO = dict(x_min=431250.5, lx=12800.0, y_min=6472100.25, ly=9600.0, lz=1024.0, angle=0.5235987755982988)


@pytest.mark.parametrize("dims", [(7, 5, 6, 12, 8, 10), (9, 4, 5, 9, 6, 8), (64, 48, 40, 80, 60, 48)])
@pytest.mark.parametrize("padding", [False, True])
def test_storm_file_matches_oracle(ctx, tmp_path, dims, padding):
    """writeStormFile from the device gives the oracle's bytes (box or padded cube); reading the
    oracle's file of the box restores the cube with the padding fill of FFTGrid(NRLib::Grid)."""
    nx, ny, nz, nxp, nyp, nzp = dims
    print("storm file", dims, padding, flush=True)
    rv = _seismic(dims, 41)
    g = _grid(ctx, dims, C.DATA)
    g.upload_real(rv)
    g.writeStormFile(str(tmp_path / "dev"), STORM_GEO, padding=padding)
    cube = rv[:nzp, :nyp, :nxp] if padding else rv[:nz, :ny, :nx]
    O.write_storm(str(tmp_path / "orc.storm"), cube, C.DATA, STORM_GEO)
    assert open(tmp_path / "dev.storm", "rb").read() == open(tmp_path / "orc.storm", "rb").read()
    if not padding:
        h = _grid(ctx, dims)
        geo = h.readStormFile(str(tmp_path / "orc.storm"))
        assert all(abs(geo[k] - STORM_GEO[k]) < 1e-6 for k in STORM_GEO)
        got = h.download_real(padded=True)
        want = O.fill_padded(cube, nxp, nyp, nzp)
        assert np.array_equal(got[:, :, :nxp].view(np.uint32), want[:, :, :nxp].view(np.uint32))
        h.close()
    g.close()


def test_storm_file_errors(ctx, tmp_path):
    dims = (9, 7, 5, 9, 7, 5)
    O.write_storm(str(tmp_path / "a.storm"), np.zeros((5, 7, 9), np.float32), 1, STORM_GEO)
    other = _grid(ctx, (8, 7, 5, 9, 7, 5))
    with pytest.raises(C.CravaError, match="not the grid's box"):
        other.readStormFile(str(tmp_path / "a.storm"))
    g = _grid(ctx, dims)
    with pytest.raises(C.CravaError, match="cannot open"):
        g.readStormFile(str(tmp_path / "missing.storm"))
    raw = open(tmp_path / "a.storm", "rb").read()
    open(tmp_path / "short.storm", "wb").write(raw[:-40])
    with pytest.raises(C.CravaError, match="ends inside the grid"):
        g.readStormFile(str(tmp_path / "short.storm"))
    open(tmp_path / "crava.storm", "wb").write(b"crava_fftgrid_binary\n" + raw)
    with pytest.raises(C.CravaError, match="not a STORM binary grid"):
        g.readStormFile(str(tmp_path / "crava.storm"))
    other.close()
    g.close()


@pytest.mark.parametrize("dims,angle", [((7, 5, 6, 12, 8, 10), 0.3), ((33, 20, 40, 40, 24, 48), 0.0),
                                        ((64, 48, 40, 80, 60, 48), -0.7)])
def test_segy_file_matches_oracle(ctx, tmp_path, dims, angle):
    """writeSegyFile from the device (samples interpolated and converted to IBM floats there) gives
    the oracle's bytes: headers, inline/crossline order, a missing trace left out, tops and
    thicknesses varying by trace."""
    nx, ny, nz, nxp, nyp, nzp = dims
    print("segy file", dims, angle, flush=True)
    rng = np.random.default_rng(nx + ny)
    rv = _seismic(dims, 43)
    g = _grid(ctx, dims, C.DATA)
    g.upload_real(rv)
    geo = dict(x0=431250.5, y0=6472100.25, dx=25.0, dy=20.0, angle=angle, il0=1000.0, xl0=2000.0,
               il_step_x=-np.sin(angle) / 20.0, il_step_y=np.cos(angle) / 20.0,
               xl_step_x=np.cos(angle) / 25.0, xl_step_y=np.sin(angle) / 25.0, dz=4.0, z0=np.float32(1996.0),
               zmax=2000.0 + 4.0 * nz * 1.2 + 40.0, top=2000.0 + rng.uniform(-6, 30, (ny, nx)),
               rel_thick=rng.uniform(0.8, 1.2, (ny, nx)), text=["cravagpu", "MI355X"])
    geo["top"][ny // 2, nx // 3] = -99999.0
    g.writeSegyFile(str(tmp_path / "dev.segy"), geo)
    want = O.segy_write(rv, nx, ny, nz, geo)
    got = open(tmp_path / "dev.segy", "rb").read()
    assert len(got) == len(want) and got == want
    # every top missing: the headers alone
    geo["top"][:] = -99999.0
    g.writeSegyFile(str(tmp_path / "none.segy"), geo)
    assert open(tmp_path / "none.segy", "rb").read() == O.segy_write(rv, nx, ny, nz, geo)
    assert len(O.segy_write(rv, nx, ny, nz, geo)) == 3600
    g.close()


def test_segy_file_one_layer_and_errors(ctx, tmp_path):
    """A grid of one layer (the interpolation clamps to index -1 and reads RMISSING there, as the
    reference's getRealValue does) matches the oracle; tables of another box and a file with no
    samples are refused."""
    dims = (6, 4, 1, 8, 6, 4)
    nx, ny, nz = dims[:3]
    rv = _seismic(dims, 44)
    g = _grid(ctx, dims, C.DATA)
    g.upload_real(rv)
    geo = dict(x0=0.0, y0=0.0, dx=10.0, dy=10.0, angle=0.0, il0=1.0, xl0=1.0, il_step_x=0.0, il_step_y=0.1,
               xl_step_x=0.1, xl_step_y=0.0, dz=4.0, z0=np.float32(100.0), zmax=120.0,
               top=np.full((ny, nx), 100.0), rel_thick=np.ones((ny, nx)))
    g.writeSegyFile(str(tmp_path / "one.segy"), geo)
    assert open(tmp_path / "one.segy", "rb").read() == O.segy_write(rv, nx, ny, nz, geo)
    with pytest.raises(ValueError, match="must be"):
        g.writeSegyFile(str(tmp_path / "bad.segy"), dict(geo, top=np.full((ny + 1, nx), 100.0)))
    with pytest.raises(C.CravaError, match="sample count"):
        g.writeSegyFile(str(tmp_path / "bad.segy"), dict(geo, zmax=90.0))
    g.close()
