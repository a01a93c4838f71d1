"""SEG-Y and STORM input (crava_amd.seismic): NRLib's SegY reading (cg_segy_read, host code) of a
file made by the SEG-Y writer's oracle, the cut to the volume, the inline / crossline lattice,
FillInData's column loop, and a STORM grid's columns (StormContGrid::FindXYIndex)."""
import os

import numpy as np

import crava_amd as C
import oracle as O
from crava_amd import seismic as SE


def _geo(nx, ny, nz, rng):
    top = 1000.0 + rng.uniform(0.0, 40.0, (ny, nx))
    # inline / crossline origins off the half-integers, so that no cell centre sits on a rounding tie
    return dict(x0=4000.0, y0=7000.0, dx=25.0, dy=25.0, angle=0.3, il0=100.2, xl0=200.3,
                il_step_x=np.cos(0.3) / 25.0, il_step_y=np.sin(0.3) / 25.0,
                xl_step_x=-np.sin(0.3) / 25.0, xl_step_y=np.cos(0.3) / 25.0, dz=4.0, z0=900.0, zmax=1400.0,
                top=top, rel_thick=np.ones((ny, nx)), text=["test"])


def test_read_back_the_writer_oracle(tmp_path):
    rng = np.random.default_rng(3)
    nx, ny, nz, nxp, nyp, nzp = 6, 5, 40, 8, 6, 48
    real = rng.standard_normal((nzp, nyp, 2 * (nxp // 2 + 1))).astype(np.float32)
    geo = _geo(nx, ny, nz, rng)
    data = O.segy_write(real, nx, ny, nz, geo)
    path = os.path.join(str(tmp_path), "a.segy")
    with open(path, "wb") as f:
        f.write(data)
    tr = SE.read_segy(path)
    n = nx * ny
    ns = int(np.ceil((geo["zmax"] - geo["z0"]) / 4.0))
    assert tr["samples"].shape == (n, ns) and tr["dz"] == 4.0
    # the samples as written: every trace's IBM words decoded, in file order
    words = np.frombuffer(data, dtype=">u4", offset=3600).reshape(n, 60 + ns)
    want = O.ibm_to_float(words[:, 60:].astype(np.uint32)).astype(np.float32)
    assert np.array_equal(tr["samples"], want)
    hdr = np.frombuffer(data, dtype=np.uint8, offset=3600).reshape(n, 240 + 4 * ns)[:, :240]
    be = lambda a, k: int.from_bytes(bytes(a[k:k + 4]), "big", signed=True)
    assert [be(h, 8) for h in hdr] == list(tr["il"]) and [be(h, 20) for h in hdr] == list(tr["xl"])
    assert [float(be(h, 72)) for h in hdr] == list(tr["x"]) and np.all(tr["start_time"] == 900.0)
    # the lattice holds every trace once, at its inline / crossline node
    lat = SE.segy_lattice(tr)
    assert (lat["node"] >= 0).sum() == n
    for t in range(n):
        i = (tr["il"][t] - lat["min_il"]) // lat["il_step"]
        j = (tr["xl"][t] - lat["min_xl"]) // lat["xl_step"]
        assert lat["node"][i, j] == t


def test_fill_number_matches_the_oracle():
    for n, np_ in ((5, 8), (4, 8), (3, 8), (20, 24), (7, 12)):
        for i in range(np_ + 2):
            want = O.fill_number(i, n, np_) if i < np_ else None
            assert SE.fill_number(i, n, np_) == want, (i, n, np_)


def test_columns_of_a_simbox_over_the_written_traces(tmp_path):
    rng = np.random.default_rng(4)
    nx, ny, nz, nxp, nyp, nzp = 6, 5, 40, 8, 6, 48
    real = rng.standard_normal((nzp, nyp, 2 * (nxp // 2 + 1))).astype(np.float32)
    geo = _geo(nx, ny, nz, rng)
    path = os.path.join(str(tmp_path), "b.segy")
    with open(path, "wb") as f:
        f.write(O.segy_write(real, nx, ny, nz, geo))
    tr = SE.read_segy(path)
    lat = SE.segy_lattice(tr)
    transform = dict(in_line0=geo["il0"], cross_line0=geo["xl0"], x0=geo["x0"], y0=geo["y0"],
                     il_step_x=geo["il_step_x"], il_step_y=geo["il_step_y"], xl_step_x=geo["xl_step_x"],
                     xl_step_y=geo["xl_step_y"])
    top = geo["top"].copy()
    top[1, 2] = SE.RMISSING
    simbox = dict(x0=geo["x0"], y0=geo["y0"], dx=geo["dx"], dy=geo["dy"], angle=geo["angle"], top=top,
                  bot=geo["top"] + nz * geo["dz"])
    ct, z0g, dzg, z0d = SE.resolve_columns(tr, lat, transform, simbox, nx, ny, nz, nxp, nyp)
    # the trace of box cell (i, j) is the one the writer wrote for it (its IL / XL from the same
    # cell centre); padding columns take their fill cell's trace; past nxp, a missing top: none
    order = {}
    cr, sr = np.cos(geo["angle"]), np.sin(geo["angle"])
    for j in range(ny):
        for i in range(nx):
            rx, ry = (i + 0.5) * 25.0, (j + 0.5) * 25.0
            xf = float(np.float32(rx * cr - ry * sr + geo["x0"]))
            yf = float(np.float32(rx * sr + ry * cr + geo["y0"]))
            il = int(0.5 + geo["il0"] + (xf - geo["x0"]) * geo["il_step_x"] + (yf - geo["y0"]) * geo["il_step_y"])
            xl = int(0.5 + geo["xl0"] + (xf - geo["x0"]) * geo["xl_step_x"] + (yf - geo["y0"]) * geo["xl_step_y"])
            order[(i, j)] = int(np.where((tr["il"] == il) & (tr["xl"] == xl))[0][0])
    for j in range(nyp):
        for i in range(2 * (nxp // 2 + 1)):
            ri, rj = SE.fill_number(i, nx, nxp), SE.fill_number(j, ny, nyp)
            if ri is None:
                assert ct[j, i] == -1
            elif (ri, rj) == (2, 1):
                assert ct[j, i] == -2  # inside the data, a missing cell: dead
            else:
                assert ct[j, i] == order[(ri, rj)], (i, j)
                assert abs(dzg[j, i] - 4.0) < 1e-6 and abs(z0g[j, i] - (top[rj, ri] + 4.0)) < 1e-3
    assert np.all(z0d == 900.0)


def test_window_traces_cuts_to_the_volume():
    # SegY::ReadTrace with a volume: samples j0..j1 around the surfaces plus half the padding,
    # traces outside the volume, with a missing surface or beyond the data not kept
    rng = np.random.default_rng(5)
    n, ns = 7, 100
    tr = dict(samples=rng.standard_normal((n, ns)).astype(np.float32), il=np.arange(n), xl=np.zeros(n, np.int32),
              x=np.zeros(n), y=np.zeros(n), start_time=np.full(n, 900.0, np.float32), dz=4.0)
    z_top = np.array([1000.0, 950.5, SE.RMISSING, 880.0, 1000.0, 1000.0, 925.0])
    z_bot = np.array([1100.0, 1200.0, 1100.0, 1000.0, 1350.0, 1100.0, 1290.0])
    inside = np.array([True, True, True, True, True, False, True])
    w = SE.window_traces(tr, z_top, z_bot, 40.0, inside=inside)
    kept = [s is not None for s in w["samples"]]
    assert kept == [True, True, False, False, False, False, True]
    # (zTop - 20 - 900) / 4 and (zBot + 20 - 900) / 4, truncated; the second clipped at ns - 1
    for t, (j0, j1) in {0: (20, 55), 1: (7, 80), 6: (1, 99)}.items():
        assert w["start"][t] == j0 and np.array_equal(w["samples"][t], tr["samples"][t, j0:j1 + 1])
    assert w["z0"] == 900.0
    # the trace's first sample: z0_ + (j0 + 0.5) dz, less dz / 2 for seismic data
    samples, z0k, ct = SE.kept_traces(w, np.array([[0, 1, 6, -2, -1]]), np.array([1.0, 2.0, 0, 0, 0, 0, 3.0]))
    assert [len(s) for s in samples] == [36, 74, 99] and list(z0k) == [1.0, 2.0, 3.0]
    assert ct.tolist() == [[0, 1, 2, -2, -1]]


def _write_storm(path, values, x_min, lx, y_min, ly, top, bot, angle_deg):
    nk, nj, ni = values.shape
    head = ("storm_petro_binary\n0 2 -999.000000\nFFTGrid\n%f %f %f %f %s %s 0.0 0.0\n%f %f\n\n%d %d %d\n"
            % (x_min, lx, y_min, ly, top, bot, float(bot) - float(top), angle_deg, ni, nj, nk))
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(np.ascontiguousarray(values, dtype=">f4").tobytes())
        f.write(b"0\n")


def test_read_storm_and_its_columns(tmp_path):
    # a STORM grid on the simbox's own lattice (same origin, angle and cell size): box cell (i, j)
    # takes STORM column (i, j); a second one shifted east leaves the western columns outside
    rng = np.random.default_rng(6)
    nx, ny, nz, nxp, nyp = 6, 5, 20, 8, 6
    vals = rng.standard_normal((30, ny, nx)).astype(np.float32)
    path = os.path.join(str(tmp_path), "a.storm")
    _write_storm(path, vals, 4000.0, nx * 25.0, 7000.0, ny * 25.0, 1000, 1150, 30.0)
    st = SE.read_storm(path)
    assert np.array_equal(st["values"], vals) and st["ni"] == nx and st["nk"] == 30
    assert abs(st["angle"] - np.radians(30.0)) < 1e-12 and st["top"] == 1000.0 and st["bot"] == 1150.0
    top = 1000.0 + rng.uniform(0.0, 20.0, (ny, nx))
    top[2, 3] = SE.RMISSING
    simbox = dict(x0=4000.0, y0=7000.0, dx=25.0, dy=25.0, angle=np.radians(30.0), top=top, bot=top + 120.0)
    traces, z0d, dzd, ct, z0g, dzg = SE.storm_columns(st, simbox, nx, ny, nz, nxp, nyp, C.DATA)
    for j in range(nyp):
        for i in range(2 * (nxp // 2 + 1)):
            ri, rj = SE.fill_number(i, nx, nxp), SE.fill_number(j, ny, nyp)
            if ri is None:
                assert ct[j, i] == -1
            elif (ri, rj) == (3, 2):
                assert ct[j, i] == -2
            else:
                assert np.array_equal(traces[ct[j, i]], vals[:, rj, ri]), (i, j)
                assert abs(z0g[j, i] - (top[rj, ri] + 6.0)) < 1e-3 and abs(dzg[j, i] - 6.0) < 1e-5
    # the first cell's centre and the distance between first and last centres over nk - 1
    assert z0d[0] == np.float32(1002.5) and abs(dzd - 5.0) < 1e-5
    _, _, _, ctp, z0p, _ = SE.storm_columns(st, simbox, nx, ny, nz, nxp, nyp, C.PARAMETER)
    assert abs(z0p[0, 0] - (top[0, 0] + 3.0)) < 1e-3  # the cell centre itself (is_seismic false)
    path2 = os.path.join(str(tmp_path), "b.storm")
    _write_storm(path2, vals, 4000.0 + 60.0 * np.cos(np.radians(30.0)), nx * 25.0, 7000.0 + 60.0 * np.sin(np.radians(30.0)),
                 ny * 25.0, 1000, 1150, 30.0)
    _, _, _, ct2, _, _ = SE.storm_columns(SE.read_storm(path2), simbox, nx, ny, nz, nxp, nyp, C.DATA)
    east = ct2[:ny, 3:nx].copy()
    assert east[2, 0] == -2  # the missing top
    east[2, 0] = 0
    assert np.all(ct2[:ny, :2] == -1) and np.all(east >= 0)


def test_trace_counts():
    ct = np.array([[0, 1, -1, -2, -1, -1], [2, -2, 1, 0, 2, -1], [-1, 0, 0, 1, 2, -1]], dtype=np.int32)
    c = SE.trace_counts(ct, 3, 2)
    # box (2 x 3): one outside, two dead; padding: rows 2 and columns 3..5
    assert c["missing_traces_simbox"] == 1 and c["dead_traces_simbox"] == 2
    assert c["missing_traces_padding"] == int((ct == -1).sum()) - 1
    traces = [np.ones(4), np.zeros(4), np.array([0.0, 0.0, 2.0, 0.0])]
    cp = SE.trace_counts(ct, 3, 2, traces, C.PARAMETER)
    assert cp["dead_traces_simbox"] == 2 + int((ct == 1).sum())  # trace 1 has no nonzero sample
    assert np.array_equal(cp["dead_traces_map"], (ct == -2) | (ct == 1))
