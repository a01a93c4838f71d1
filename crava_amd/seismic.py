"""Seismic and parameter input from SEG-Y files onto the grid: NRLib's SegY reading
(cg_segy_read), its cut of the traces to the inversion volume (window_traces) and
CommonData::FillInData's column loop (commondata.cpp:6451-6775) for the seismic (DATA) branch and
the branch for other grids (background and parameter cubes, PARAMETER), whose per-trace work runs
on the device (FFTGrid.fillInData, cg_grid_fill_from_traces).

The column loop follows the reference: column (i, j) of the padded grid, i < rnxp, takes the box
cell refi = GetFillNumber(i, nx, nxp), refj = GetFillNumber(j, ny, nyp) (no cell for i >= nxp),
its centre from Simbox::getCoord(refi, refj, 0) (simbox.cpp:814-842), the SEG-Y trace nearest to
that centre (SegyGeometry::FindContIndex / FindIndex and SegY::GetNearestTrace, segy.cpp:1031-1050,
segygeometry.cpp:955-1035) and z0 + dz / 2 and dz of the cell column (commondata.cpp:6555-6640;
z0 itself for the other grids, whose FillInData call leaves is_seismic false).  The trace's first
sample is at z0_ + (j0 + 0.5) dz, less dz / 2 for seismic data (commondata.cpp:6582-6585), z0_ the
file's start time (the first trace's, SegY::ReadTrace) and j0 the first sample the volume cut kept.
The lattice of the SEG-Y volume is the one its inline and crossline numbers span, with the
first inline and crossline the smallest ones and inlines along the first axis (NRLib's
SegyGeometry when IL0 == minIL, XL0 == minXL and firstAxisIL); the continuous inline / crossline
of a point is the caller's linear map, as SegyGeometry::FindContILXL computes it."""
import ctypes
import math

import numpy as np

from . import DATA, PARAMETER, FFTGrid, CravaError, lib  # noqa: F401

RMISSING = -99999.0


class SegyFormat(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("il_pos", "xl_pos", "x_pos", "y_pos", "scalco_pos", "start_time_pos")]


class SegyTraces(ctypes.Structure):
    _fields_ = [("ntraces", ctypes.c_int), ("nsamples", ctypes.c_int), ("dz", ctypes.c_float),
                ("format_code", ctypes.c_int), ("samples", ctypes.POINTER(ctypes.c_float)),
                ("il", ctypes.POINTER(ctypes.c_int)), ("xl", ctypes.POINTER(ctypes.c_int)),
                ("x", ctypes.POINTER(ctypes.c_double)), ("y", ctypes.POINTER(ctypes.c_double)),
                ("start_time", ctypes.POINTER(ctypes.c_float))]


def read_segy(path, header_positions=None):
    """Every trace of a SEG-Y file (cg_segy_read): dict of samples (ntraces, nsamples), il, xl,
    x, y, start_time (ms), dz (ms).  header_positions: dict of 1-based trace-header byte
    positions il_pos, xl_pos, x_pos, y_pos, scalco_pos, start_time_pos (SeisWorks by default)."""
    L = lib()
    L.cg_segy_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(SegyFormat), ctypes.POINTER(SegyTraces)]
    L.cg_segy_read.restype = ctypes.c_int
    fmt = SegyFormat(**(header_positions or {}))
    t = SegyTraces()
    rc = L.cg_segy_read(path.encode(), ctypes.byref(fmt), ctypes.byref(t))
    if rc != 0:
        raise CravaError("cg_segy_read(%s) failed (%d)" % (path, rc))
    n, ns = t.ntraces, t.nsamples
    out = dict(samples=np.zeros((n, ns), np.float32), il=np.zeros(n, np.int32), xl=np.zeros(n, np.int32),
               x=np.zeros(n), y=np.zeros(n), start_time=np.zeros(n, np.float32), dz=float(t.dz))
    t.samples = out["samples"].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    t.il = out["il"].ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    t.xl = out["xl"].ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    t.x = out["x"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    t.y = out["y"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    t.start_time = out["start_time"].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    rc = L.cg_segy_read(path.encode(), ctypes.byref(fmt), ctypes.byref(t))
    if rc != 0:
        raise CravaError("cg_segy_read(%s) failed (%d)" % (path, rc))
    return out


def window_traces(traces, z_top, z_bot, z_pad, inside=None, offset=None):
    """SegY::ReadAllTraces / ReadTrace with a volume (segy.cpp:619-890): each trace cut to the
    samples j0..j1 around the volume's top and base at the trace, j0 = max(0, int((zTop - pad -
    z0_) / dz)), j1 = min(ns - 1, (zBot + pad - z0_) / dz) (truncated), pad = z_pad / 2, in float as
    the reader computes them; z0_ the offset if given, else the first trace's start time.  A trace
    outside the volume (inside[t] false: only_volume), with a missing surface value (RMISSING in
    z_top or z_bot) or with the volume reaching past its data (zTop < z0_ or zBot > z0_ + ns dz)
    is not kept.  Returns the traces as read_segy does, the samples as a list of the kept traces
    (None where not kept), with start (j0 per trace) and z0 (z0_)."""
    f32 = np.float32
    ns = traces["samples"].shape[1]
    dz = f32(traces["dz"])
    z0 = f32(offset) if offset is not None else f32(traces["start_time"][0])
    pad = f32(0.5 * z_pad)
    s_bot = f32(z0 + f32(f32(ns) * dz))
    kept, start = [], np.zeros(len(traces["il"]), dtype=np.int64)
    for t in range(len(traces["il"])):
        zt, zb = f32(z_top[t]), f32(z_bot[t])
        if (inside is not None and not inside[t]) or zt == f32(RMISSING) or zb == f32(RMISSING) or zt < z0 or \
                zb > s_bot:
            kept.append(None)
            continue
        j0 = max(0, int(f32(f32(f32(zt - pad) - z0) / dz)))
        j1 = min(ns - 1, int(f32(f32(f32(zb + pad) - z0) / dz)))
        if j0 > j1:
            raise CravaError("window_traces: lower horizon above the SEG-Y data or upper horizon below it")
        kept.append(np.asarray(traces["samples"][t, j0:j1 + 1], dtype=np.float32))
        start[t] = j0
    return dict(traces, samples=kept, start=start, z0=float(z0))


def fill_number(i, n, np_):
    """CommonData::GetFillNumber (commondata.cpp:6777-6814): the box index whose trace fills
    padded index i, None at or past np_."""
    if i >= np_:
        return None
    if i < n:
        return i
    below, above = np_ - i - 1, i - n + 1
    return max(n - above, n // 2) if above < below else min(below, n // 2)


def segy_lattice(traces):
    """The inline / crossline lattice of the traces (steps the smallest positive differences) and
    the trace at each lattice node ((nx, ny) array, -1 where none)."""
    def step(v):
        u = np.unique(v)
        d = np.diff(u)
        return int(d[d > 0].min()) if len(u) > 1 else 1
    il, xl = traces["il"], traces["xl"]
    lat = dict(min_il=int(il.min()), max_il=int(il.max()), il_step=step(il),
               min_xl=int(xl.min()), max_xl=int(xl.max()), xl_step=step(xl))
    lat["nx"] = (lat["max_il"] - lat["min_il"]) // lat["il_step"] + 1
    lat["ny"] = (lat["max_xl"] - lat["min_xl"]) // lat["xl_step"] + 1
    node = np.full((lat["nx"], lat["ny"]), -1, dtype=np.int64)
    i = (il - lat["min_il"]) // lat["il_step"]
    j = (xl - lat["min_xl"]) // lat["xl_step"]
    node[i, j] = np.arange(len(il))
    lat["node"] = node
    return lat


def resolve_columns(traces, lattice, transform, simbox, nx, ny, nz, nxp, nyp, grid_type=DATA):
    """FillInData's column loop: (column_trace, z0_grid, dz_grid), each (nyp, rnxp), and the
    traces' z0_data (ms).  column_trace: the trace of the column, -1 outside the data, -2 a dead
    column (no trace kept at its lattice node, or a missing cell).  transform: in_line0,
    cross_line0, x0, y0, il_step_x, il_step_y, xl_step_x, xl_step_y (SegyGeometry::FindContILXL);
    simbox: x0, y0, dx, dy, angle and top, bot (ny, nx) of the box cells (RMISSING where missing).
    traces: read_segy's, or window_traces' (a trace not kept has no samples).  grid_type DATA: the
    seismic call (is_seismic true); PARAMETER: the other grids'.  Vectorised over the columns, each
    in the reference's arithmetic (the cell centre in double, rounded to float as the reference
    passes it on; the lattice index truncated)."""
    seismic = grid_type == DATA
    rnxp = 2 * (nxp // 2 + 1)
    ri = np.array([-1 if fill_number(i, nx, nxp) is None else fill_number(i, nx, nxp) for i in range(rnxp)])
    rj = np.array([fill_number(j, ny, nyp) for j in range(nyp)])
    RI, RJ = np.meshgrid(ri, rj)  # (nyp, rnxp)
    has_cell = RI >= 0
    RIc = np.where(has_cell, RI, 0)
    cr, sr = math.cos(simbox["angle"]), math.sin(simbox["angle"])
    rx, ry = (RIc + 0.5) * simbox["dx"], (RJ + 0.5) * simbox["dy"]
    xf = (rx * cr - ry * sr + simbox["x0"]).astype(np.float32).astype(np.float64)
    yf = (rx * sr + ry * cr + simbox["y0"]).astype(np.float32).astype(np.float64)
    tf = transform
    il = tf["in_line0"] + (xf - tf["x0"]) * tf["il_step_x"] + (yf - tf["y0"]) * tf["il_step_y"]
    xl = tf["cross_line0"] + (xf - tf["x0"]) * tf["xl_step_x"] + (yf - tf["y0"]) * tf["xl_step_y"]
    xind = 0.5 + (il - lattice["min_il"]) / lattice["il_step"]
    yind = 0.5 + (xl - lattice["min_xl"]) / lattice["xl_step"]
    inside = has_cell & (xind >= 0.0) & (xind < lattice["nx"]) & (yind >= 0.0) & (yind < lattice["ny"])
    xi = np.where(inside, xind, 0.0).astype(np.int64)
    yi = np.where(inside, yind, 0.0).astype(np.int64)
    t = np.where(inside, lattice["node"][xi, yi], -1)
    samples = traces["samples"]
    kept = np.array([samples[q] is not None for q in range(len(samples))]) if isinstance(samples, list) else \
        np.ones(len(traces["il"]), dtype=bool)
    top = np.asarray(simbox["top"], dtype=np.float64)[RJ, RIc]
    bot = np.asarray(simbox["bot"], dtype=np.float64)[RJ, RIc]
    live = (t >= 0) & (top != RMISSING) & (bot != RMISSING)
    live &= kept[np.maximum(t, 0)]
    dz = np.where(live, (bot - top) / nz, 1.0)
    z0 = top + 0.5 * dz  # Simbox::getCoord(refi, refj, 0)
    ct = np.where(live, t, np.where(inside, -2, -1)).astype(np.int32)
    z0g = z0.astype(np.float32)
    if seismic:
        z0g = z0g + np.float32(0.5) * dz.astype(np.float32)
    z0g = np.where(live, z0g, 0.0).astype(np.float32)
    dzg = dz.astype(np.float32)
    # SegY::GetNearestTrace's start time (segy.cpp:1045), less half a sample for seismic data
    f32 = np.float32
    dzd = f32(traces["dz"])
    z0f = f32(traces.get("z0", traces["start_time"][0]))
    start = np.asarray(traces.get("start", np.zeros(len(traces["il"]))), dtype=np.float32)
    z0d = (z0f + ((start + f32(0.5)).astype(f32) * dzd).astype(f32)).astype(f32)
    if seismic:
        z0d = (z0d - f32(f32(0.5) * dzd)).astype(f32)
    return ct, z0g, dzg, z0d


def kept_traces(traces, column_trace, z0_data):
    """The traces the columns use, as fillInData takes them: (list of samples, their z0_data, the
    columns renumbered); traces window_traces did not keep are left out."""
    s = traces["samples"]
    keep = [q for q in range(len(s)) if s[q] is not None]
    remap = np.full(len(s) + 1, -1, dtype=np.int32)
    remap[keep] = np.arange(len(keep), dtype=np.int32)
    ct = np.where(column_trace >= 0, remap[np.maximum(column_trace, 0)], column_trace).astype(np.int32)
    return [np.asarray(s[q], dtype=np.float32) for q in keep], np.asarray(z0_data, np.float32)[keep], ct


def fill_in_data_from_segy(grid, path, transform, simbox, smooth_length, header_positions=None, grid_type=DATA,
                           window=None):
    """CommonData::FillInData for an FFTGrid from a SEG-Y file: the traces read (and cut to the
    volume when window = dict(z_top, z_bot, z_pad[, inside, offset]) gives it), the columns
    resolved on the host, the per-trace work on the device.  grid_type DATA: seismic data;
    PARAMETER: a background or parameter cube (CravaError if the data do not cover the box)."""
    tr = read_segy(path, header_positions)
    if window is not None:
        tr = window_traces(tr, **window)
    lat = segy_lattice(tr)
    ct, z0g, dzg, z0d = resolve_columns(tr, lat, transform, simbox, grid.nx, grid.ny, grid.nz, grid.nxp, grid.nyp,
                                        grid_type)
    samples, z0k, ctk = kept_traces(tr, ct, z0d)
    grid.fillInData(samples, z0k, tr["dz"], ctk, z0g, dzg, smooth_length, grid_type)
    return tr, ct


# ---------------------------------------------------------------- STORM grids as input
def read_storm(path):
    """StormContGrid::ReadFromFile (stormcontgrid.cpp:82-170) of a binary STORM grid whose top and
    base surfaces are constants (a surface given by file name raises CravaError): dict of values
    (nk, nj, ni) float32 (i fastest in the file, big-endian), x_min, lx, y_min, ly, top, bot, lz,
    angle (radians; the header holds degrees), ni, nj, nk."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0

    def token():
        nonlocal pos
        while pos < len(data) and data[pos:pos + 1].isspace():
            pos += 1
        st = pos
        while pos < len(data) and not data[pos:pos + 1].isspace():
            pos += 1
        return data[st:pos].decode("latin-1")
    if token() != "storm_petro_binary":
        raise CravaError("read_storm: %s is not a binary STORM grid" % path)
    tok = [token() for _ in range(17)]  # 4 header tokens, the volume (8), lz, angle, ni nj nk
    nl = data.find(b"\n", pos)  # the rest of the size line
    if not tok[-1] or nl < 0:
        raise CravaError("read_storm: %s ends inside the header" % path)
    pos = nl + 1
    try:
        top, bot = float(tok[8]), float(tok[9])
    except ValueError:
        raise CravaError("read_storm: surfaces given by file name (%s, %s) are not read" % (tok[8], tok[9]))
    ni, nj, nk = int(tok[14]), int(tok[15]), int(tok[16])
    if len(data) - pos < 4 * ni * nj * nk:
        raise CravaError("read_storm: %s ends inside the grid" % path)
    values = np.frombuffer(data, dtype=">f4", count=ni * nj * nk, offset=pos).astype(np.float32)
    return dict(values=values.reshape(nk, nj, ni), x_min=float(tok[4]), lx=float(tok[5]), y_min=float(tok[6]),
                ly=float(tok[7]), top=top, bot=bot, lz=float(tok[12]), angle=math.radians(float(tok[13])),
                ni=ni, nj=nj, nk=nk)


def storm_columns(storm, simbox, nx, ny, nz, nxp, nyp, grid_type=DATA):
    """FillInData's column loop with a STORM grid as the source (is_storm, commondata.cpp:6560-6611):
    column (i, j) of the padded grid at the centre of box cell GetFillNumber(i), GetFillNumber(j)
    takes the STORM column StormContGrid::FindXYIndex gives for that point when
    StormContGrid::IsInside holds (the point as float, stormcontgrid.cpp:291-322, volume.cpp:458-575),
    -1 outside, -2 where the box cell's top or base is missing.  The trace's first sample is at the
    centre of the STORM column's first cell and its sample interval the distance between the
    centres of its first and last cells over nk - 1 (FindCenterOfCell, :494-503; in float).
    Returns (traces, z0_data, dz_data, column_trace, z0_grid, dz_grid): the STORM columns used,
    in order of first use."""
    seismic = grid_type == DATA
    f32 = np.float32
    rnxp = 2 * (nxp // 2 + 1)
    ca, sa = math.cos(simbox["angle"]), math.sin(simbox["angle"])
    cs, ss = math.cos(storm["angle"]), math.sin(storm["angle"])
    dxs, dys = storm["lx"] / storm["ni"], storm["ly"] / storm["nj"]
    nk = storm["nk"]
    dzs = (storm["bot"] - storm["top"]) / nk
    z_min = f32(storm["top"] + 0.5 * dzs)
    z_max = f32(storm["top"] + (nk - 1 + 0.5) * dzs)
    dz_data = f32(f32(z_max - z_min) / f32(nk - 1))
    ct = np.full((nyp, rnxp), -1, dtype=np.int32)
    z0g = np.zeros((nyp, rnxp), dtype=np.float32)
    dzg = np.ones((nyp, rnxp), dtype=np.float32)
    index, traces = {}, []
    top, bot = np.asarray(simbox["top"], np.float64), np.asarray(simbox["bot"], np.float64)
    for j in range(nyp):
        rj = fill_number(j, ny, nyp)
        for i in range(rnxp):
            ri = fill_number(i, nx, nxp)
            if ri is None:
                continue
            rx, ry = (ri + 0.5) * simbox["dx"], (rj + 0.5) * simbox["dy"]
            x = rx * ca - ry * sa + simbox["x0"]
            y = rx * sa + ry * ca + simbox["y0"]
            xf, yf = float(f32(x)), float(f32(y))
            lxf = (xf - storm["x_min"]) * cs + (yf - storm["y_min"]) * ss
            lyf = -(xf - storm["x_min"]) * ss + (yf - storm["y_min"]) * cs
            if lxf < 0.0 or lxf > storm["lx"] or lyf < 0.0 or lyf > storm["ly"]:
                continue  # outside the STORM volume: zeros
            t_, b_ = top[rj, ri], bot[rj, ri]
            if t_ == RMISSING or b_ == RMISSING:
                ct[j, i] = -2
                continue
            lx_ = cs * (x - storm["x_min"]) + ss * (y - storm["y_min"])
            ly_ = -ss * (x - storm["x_min"]) + cs * (y - storm["y_min"])
            ii, jj = int(lx_ / dxs), int(ly_ / dys)
            if lx_ < 0.0 or ly_ < 0.0 or ii >= storm["ni"] or jj >= storm["nj"]:
                raise CravaError("storm_columns: an xy-position outside the STORM grid (FindXYIndex)")
            if (ii, jj) not in index:
                index[(ii, jj)] = len(traces)
                traces.append(np.asarray(storm["values"][:, jj, ii], dtype=np.float32))
            dz = (b_ - t_) / nz
            z0 = t_ + 0.5 * dz  # Simbox::getCoord(refi, refj, 0)
            ct[j, i] = index[(ii, jj)]
            z0g[j, i] = f32(z0) + f32(0.5) * f32(dz) if seismic else f32(z0)
            dzg[j, i] = f32(dz)
    z0_data = np.full(len(traces), z_min, dtype=np.float32)
    return traces, z0_data, float(dz_data), ct, z0g, dzg


def fill_in_data_from_storm(grid, path, simbox, smooth_length, grid_type=DATA):
    """CommonData::FillInData for an FFTGrid from a binary STORM grid of any lateral and vertical
    geometry (constant surfaces): the columns resolved on the host, the per-trace work on the device
    (a trace's trimmed zeros leave its start where it is).  Returns (storm, column_trace)."""
    storm = read_storm(path)
    traces, z0d, dz_data, ct, z0g, dzg = storm_columns(storm, simbox, grid.nx, grid.ny, grid.nz, grid.nxp, grid.nyp,
                                                       grid_type)
    if not traces:
        raise CravaError("fill_in_data_from_storm: the STORM grid does not cover the box")
    grid.fillInData(traces, z0d, dz_data, ct, z0g, dzg, smooth_length, grid_type, trim_keeps_z0=True)
    return storm, ct


def trace_counts(column_trace, nx, ny, traces=None, grid_type=DATA):
    """FillInData's counters (commondata.cpp:6456-6460, 6731-6770): missing_traces_simbox (box
    columns outside the data), missing_traces_padding (padding columns outside it),
    dead_traces_simbox (columns inside the data without a trace -- for a non-DATA grid also those
    whose trace has no nonzero sample) and dead_traces_map ((nyp, rnxp) bool).  The reference
    stops a non-DATA grid at its first box column outside the data; the counters here cover every
    column (cg_grid_fill_from_traces refuses such a grid)."""
    ct = np.asarray(column_trace)
    nyp, rnxp = ct.shape
    box = np.zeros_like(ct, dtype=bool)
    box[:ny, :nx] = True
    dead = ct == -2
    if grid_type != DATA and traces is not None:
        empty = np.array([not np.any(np.asarray(t) != 0) for t in traces] + [False])
        dead |= (ct >= 0) & empty[np.where(ct >= 0, ct, len(traces))]
    out = ct == -1
    return dict(missing_traces_simbox=int((out & box).sum()), missing_traces_padding=int((out & ~box).sum()),
                dead_traces_simbox=int(dead.sum()), dead_traces_map=dead)
