"""The oracle's SEG-Y writer (FFTGrid::writeSegyFile + NRLib's SegY writer) pinned by what the
format fixes: IBM floats of known values and their definition, the byte positions of the binary
and trace headers, the linear interpolation of a linear trace, the inline/crossline order and the
traces left out where the top is missing."""
import math
import struct

import numpy as np
import pytest

import oracle as O


def test_ibm_known_values():
    vals = np.array([1.0, -118.625, 0.1, 0.0, -0.0, 0.5, 16.0, 1e-30, 3.0e38, -2.5e-5], np.float32)
    ibm = O.ieee_to_ibm(vals)
    assert [hex(v) for v in ibm[:7]] == ['0x41100000', '0xc276a000', '0x40199999', '0x0', '0x0', '0x40800000',
                                         '0x42100000']
    # truncation: the decoded value is the float with the fraction cut to 24 bits after the hexadecimal
    # normalisation, never above it in magnitude and within 2^-20 relative
    dec = O.ibm_to_float(ibm)
    v64 = vals.astype(np.float64)
    assert np.all(np.abs(dec) <= np.abs(v64)) and np.all(np.sign(dec) == np.sign(v64))
    assert np.all(np.abs(dec - v64) <= 2.0 ** -20 * np.abs(v64))


def test_ibm_random_floats_against_exact_arithmetic():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000)).astype(np.float32)
    ibm = O.ieee_to_ibm(x)
    for v, b in zip(x[:2000], ibm[:2000]):
        m, e = math.frexp(float(v))  # v = m 2^e, 0.5 <= |m| < 1
        q = -((-e) // 4)  # ceil(e / 4)
        frac = int(abs(m) * 2 ** 24) >> (4 * q - e)
        want = (0x80000000 if v < 0 else 0) | ((q + 64) << 24) | frac
        assert int(b) == want
    assert np.all((ibm & 0x00F00000) != 0)  # normalised: a non-zero leading hex digit


def _geo(nx, ny, rng, angle=0.3):
    return dict(x0=1000.0, y0=2000.0, dx=25.0, dy=20.0, angle=angle, il0=100.0, xl0=10.0,
                il_step_x=-math.sin(angle) / 20.0, il_step_y=math.cos(angle) / 20.0,
                xl_step_x=math.cos(angle) / 25.0, xl_step_y=math.sin(angle) / 25.0, dz=4.0, z0=np.float32(1500.0),
                zmax=1500.0 + 4.0 * 30 + 20.0, top=1500.0 + rng.uniform(0, 20, (ny, nx)),
                rel_thick=rng.uniform(0.8, 1.2, (ny, nx)), text=["cravagpu test", "line 3"])


def _parse(data, ns):
    assert len(data) >= 3600
    bh = data[3200:3600]
    traces = []
    pos = 3600
    while pos < len(data):
        th = data[pos:pos + 240]
        s = np.frombuffer(data[pos + 240:pos + 240 + 4 * ns], dtype=">u4")
        traces.append(dict(il=struct.unpack_from(">i", th, 8)[0], xl=struct.unpack_from(">i", th, 20)[0],
                           scalco=struct.unpack_from(">h", th, 70)[0], x=struct.unpack_from(">i", th, 72)[0],
                           y=struct.unpack_from(">i", th, 76)[0], t0=struct.unpack_from(">h", th, 108)[0],
                           ns=struct.unpack_from(">h", th, 114)[0], dt=struct.unpack_from(">H", th, 116)[0],
                           v=O.ibm_to_float(s)))
        pos += 240 + 4 * ns
    return bh, traces


def test_segy_layout_headers_and_linear_traces():
    rng = np.random.default_rng(5)
    nx, ny, nz, nxp, nyp, nzp = 5, 4, 30, 8, 6, 36
    rnxp = 2 * (nxp // 2 + 1)
    # a linear trace v(k) = a(i, j) + b(i, j) k: the interpolation between cells is exact up to float rounding
    a = rng.uniform(-1, 1, (ny, nx))
    b = rng.uniform(-0.1, 0.1, (ny, nx))
    real = np.zeros((nzp, nyp, rnxp), np.float32)
    kk = np.arange(nzp)[:, None, None]
    real[:, :ny, :nx] = (a[None] + b[None] * kk).astype(np.float32)
    geo = _geo(nx, ny, rng)
    geo["top"][1, 2] = -99999.0  # a missing trace is left out
    data = O.segy_write(real, nx, ny, nz, geo)
    ns = int(math.ceil((geo["zmax"] - 1500.0) / 4.0))
    assert data[:3200].decode("cp037").startswith("C 1  ") and "cravagpu test" in data[:3200].decode("cp037")
    bh, traces = _parse(data, ns)
    assert struct.unpack_from(">iii", bh, 0) == (9999, 1, 1)
    assert struct.unpack_from(">hhhhh", bh, 12) == (nx, 0, 4000, 0, ns)
    assert struct.unpack_from(">hhh", bh, 24) == (1, 1, 4) and struct.unpack_from(">h", bh, 302)[0] == 1
    assert len(traces) == nx * ny - 1
    assert [(t["il"], t["xl"]) for t in traces] == sorted((t["il"], t["xl"]) for t in traces)
    assert all(t["scalco"] == 1 and t["ns"] == ns and t["dt"] == 4000 and t["t0"] == 1500 for t in traces)
    cr, sr = math.cos(geo["angle"]), math.sin(geo["angle"])
    seen = set()
    for t in traces:
        # the cell whose centre the header's coordinates truncate
        ij = [(i, j) for j in range(ny) for i in range(nx)
              if int(np.float32((i + .5) * 25 * cr - (j + .5) * 20 * sr + 1000)) == t["x"]
              and int(np.float32((i + .5) * 25 * sr + (j + .5) * 20 * cr + 2000)) == t["y"]]
        assert len(ij) == 1
        i, j = ij[0]
        seen.add((i, j))
        top, gdz = geo["top"][j, i], 4.0 * geo["rel_thick"][j, i]
        first = int(math.floor(0.5 + (top - 1500.0) / 4.0))
        end = min(ns, int(math.floor(0.5 + (top - 1500.0 + nz * gdz) / 4.0)))
        k = np.arange(ns)
        pos = (1500.0 + 4.0 * k - top) / gdz
        inside = (k >= first) & (k < end)
        # the reference's index is the position truncated toward zero: a position in (-1, 0) keeps
        # index 0 and extrapolates linearly, one in (-2, -1] reads layer 0, one in [nz - 1, nz)
        # reads layer nz - 1
        kg = np.trunc(pos)
        eff = np.where(kg == -1, 0.0, np.where(kg == nz - 1, nz - 1.0, pos))
        assert np.all((kg >= -1) & (kg <= nz - 1) | ~inside)
        want = np.where(inside, a[j, i] + b[j, i] * eff, 0.0)
        assert np.allclose(t["v"], want, atol=2e-6, rtol=1e-5)
        assert np.all(t["v"][~inside] == 0)
    assert (2, 1) not in seen


def test_segy_trace_beyond_the_grid_is_missing():
    """A top far above the file start puts samples more than one cell beyond the grid's last
    layer: getRegularZInterpolatedRealValue returns RMISSING there, and the writer keeps it."""
    rng = np.random.default_rng(6)
    nx, ny, nz = 2, 2, 10
    real = np.ones((12, 4, 4), np.float32)
    geo = _geo(nx, ny, rng, angle=0.0)
    geo["top"][:] = 1500.0
    geo["rel_thick"][:] = 1.0
    geo["zmax"] = 1500.0 + 4.0 * 40
    data = O.segy_write(real, nx, ny, nz, geo)
    _, traces = _parse(data, 40)
    v = traces[0]["v"]
    assert np.all(v[:10] == 1.0) and np.all(v[10:] == 0.0)  # end = nz samples: the rest written as 0


def test_textual_header_is_nrlibs():
    """NRLib's TextualHeader (commonheaders.cpp:60-213): the caller's lines go on the empty header,
    whose last line "C40 END EBCDIC" (one space) is 79 characters followed by the string's
    terminating zero; with no lines the standard header carries the writing program's credits
    (lines 12-28) and a full last line."""
    rng = np.random.default_rng(7)
    real = np.zeros((12, 4, 4), np.float32)
    geo = _geo(2, 2, rng, angle=0.0)
    with_text = O.segy_write(real, 2, 2, 10, geo)[:3200]
    lines = [with_text[80 * l:80 * l + 80] for l in range(40)]
    assert lines[0].decode("cp037") == "C 1" + " " * 77
    assert lines[1].decode("cp037").startswith("C 2  cravagpu test")
    assert lines[39][:79].decode("cp037") == "C40 END EBCDIC" + " " * 65 and lines[39][79] == 0
    std = O.segy_write(real, 2, 2, 10, dict(geo, text=[]))[:3200].decode("cp037")
    sl = [std[80 * l:80 * l + 80] for l in range(40)]
    assert sl[13] == "C14" + " " * 9 + "THE CRAVA PROGRAM" + " " * 51
    assert sl[11].rstrip() == "C12         THIS SEGY FILE WAS WRITTEN BY" and sl[27].rstrip() == "C28         2003"
    assert sl[39] == "C40 END EBCDIC" + " " * 66 and sl[10] == "C11" + " " * 77
    assert sum(1 for s in sl if s[3:].strip()) == 10
