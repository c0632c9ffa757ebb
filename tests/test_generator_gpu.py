"""GPU parity of the initial-condition generator (i3g_markers_paint,
i3g_temperature_boxes) with the reference's own in3mg run on
g16_init.t3c (golden g16): the regular marker lattice painted with the
deck's rock bodies, the node temperatures from its temperature boxes, the
marker temperatures interpolated from them (in3mg.c:954-966).  Types, melt
extraction and immobile origins exact; the strain equal in magnitude (the
generator's two ronurecalc passes set its sign by viscalc's yield test);
marker temperatures within 1e-6 relative (the device's erf and pow)."""
import os

import numpy as np
import pytest

import i3elvis_amd as I
from conftest import golden

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIZES = (2.4e5, 1.2e5, 1.8e5)


def _grid_lines():
    """the generator's grid lines as doubles (the VTR file holds them
    exactly; the .prn only as floats)"""
    data = open(os.path.join(HERE, "g16_nodal.vtr"), "rb").read()
    key = b'<AppendedData encoding="raw">\n_'
    i = data.index(key) + len(key)
    out = []
    for _ in range(3):
        n = int(np.frombuffer(data[i:i + 4], np.int32)[0])
        out.append(np.frombuffer(data[i + 4:i + 4 + n], np.float64).copy())
        i += 4 + n
    return out


def _tkcalc(tk, gx, gy, gz, x, y, z):
    """tkcalc1 (mark3mg.c:2844-2910) at points: the checker's interpolation
    of the device field, the eight weights summed in the reference's order"""
    nx, ny, nz = len(gx), len(gy), len(gz)

    def serch(g, v):
        m = np.searchsorted(g, v, side="right") - 1
        return np.clip(m, 0, len(g) - 2)
    m1, m2, m3 = serch(gx, x), serch(gy, y), serch(gz, z)
    dx, dy, dz = gx[m1 + 1] - gx[m1], gy[m2 + 1] - gy[m2], gz[m3 + 1] - gz[m3]
    s = np.zeros(len(x))
    for c in (0, 1):
        ddz = 1.0 - np.abs((gz[m3 + c] - z) / dz)
        for a in (0, 1):
            ddx = 1.0 - np.abs((gx[m1 + a] - x) / dx)
            for b in (0, 1):
                ddy = 1.0 - np.abs((gy[m2 + b] - y) / dy)
                s = s + ddx * ddy * ddz * tk[(m3 + c) * nx * ny + (m1 + a) * ny + m2 + b].astype(np.float64)
    return s


def test_generator_matches_reference(gpu):
    g = golden("g16_generator.npz")
    mark, t_ref = g["mark"], g["markt"]
    n = len(t_ref)
    gx, gy, gz = _grid_lines()
    print(f"case: generator g16, {n} markers, grid {len(gx)}x{len(gy)}x{len(gz)}", flush=True)
    rb, tb = I.deck_boxes(os.path.join(HERE, "g16_init.t3c"), *SIZES)
    M = I.Markers(gx, gy, gz, n)
    zero = np.zeros(n, np.float32)
    M.upload(mark[0], mark[1], mark[2], zero, zero, np.zeros(n, np.int8))
    ex, od, ow = M.paint(*SIZES, rb, zero, zero, zero)
    d = M.download()
    e = M.get_strain()
    assert np.array_equal(d["id"], np.arange(n))
    assert np.array_equal(d["type"], t_ref)
    # .prn marker records (load3mg.c:1037-1048): x y z k w d v e ex ft rr rn
    assert np.array_equal(ex, mark[8])
    assert np.array_equal(od, mark[5]) and np.array_equal(ow, mark[4])
    assert np.array_equal(np.abs(e), np.abs(mark[7]))
    imm = t_ref >= 100
    assert imm.any() and np.array_equal(d["tk"][imm], mark[3][imm])
    M.close()
    # node temperatures from the boxes, then the marker temperatures
    N = len(gx) * len(gy) * len(gz)
    tk = I.temperature_boxes(gx, gy, gz, *SIZES, tb, np.zeros(N, np.float32))
    inside = (~imm & (mark[0] > 0) & (mark[1] > 0) & (mark[2] > 0) & (mark[0] < SIZES[0]) & (mark[1] < SIZES[1]) &
              (mark[2] < SIZES[2]))
    got = _tkcalc(tk, gx, gy, gz, mark[0][inside].astype(np.float64), mark[1][inside].astype(np.float64),
                  mark[2][inside].astype(np.float64)).astype(np.float32)
    ref = mark[3][inside]
    rel = np.abs(got.astype(np.float64) - ref) / np.abs(ref)
    print(f"  marker T: max rel {rel.max():.2e}, exact {int((got == ref).sum())} of {inside.sum()}", flush=True)
    assert rel.max() <= 1e-6


def test_temperature_box_types_cover_the_deck(gpu):
    """every temperature box type of the deck changes the field where it
    should: a field built without the transition box differs only inside
    that box"""
    gx, gy, gz = _grid_lines()
    rb, tb = I.deck_boxes(os.path.join(HERE, "g16_init.t3c"), *SIZES)
    N = len(gx) * len(gy) * len(gz)
    print("case: temperature boxes with and without the transition box", flush=True)
    full = I.temperature_boxes(gx, gy, gz, *SIZES, tb, np.zeros(N, np.float32))
    keep = [b for b in tb if b.type != 5]
    part = I.temperature_boxes(gx, gy, gz, *SIZES, (I.TempBox * len(keep))(*keep), np.zeros(N, np.float32))
    diff = full != part
    assert diff.any()
    t5 = [b for b in tb if b.type == 5][0]
    xs = np.broadcast_to(gx[None, :, None], (len(gz), len(gx), len(gy))).ravel()
    assert np.all((xs[diff] >= min(t5.x0, t5.x5) - 1) & (xs[diff] <= max(t5.x2, t5.x7) + 1))
