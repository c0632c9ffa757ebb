"""The initial-condition deck reader and the ParaView writer against the
reference's own generator run (golden g16: in3mg on g16_init.t3c, its
.prn state and the VTR files its saver_paraview wrote).  CPU only: both
are host code."""
import os

import numpy as np

import i3elvis_amd as I
from conftest import golden

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIZES = (2.4e5, 1.2e5, 1.8e5)


def _vtr_blocks(path):
    """header text and the raw appended blocks (32-bit length prefixed)"""
    data = open(path, "rb").read()
    key = b'<AppendedData encoding="raw">\n_'
    i = data.index(key) + len(key)
    head = data[:i]
    blocks = []
    while data[i:i + 2] != b"  ":
        n = int(np.frombuffer(data[i:i + 4], np.int32)[0])
        blocks.append(data[i + 4:i + 4 + n])
        i += 4 + n
    return head, blocks, data[i:]


def test_nodal_vtr_is_the_reference_bytes(tmp_path):
    """the reference's nodal file rewritten from its own grid lines and
    seven fields (x fastest in the file, back to the node order for the
    writer): equal to the byte"""
    ref = os.path.join(HERE, "g16_nodal.vtr")
    head, blocks, tail = _vtr_blocks(ref)
    gx, gy, gz = (np.frombuffer(b, np.float64) for b in blocks[:3])
    nx, ny, nz = len(gx), len(gy), len(gz)
    fields = {}
    names = ("viscosity", "temperature", "density", "vy", "vx", "vz", "pressure")
    assert len(blocks) == 3 + len(names)
    for name, b in zip(names, blocks[3:]):
        dt = np.float64 if name == "pressure" else np.float32
        pv = np.frombuffer(b, dt).reshape(nz, ny, nx)
        fields[name] = np.ascontiguousarray(pv.transpose(0, 2, 1)).ravel()  # (k, i, j): node m3*nx*ny + m1*ny + m2
    g = golden("g16_generator.npz")
    assert np.array_equal(np.asarray(g["gx"], np.float32), gx.astype(np.float32))
    out = tmp_path / "g16.vtr"
    I.write_vtr_nodal(out, gx, gy, gz, fields)
    assert open(out, "rb").read() == open(ref, "rb").read()


def test_composition_vtr_is_the_reference_bytes(tmp_path):
    """the composition map of the generator's markers (claim rule and
    neighbour fill) on the 1.5x grid: the reference's file to the byte"""
    g = golden("g16_generator.npz")
    mark = g["mark"]
    nx, ny, nz = len(g["gx"]), len(g["gy"]), len(g["gz"])
    mx, my, mz = int(1.5 * (nx - 1)), int(1.5 * (ny - 1)), int(1.5 * (nz - 1))
    out = tmp_path / "comp.vtr"
    I.write_vtr_composition(out, *SIZES, mx, my, mz, mark[0], mark[1], mark[2], g["markt"])
    ref = open(os.path.join(HERE, "g16_comp.vtr"), "rb").read()
    got = open(out, "rb").read()
    assert got == ref
    _, blocks, _ = _vtr_blocks(os.path.join(HERE, "g16_comp.vtr"))
    compo = np.frombuffer(blocks[3], np.int32)
    assert len(np.unique(compo)) > 5


def test_pvd_collection(tmp_path):
    p = tmp_path / "run.pvd"
    I.pvd_open(p)
    I.pvd_append(p, 0.0, "a.vtr")
    I.pvd_append(p, 1.5e3, "b.vtr", "out")
    txt = open(p).read().splitlines()
    assert txt == ['<?xml version="1.0"?>', '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">',
                   "<Collection>", '  <DataSet timestep="0.000000e+00" file="./a.vtr"/>',
                   '  <DataSet timestep="1.500000e+03" file="out/b.vtr"/>', "</Collection>", "</VTKFile>"]


def test_deck_boxes_read_as_the_generator_reads_them():
    """rock and temperature boxes of g16_init.t3c: counts, shapes, the 'm'
    (metres) and fraction conversions, the shell parameters"""
    rb, tb = I.deck_boxes(os.path.join(HERE, "g16_init.t3c"), *SIZES)
    assert len(rb) == 10 and len(tb) == 6
    assert [b.shape for b in rb] == [0, 0, 0, 0, 0, 0, 1, 2, 0, 0]
    assert [b.type for b in rb] == [10, 9, 5, 2, 12, 7, 11, 13, 114, 8]
    assert rb[1].y[1] == 62000 / 1.2e5 and rb[1].x[2] == 1.0 and rb[1].ex == 0.05 and rb[1].e == 0.10
    assert rb[6].x[0] == 0.62 and rb[6].x[5] == 0.09 and rb[6].z[7] == 0.9
    assert [b.type for b in tb] == [0, 4, 4, 5, 2, 3]
    assert tb[1].x2 == 0.55 * 2.4e5 and tb[1].y1 == 80000.0 and tb[1].t[6] == 1e-6
    assert tb[4].x2 == 6000.0 and tb[4].x0 == 0.62 * 2.4e5 and tb[4].x7 == 1.0 * 2.4e5
