"""Snapshot writers on the CPU: the oracle (oracle/vti_oracle.py) pinned by a
hand-built file, its own decoder and a literal manifest; the product's host
half (XML header, manifest) against the oracle; and the product refusing to
write without the GPU path (no host fallback)."""
import struct

import numpy as np
import pytest

from glide_amd import io as gio
from oracle import vti_oracle as V


def test_oracle_known_answer_file():
    # 2 x 3 scalar field, flipped: the file's rows are source rows 1 then 0
    a = np.array([[0.0, 1.0, 2.0], [3.0, 4.0, 5.0]])
    got = V.vti_bytes({"thk": a}, 2.0)
    want = (b"<?xml version='1.0' encoding='utf-8'?>\n"
            b'<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian">'
            b'<ImageData WholeExtent="0 2 0 1 0 0" Origin="0.0 0.0 0" Spacing="2.0 2.0 1.0">'
            b'<Piece Extent="0 2 0 1 0 0"><PointData Scalars="thk">'
            b'<DataArray type="Float32" Name="thk" NumberOfComponents="1" format="appended" offset="0" />'
            b"</PointData></Piece></ImageData>"
            b'  <AppendedData encoding="raw">\n   _'
            + struct.pack("<I", 24) + struct.pack("<6f", 3, 4, 5, 0, 1, 2)
            + b"\n  </AppendedData>\n</VTKFile>\n")
    assert got == want


def test_oracle_known_answer_vector_and_time():
    u = np.array([[1.0, 2.0]])
    v = np.array([[-1.0, 0.5]])
    got = V.vti_bytes({"vel": [u, v]}, 10, time_value=3, flip_y=False)
    want = (b"<?xml version='1.0' encoding='utf-8'?>\n"
            b'<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian">'
            b'<ImageData WholeExtent="0 1 0 0 0 0" Origin="0.0 0.0 0" Spacing="10 10 1.0">'
            b'<Piece Extent="0 1 0 0 0 0"><FieldData>'
            b'<DataArray type="Float32" Name="TimeValue" NumberOfComponents="1" format="ascii">\n3.0\n</DataArray>'
            b"</FieldData><PointData>"
            b'<DataArray type="Float32" Name="vel" NumberOfComponents="2" format="appended" offset="0" />'
            b"</PointData></Piece></ImageData>"
            b'  <AppendedData encoding="raw">\n   _'
            + struct.pack("<I", 16) + struct.pack("<4f", 1.0, -1.0, 2.0, 0.5)
            + b"\n  </AppendedData>\n</VTKFile>\n")
    assert got == want


def _mixed(rng, ny=5, nx=7):
    return {"vel": [rng.normal(size=(ny, nx)), rng.normal(size=(ny, nx))],
            "thk": rng.uniform(0, 3000, size=(ny, nx)),
            "w": [rng.normal(size=(ny, nx)) for _ in range(3)],
            "srf": rng.uniform(-500, 3000, size=(ny, nx))}


@pytest.mark.parametrize("flip", [True, False])
def test_oracle_decodes_to_its_inputs(flip):
    rng = np.random.default_rng(2)
    data = _mixed(rng)
    blob = V.vti_bytes(data, 1500.0, origin=(-3.5, 2), time_value=12.5, flip_y=flip)
    d = V.decode_vti(blob)
    f = (lambda a: np.flip(a, 0)) if flip else (lambda a: a)
    assert list(d["arrays"]) == ["thk", "srf", "vel", "w"]          # scalars first, then vectors
    assert d["attrs"]["Scalars"] == "thk" and d["attrs"]["Vectors"] == "w"
    assert d["attrs"]["Origin"] == "-3.5 2 0" and d["attrs"]["Spacing"] == "1500.0 1500.0 1.0"
    assert d["time"] == 12.5
    assert np.array_equal(d["arrays"]["thk"], f(data["thk"]).astype(np.float32))
    assert np.array_equal(d["arrays"]["srf"], f(data["srf"]).astype(np.float32))
    assert np.array_equal(d["arrays"]["vel"], np.stack([f(c) for c in data["vel"]], -1).astype(np.float32))
    assert np.array_equal(d["arrays"]["w"], np.stack([f(c) for c in data["w"]], -1).astype(np.float32))


def test_oracle_manifest_literal():
    got = V.pvd_text([(1.0, "out_0001.vti"), (2.5, "out_0002.vti")])
    want = ('<?xml version="1.0" ?>\n'
            '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
            '  <Collection>\n'
            '    <DataSet timestep="1.0" group="" part="0" file="out_0001.vti"/>\n'
            '    <DataSet timestep="2.5" group="" part="0" file="out_0002.vti"/>\n'
            '  </Collection>\n'
            '</VTKFile>\n')
    assert got == want


@pytest.mark.parametrize("time_value", [None, 7.25])
def test_product_header_and_manifest_match_oracle(time_value):
    rng = np.random.default_rng(3)
    data = _mixed(rng, 4, 6)
    blob = V.vti_bytes(data, 250.0, dy=125.0, origin=(1.0, -2.0), time_value=time_value)
    scalars, vectors = gio.split_fields(data)
    head = gio.vti_header(4, 6, 250.0, 125.0, (1.0, -2.0), time_value, [n for n, _ in scalars],
                          [(n, len(c)) for n, c in vectors])
    assert blob.startswith(head + gio.APPENDED_OPEN)
    assert blob.endswith(gio.APPENDED_CLOSE)
    recs = [(0.0, "a_0000.vti"), (10.0, "a_0001.vti")]
    assert gio.pvd_text(recs) == V.pvd_text(recs)


def test_product_rejects_bad_inputs_before_any_device_work():
    with pytest.raises(ValueError):
        gio.split_fields({})
    with pytest.raises(ValueError):
        gio.split_fields({"v": [np.zeros((2, 2))] * 4})
