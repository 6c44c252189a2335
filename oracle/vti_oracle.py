"""Oracle for the snapshot writers (test infrastructure: only tests/ import it).

A NumPy restatement of the reference's VTK writers:
- vti_bytes   : glide/io.py:20-131 (write_vti) - the file's bytes for NumPy
                fields: float32 conversion, flip along y, scalars then
                vectors (components interleaved), raw appended block with a
                UInt32 byte count before each array;
- pvd_text    : glide/io.py:183-202 (VTIWriter.write_pvd) - minidom's
                pretty-printed collection manifest;
- decode_vti  : an independent reader of such a file (header parsed as XML,
                each array read at its offset), used to pin vti_bytes.

Pinned in tests/test_vti.py by a hand-built known-answer file, by the decoder
round trip, and by a literal manifest.
"""
from __future__ import annotations

import struct
import xml.etree.ElementTree as ET
from xml.dom import minidom

import numpy as np


def vti_bytes(data: dict, dx, dy=None, origin=(0.0, 0.0), time_value=None, flip_y=True) -> bytes:
    """glide/io.py:20-131 for NumPy inputs (lists are vector fields)."""
    if dy is None:
        dy = dx
    scalars, vectors = {}, {}
    for name, value in data.items():                                   # io.py:50-62
        if isinstance(value, (list, tuple)):
            comps = [np.asarray(c).astype(np.float32) for c in value]
            vectors[name] = [np.flip(c, axis=0) for c in comps] if flip_y else comps
        else:
            a = np.asarray(value).astype(np.float32)
            scalars[name] = np.flip(a, axis=0) if flip_y else a
    first = next(iter(scalars.values())) if scalars else next(iter(vectors.values()))[0]
    ny, nx = first.shape                                               # io.py:64-69
    root = ET.Element("VTKFile", type="ImageData", version="1.0", byte_order="LittleEndian")
    img = ET.SubElement(root, "ImageData", WholeExtent=f"0 {nx-1} 0 {ny-1} 0 0",
                        Origin=f"{origin[0]} {origin[1]} 0", Spacing=f"{dx} {dy} 1.0")
    piece = ET.SubElement(img, "Piece", Extent=f"0 {nx-1} 0 {ny-1} 0 0")
    if time_value is not None:                                         # io.py:79-84
        fd = ET.SubElement(piece, "FieldData")
        da = ET.SubElement(fd, "DataArray", type="Float32", Name="TimeValue", NumberOfComponents="1",
                           format="ascii")
        da.text = f"\n{float(time_value)}\n"
    attrs = {}
    if scalars:                                                        # io.py:86-92
        attrs["Scalars"] = next(iter(scalars))
    for name, comps in vectors.items():
        if len(comps) == 3:
            attrs["Vectors"] = name
            break
    pd = ET.SubElement(piece, "PointData", **attrs)
    blocks, offset = [], 0
    for name, arr in scalars.items():                                  # io.py:97-107
        b = arr.ravel(order="C").tobytes()
        ET.SubElement(pd, "DataArray", type="Float32", Name=name, NumberOfComponents="1", format="appended",
                      offset=str(offset))
        blocks.append(b)
        offset += len(b) + 4
    for name, comps in vectors.items():                                # io.py:109-119
        b = np.stack(comps, axis=-1).astype(np.float32).ravel(order="C").tobytes()
        ET.SubElement(pd, "DataArray", type="Float32", Name=name, NumberOfComponents=str(len(comps)),
                      format="appended", offset=str(offset))
        blocks.append(b)
        offset += len(b) + 4
    text = ET.tostring(root, encoding="utf-8", xml_declaration=True).decode("utf-8")
    if text.endswith("</VTKFile>"):                                    # io.py:123-128
        text = text[: -len("</VTKFile>")]
    elif text.endswith("</VTKFile>\n"):
        text = text[: -len("</VTKFile>\n")]
    out = [text.encode("utf-8"), b'  <AppendedData encoding="raw">\n   _']
    for b in blocks:                                                   # io.py:130-136
        out.append(struct.pack("<I", len(b)))
        out.append(b)
    out.append(b"\n  </AppendedData>\n</VTKFile>\n")
    return b"".join(out)


def pvd_text(records) -> str:
    """glide/io.py:183-202: records = [(time, file name)]."""
    root = ET.Element("VTKFile", type="Collection", version="0.1", byte_order="LittleEndian")
    coll = ET.SubElement(root, "Collection")
    for t, fname in records:
        ET.SubElement(coll, "DataSet", timestep=str(t), group="", part="0", file=str(fname))
    return minidom.parseString(ET.tostring(root)).toprettyxml(indent="  ")


def decode_vti(blob: bytes) -> dict:
    """Read a file written as above: returns {"attrs": {...}, "time": t or
    None, "arrays": {name: float32 array (ny, nx) or (ny, nx, n_comp)}}.
    Independent of vti_bytes: the XML is parsed, every array is located by
    its offset attribute and checked against its UInt32 byte count."""
    marker = b'<AppendedData encoding="raw">\n   _'
    cut = blob.index(marker)
    header = blob[:cut].decode("utf-8").rstrip() + "</VTKFile>"
    data = blob[cut + len(marker):]
    if not blob.endswith(b"\n  </AppendedData>\n</VTKFile>\n"):
        raise ValueError("bad file tail")
    root = ET.fromstring(header.split("?>", 1)[1])
    img = root.find("ImageData")
    ext = [int(x) for x in img.get("WholeExtent").split()]
    nx, ny = ext[1] + 1, ext[3] + 1
    piece = img.find("Piece")
    tv = piece.find("FieldData/DataArray")
    pd = piece.find("PointData")
    arrays = {}
    for da in pd.findall("DataArray"):
        off, nc = int(da.get("offset")), int(da.get("NumberOfComponents"))
        (n,) = struct.unpack("<I", data[off:off + 4])
        if n != 4 * nc * ny * nx:
            raise ValueError(f"{da.get('Name')}: byte count {n}")
        a = np.frombuffer(data[off + 4: off + 4 + n], dtype="<f4").reshape((ny, nx) if nc == 1 else (ny, nx, nc))
        arrays[da.get("Name")] = a
    return {"attrs": {"WholeExtent": img.get("WholeExtent"), "Origin": img.get("Origin"),
                      "Spacing": img.get("Spacing"), **dict(pd.attrib)},
            "time": float(tv.text) if tv is not None else None, "arrays": arrays}
