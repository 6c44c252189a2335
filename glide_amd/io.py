"""Snapshot output: VTK ImageData (.vti) files, PVD time-series manifests and
an HDF5 time series (counterpart of glide/io.py:14-282).

The files are the reference's, byte for byte: the same XML header, the arrays
as float32 with their UInt32 byte counts in one raw appended block (scalars
first, then vectors, in the order given), rows flipped along y by default.
What differs is where the bytes are made: the fields stay FP64 in HBM and one
kernel (libglide_hip.so glide_vti_pack) writes the whole appended block on
the GPU, so the host only copies it out once and writes it.  Host (NumPy)
arrays are uploaded first.
"""
from __future__ import annotations

import ctypes
import xml.etree.ElementTree as ET
from pathlib import Path
from xml.dom import minidom

import numpy as np
import torch

from . import _lib

APPENDED_OPEN = b'  <AppendedData encoding="raw">\n   _'
APPENDED_CLOSE = b"\n  </AppendedData>\n</VTKFile>\n"


def _is_vector(value) -> bool:
    return isinstance(value, (list, tuple))


def split_fields(data: dict):
    """(scalars, vectors) as lists of (name, [components]) in the order the
    file stores them: every scalar before every vector (glide/io.py:47-62)."""
    if not data:
        raise ValueError("no fields to write")
    scalars, vectors = [], []
    for name, value in data.items():
        if _is_vector(value):
            if not 1 <= len(value) <= 3:
                raise ValueError(f"vector field {name!r} needs 1 to 3 components")
            vectors.append((name, list(value)))
        else:
            scalars.append((name, [value]))
    return scalars, vectors


def vti_header(ny: int, nx: int, dx, dy, origin, time_value, scalars, vectors) -> bytes:
    """XML part of a .vti file (everything before the appended block).
    ``scalars`` = [name], ``vectors`` = [(name, n_components)]; the file
    holds the scalars first.  Element and attribute order follow
    glide/io.py:72-115; ElementTree serialises it with its XML declaration,
    and the closing </VTKFile> is left for after the appended data."""
    root = ET.Element("VTKFile", type="ImageData", version="1.0", byte_order="LittleEndian")
    image = ET.SubElement(root, "ImageData", WholeExtent=f"0 {nx - 1} 0 {ny - 1} 0 0",
                          Origin=f"{origin[0]} {origin[1]} 0", Spacing=f"{dx} {dy} 1.0")
    piece = ET.SubElement(image, "Piece", Extent=f"0 {nx - 1} 0 {ny - 1} 0 0")
    if time_value is not None:
        fd = ET.SubElement(piece, "FieldData")
        tv = ET.SubElement(fd, "DataArray", type="Float32", Name="TimeValue", NumberOfComponents="1", format="ascii")
        tv.text = f"\n{float(time_value)}\n"
    attrs = {}
    if scalars:
        attrs["Scalars"] = scalars[0]
    three = [name for name, nc in vectors if nc == 3]
    if three:
        attrs["Vectors"] = three[0]
    pd = ET.SubElement(piece, "PointData", **attrs)
    offset = 0
    for name, nc in [(n, 1) for n in scalars] + list(vectors):
        ET.SubElement(pd, "DataArray", type="Float32", Name=name, NumberOfComponents=str(nc), format="appended",
                      offset=str(offset))
        offset += 4 * nc * ny * nx + 4
    text = ET.tostring(root, encoding="utf-8", xml_declaration=True).decode("utf-8")
    for tail in ("</VTKFile>", "</VTKFile>\n"):
        if text.endswith(tail):
            text = text[: -len(tail)]
            break
    return text.encode("utf-8")


def _device(x, device) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=device)


def _pick_device(data: dict) -> torch.device:
    for value in data.values():
        for c in (value if _is_vector(value) else [value]):
            if isinstance(c, torch.Tensor) and c.is_cuda:
                return c.device
    return torch.device("cuda")


def pack_block(fields, ny: int, nx: int, flip_y: bool, device) -> np.ndarray:
    """The appended block of ``fields`` = [[component tensors]] (file order),
    packed on the GPU and copied to host memory (uint8)."""
    arr = (_lib.GlideVtiArray * len(fields))()
    for a, comps in enumerate(fields):
        arr[a].n_comp = len(comps)
        for q, c in enumerate(comps):
            arr[a].comp[q] = c.data_ptr()
    nbytes = _lib.fn("glide_vti_block_bytes")(arr, len(fields), ny, nx)
    if nbytes == 0:
        raise ValueError(f"invalid snapshot layout ({len(fields)} arrays of {ny} x {nx})")
    block = torch.empty(nbytes, dtype=torch.uint8, device=device)
    stream = torch.cuda.current_stream(device)
    _lib.call("glide_vti_pack", arr, len(fields), ny, nx, int(bool(flip_y)), ctypes.c_void_p(block.data_ptr()),
              nbytes, ctypes.c_void_p(stream.cuda_stream))
    host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
    host.copy_(block, non_blocking=True)
    stream.synchronize()
    return host.numpy()


def write_vti(filename, data: dict, dx, dy=None, origin=(0.0, 0.0), time_value=None, flip_y=True) -> Path:
    """Write ``data`` (name -> field, or name -> list of 1-3 component
    fields) as one ImageData snapshot (glide/io.py:20-131).  Fields are
    (ny, nx) torch tensors (any device; FP64 on the GPU is the fast path) or
    NumPy arrays; all must share one shape."""
    if dy is None:
        dy = dx
    path = Path(filename)
    path.parent.mkdir(parents=True, exist_ok=True)
    scalars, vectors = split_fields(data)
    device = _pick_device(data)
    fields, shape = [], None
    for name, comps in scalars + vectors:
        dev = [_device(c, device) for c in comps]
        for t in dev:
            if t.ndim != 2:
                raise ValueError(f"field {name!r} is not two-dimensional")
            if shape is None:
                shape = tuple(t.shape)
            elif tuple(t.shape) != shape:
                raise ValueError(f"field {name!r} has shape {tuple(t.shape)}, expected {shape}")
        fields.append(dev)
    ny, nx = shape
    header = vti_header(ny, nx, dx, dy, origin, time_value, [n for n, _ in scalars],
                        [(n, len(c)) for n, c in vectors])
    block = pack_block(fields, ny, nx, flip_y, device)
    with open(path, "wb") as f:
        f.write(header)
        f.write(APPENDED_OPEN)
        f.write(memoryview(block))
        f.write(APPENDED_CLOSE)
    return path


def pvd_text(records) -> str:
    """PVD manifest for ``records`` = [(time, file name)] (glide/io.py:183-202),
    in minidom's pretty-printed form."""
    root = ET.Element("VTKFile", type="Collection", version="0.1", byte_order="LittleEndian")
    coll = ET.SubElement(root, "Collection")
    for t, fname in records:
        ET.SubElement(coll, "DataSet", timestep=str(t), group="", part="0", file=str(fname))
    return minidom.parseString(ET.tostring(root)).toprettyxml(indent="  ")


class VTIWriter:
    """A numbered series of .vti snapshots plus its .pvd manifest
    (glide/io.py:134-202): ``write_step(i, t, data)`` writes
    ``{base}_{i:04d}.vti``; ``write_pvd()`` lists every step written."""

    def __init__(self, out_dir, base="output", dx=1.0, dy=None, origin=(0.0, 0.0), flip_y=True):
        self.out_dir = Path(out_dir)
        self.out_dir.mkdir(parents=True, exist_ok=True)
        self.base = base
        self.dx = dx
        self.dy = dy if dy is not None else dx
        self.origin = origin
        self.flip_y = flip_y
        self.records: list = []

    def write_step(self, step_idx, time_value, data) -> Path:
        fname = f"{self.base}_{step_idx:04d}.vti"
        path = write_vti(self.out_dir / fname, data, self.dx, self.dy, self.origin, time_value=time_value,
                         flip_y=self.flip_y)
        self.records.append((float(time_value), fname))
        return path

    def write_pvd(self, pvd_name=None) -> Path:
        name = pvd_name if pvd_name is not None else f"{self.base}.pvd"
        path = self.out_dir / name
        with open(path, "w") as f:
            f.write(pvd_text(self.records))
        return path


class HDF5Writer:
    """A time series in one HDF5 file, one group ``t{k:04d}`` per step with a
    ``time`` attribute and one float32 dataset per field (vector fields as
    ``name_x``, ``name_y``, ``name_z``), file attributes dx, dy, origin_x,
    origin_y (glide/io.py:205-282).  Needs h5py, which this image does not
    ship: constructing one raises ImportError here, and the layout is not
    exercised by the tests."""

    def __init__(self, filename, dx=1.0, dy=None, origin=(0.0, 0.0), compression="lzf", compression_opts=None):
        try:
            import h5py
        except ImportError as e:  # pragma: no cover - depends on the image
            raise ImportError("HDF5Writer needs h5py, which is not installed") from e
        self._h5py = h5py
        self.filename = Path(filename)
        self.filename.parent.mkdir(parents=True, exist_ok=True)
        self.dx = dx
        self.dy = dy if dy is not None else dx
        self.origin = origin
        self.compression, self.compression_opts = compression, compression_opts
        self.step_count = 0
        with h5py.File(self.filename, "w") as f:
            f.attrs["dx"] = dx
            f.attrs["dy"] = self.dy
            f.attrs["origin_x"] = origin[0]
            f.attrs["origin_y"] = origin[1]

    def write_step(self, time_value, data) -> str:
        group = f"t{self.step_count:04d}"
        with self._h5py.File(self.filename, "a") as f:
            grp = f.create_group(group)
            grp.attrs["time"] = float(time_value)
            for name, value in data.items():
                comps = list(value) if _is_vector(value) else [value]
                names = [name + s for s in ("_x", "_y", "_z")[: len(comps)]] if _is_vector(value) else [name]
                for nm, c in zip(names, comps):
                    arr = (c.detach().cpu().numpy() if isinstance(c, torch.Tensor) else np.asarray(c))
                    grp.create_dataset(nm, data=arr.astype(np.float32), compression=self.compression,
                                       compression_opts=self.compression_opts)
        self.step_count += 1
        return group

    def close(self):
        pass

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        self.close()
        return False
