"""Input preparation (counterpart of glide/data.py:220-473, the parts that do
not fetch remote data): the BedMachine thickness recomputation, observations
interpolated to the model grid, and the crop to a multigrid-compatible size.

Interpolation and thickness run on the GPU (libglide_hip.so); the crop is
index arithmetic on whatever arrays it is given.  The loaders that download
or open BedMachine / velocity mosaics / SMB files (glide/data.py:31-385) need
the network and xarray / rasterio / h5py, none of which this image has.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dev(x, device) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=device)


def _device_of(*xs) -> torch.device:
    for x in xs:
        if isinstance(x, torch.Tensor) and x.is_cuda:
            return x.device
    return torch.device("cuda")


def _direction(coords, name: str) -> int:
    """0 for strictly ascending, 1 for strictly descending coordinates (the
    interpolator's requirement, scipy RegularGridInterpolator)."""
    c = coords.detach().cpu().numpy() if isinstance(coords, torch.Tensor) else np.asarray(coords, dtype=np.float64)
    if c.ndim != 1 or c.size < 2:
        raise ValueError(f"{name} must be a 1-D array of at least 2 coordinates")
    d = np.diff(c)
    if np.all(d > 0):
        return 0
    if np.all(d < 0):
        return 1
    raise ValueError(f"{name} must be strictly ascending or descending")


def interpolate_to_grid(data, x_data, y_data, x_target, y_target, fill_value=0.0, dtype=torch.float32):
    """``data`` (len(y_data), len(x_data)) linearly interpolated to the grid
    (y_target x x_target) on the GPU; 0 outside the data's extent
    (glide/data.py:386-431: a linear RegularGridInterpolator with
    bounds_error=False, fill_value=0, evaluated at meshgrid(x_target,
    y_target), returned as float32).  Returns a device tensor of ``dtype``
    (float32 as the reference; float64 keeps the kernel's result)."""
    dev = _device_of(data, x_data, y_data, x_target, y_target)
    ydir, xdir = _direction(y_data, "y_data"), _direction(x_data, "x_data")
    d = _dev(data, dev)
    yd, xd = _dev(y_data, dev).reshape(-1), _dev(x_data, dev).reshape(-1)
    if tuple(d.shape) != (yd.numel(), xd.numel()):
        raise ValueError(f"data has shape {tuple(d.shape)}, the coordinates ({yd.numel()}, {xd.numel()})")
    yt, xt = _dev(y_target, dev).reshape(-1), _dev(x_target, dev).reshape(-1)
    if yt.numel() == 0 or xt.numel() == 0:
        raise ValueError("empty target grid")
    out = torch.empty((yt.numel(), xt.numel()), dtype=torch.float64, device=dev)
    _lib.call("glide_interp_grid", d.data_ptr(), d.shape[0], d.shape[1], yd.data_ptr(), ydir, xd.data_ptr(), xdir,
              yt.data_ptr(), yt.numel(), xt.data_ptr(), xt.numel(), float(fill_value), out.data_ptr(), _stream(dev))
    return out if dtype == torch.float64 else out.to(dtype)


def bedmachine_thickness(bed, surface, thklim=0.1) -> torch.Tensor:
    """Ice thickness from surface and bed as the BedMachine loader recomputes
    it (glide/data.py:257-258): base = max(bed, surface (1 - 1/(1 - 0.917))),
    thickness = surface - base + thklim, so floating ice is hydrostatic.  On
    the GPU in FP64."""
    dev = _device_of(bed, surface)
    b, s = _dev(bed, dev), _dev(surface, dev)
    if b.shape != s.shape:
        raise ValueError(f"bed {tuple(b.shape)} and surface {tuple(s.shape)} differ")
    out = torch.empty_like(b)
    _lib.call("glide_bedmachine_thickness", b.data_ptr(), s.data_ptr(), b.numel(), float(thklim), out.data_ptr(),
              _stream(dev))
    return out


def prepare_grid(geometry_data: dict, n_levels: int = 5) -> dict:
    """Crop to the centred sub-grid whose sizes are the largest multiples of
    2**n_levels (glide/data.py:433-473).  ``geometry_data`` holds x, y (1-D),
    bed, thickness, surface (y, x) and dx; the result adds ny and nx."""
    x, y = geometry_data["x"], geometry_data["y"]
    factor = 2 ** int(n_levels)
    nx_t = (len(x) // factor) * factor
    ny_t = (len(y) // factor) * factor
    if nx_t == 0 or ny_t == 0:
        raise ValueError(f"grid {len(y)} x {len(x)} is smaller than 2**{n_levels}")
    x0 = (len(x) - nx_t) // 2
    y0 = (len(y) - ny_t) // 2
    xs, ys = slice(x0, x0 + nx_t), slice(y0, y0 + ny_t)
    return {"x": x[xs], "y": y[ys], "bed": geometry_data["bed"][ys, xs],
            "thickness": geometry_data["thickness"][ys, xs], "surface": geometry_data["surface"][ys, xs],
            "dx": geometry_data["dx"], "ny": ny_t, "nx": nx_t}
