"""Matched-filter volumes on disk and on the device: the per-frame
``sarimage.bin`` cube and the accumulated ``<obj>_sub<N>.npy`` volume
(docs/PrepareData.md:46-66), the grid they live on (``grid_num``,
geraf/utils/matched_filter/cuda_tools.py:4-29), the loader's trilinear lookup
at target points (geraf/datasets/transforms/loading.py:19-25,28-61,79-100), and
a producer that backprojects a frame's ADC data into such a cube with the
MI355X matched filter.

Format: float32, C order over (Nx, Ny, Nz), no header; the grid spans
``bound = [xmin, xmax, ymin, ymax, zmin, zmax]`` with ``grid_num`` points per
axis (both ends included).  The stored value is the magnitude of the complex
matched-filter output.
"""
from __future__ import annotations

import os
from typing import Optional, Sequence, Tuple

import numpy as np
import torch
import torch.nn.functional as F

from geraf.ops.RFMatchedFilter import matched_filter_cuda


def grid_num(max_val: float, min_val: float, res: float) -> int:
    """Grid points from min_val to max_val at spacing res, both ends included."""
    if res <= 0:
        raise ValueError("Resolution 'res' must be greater than zero.")
    return int(round((max_val - min_val) / res)) + 1


def grid_shape(bound: Sequence[float], mf_res: Sequence[float]) -> Tuple[int, int, int]:
    return (grid_num(bound[1], bound[0], mf_res[0]), grid_num(bound[3], bound[2], mf_res[1]),
            grid_num(bound[5], bound[4], mf_res[2]))


def grid_axes(bound: Sequence[float], mf_res: Sequence[float]):
    """The three axis coordinate vectors (float64) of the grid."""
    shape = grid_shape(bound, mf_res)
    return tuple(np.linspace(bound[2 * i], bound[2 * i + 1], shape[i]) for i in range(3))


def write_sarimage(path: str, volume) -> None:
    """Write a real [Nx, Ny, Nz] volume as raw float32 in C order."""
    v = volume.detach().cpu().numpy() if torch.is_tensor(volume) else np.asarray(volume)
    if v.ndim != 3:
        raise ValueError(f"expected a 3-D volume, got shape {v.shape}")
    np.ascontiguousarray(v, dtype=np.float32).tofile(path)


def read_sarimage(path: str, bound: Sequence[float], mf_res: Sequence[float],
                  trans_power: Optional[float] = None) -> np.ndarray:
    """Read a sarimage.bin cube; divided by trans_power when it is given
    (LoadRF with normalization=True)."""
    shape = grid_shape(bound, mf_res)
    v = np.fromfile(path, dtype=np.float32)
    if v.size != shape[0] * shape[1] * shape[2]:
        raise ValueError(f"{path}: {v.size} floats, grid {shape} needs "
                         f"{shape[0] * shape[1] * shape[2]}")
    v = v.reshape(shape)
    if trans_power is not None:
        v /= trans_power
    return v


def accumulate(volumes) -> np.ndarray:
    """The accumulated volume (<obj>_sub<N>.npy): the sum of the frames' cubes."""
    it = iter(volumes)
    acc = np.array(next(it), dtype=np.float32, copy=True)
    for v in it:
        acc += v
    return acc


def sample_volume(volume, points, mode: str = "bilinear", align_corners: bool = True):
    """Volume values at points given in the normalised [-1, 1]^3 frame of the
    grid, as the loader's `_sample_volume`: grid_sample with the (z, y, x)
    coordinate order over the [Nx, Ny, Nz] tensor.  Works on any device."""
    vol = torch.as_tensor(volume, dtype=torch.float32)
    pts = torch.as_tensor(points, dtype=torch.float32, device=vol.device).reshape(-1, 3)
    grid = pts[:, [2, 1, 0]].reshape(1, 1, 1, -1, 3)
    out = F.grid_sample(vol[None, None], grid, mode=mode, align_corners=align_corners)
    return out.reshape(-1)


def backproject(real: torch.Tensor, imag: torch.Tensor, t_pos: torch.Tensor, r_pos: torch.Tensor,
                bound: Sequence[float], mf_res: Sequence[float], radar_cfg: dict,
                slab_voxels: int = 1 << 22, algo: str = "tiled",
                tol: float = 1e-6) -> torch.Tensor:
    """|matched filter| of one frame over the grid, [Nx, Ny, Nz] float32 on the
    echo's device.  The grid is evaluated in slabs of x-planes of about
    `slab_voxels` voxels, so the voxel coordinates never occupy more than a
    slab's worth of memory.  ``algo``: "tiled" (the per-tile expansion form,
    truncation bound ``tol`` of sum |s| per antenna; a grid tiles perfectly)
    or "direct" (the drop-in op's kernels)."""
    if algo not in ("tiled", "direct"):
        raise ValueError(f"unknown matched-filter algorithm {algo!r}")
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import tiled_forward

    nx, ny, nz = grid_shape(bound, mf_res)
    dev = real.device
    ax = [torch.tensor(a, dtype=torch.float32, device=dev) for a in grid_axes(bound, mf_res)]
    out = torch.empty((nx, ny, nz), dtype=torch.float32, device=dev)
    A, K = real.shape
    planes = max(1, slab_voxels // (ny * nz))
    for x0 in range(0, nx, planes):
        x1 = min(nx, x0 + planes)
        gx, gy, gz = torch.meshgrid(ax[0][x0:x1], ax[1], ax[2], indexing="ij")
        vox = torch.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)], dim=1)
        if algo == "tiled":
            re, im, _ = tiled_forward(real, imag, t_pos, r_pos, vox, K,
                                      radar_cfg["adcSampleRate"], radar_cfg["As_sci"], tol)
        else:
            re, im = matched_filter_cuda(real, imag, t_pos, r_pos, vox, A, K,
                                         radar_cfg["adcSampleRate"], radar_cfg["As_sci"],
                                         vox.shape[0])
        out[x0:x1] = torch.hypot(re, im).reshape(x1 - x0, ny, nz)
    return out


def produce_frame(frame_dir: str, bound: Sequence[float], mf_res: Sequence[float],
                  radar_cfg: dict, device: str = "cuda", algo: str = "tiled") -> np.ndarray:
    """Backproject a frame folder (adcr.npy, adci.npy [M, K]; tpos.npy, rpos.npy
    [M, 3]) into its sarimage.bin and return the cube."""
    load = lambda n: torch.tensor(np.load(os.path.join(frame_dir, n)).astype(np.float32),  # noqa
                                  device=device)
    real, imag = load("adcr.npy"), load("adci.npy")
    real = real.reshape(real.shape[0], -1)
    imag = imag.reshape(imag.shape[0], -1)
    t_pos, r_pos = load("tpos.npy").reshape(-1, 3), load("rpos.npy").reshape(-1, 3)
    vol = backproject(real, imag, t_pos, r_pos, bound, mf_res, radar_cfg, algo=algo)
    write_sarimage(os.path.join(frame_dir, "sarimage.bin"), vol)
    return vol.cpu().numpy()
