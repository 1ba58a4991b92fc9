"""Matched-filter dispatch helpers (drop-in for
geraf/models/model_utils/matched_filter.py:7-108)."""
from typing import Dict

import torch


def base_matched_filter(real: torch.Tensor, image: torch.Tensor, t_pos: torch.Tensor,
                        r_pos: torch.Tensor, points: torch.Tensor, radar_cfg: Dict[str, float],
                        backend: str = "cuda"):
    """Backproject echoes [A,K] onto points [N,3]; returns (sar_real, sar_image) [N]."""
    if backend != "cuda":
        raise ValueError(f"Unsupported matched filter backend: {backend!r}. "
                         "Only 'cuda' is retained.")
    from geraf.ops.RFMatchedFilter.matched_filter_cuda import matched_filter_cuda

    return matched_filter_cuda(real, image, t_pos, r_pos, points, t_pos.shape[0],
                               radar_cfg["numADCSample"], radar_cfg["adcSampleRate"],
                               radar_cfg["As_sci"], points.shape[0])


def faster_matched_filter(real_sel, image_sel, real_unsel, image_unsel, tpos_sel, rpos_sel,
                          tpos_unsel, rpos_unsel, points, radar_cfg: Dict[str, float],
                          backend: str = "cuda"):
    """Matched filter over a selected and an unselected antenna bank; only the
    selected echoes receive gradients."""
    if backend != "cuda":
        raise ValueError(f"Unsupported faster matched filter backend: {backend!r}. "
                         "Only 'cuda' is retained.")
    from geraf.ops.RFMatchedFilter.faster_matched_filter_cuda import faster_matched_filter_cuda

    num_sel = tpos_sel.shape[0]
    total = num_sel + tpos_unsel.shape[0]
    return faster_matched_filter_cuda(real_sel, image_sel, real_unsel, image_unsel, tpos_sel,
                                      rpos_sel, tpos_unsel, rpos_unsel, points, total, num_sel,
                                      radar_cfg["numADCSample"], radar_cfg["adcSampleRate"],
                                      radar_cfg["As_sci"], points.shape[0])
