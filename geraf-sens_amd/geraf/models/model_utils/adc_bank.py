"""Device-resident ADC bank of the RF training stages.

The stages split the antennas into ``bank_size`` interleaved groups
(``generate_series``: group i = antennas i, i + bank_size, ...;
rf_rendering.py:275-276), render one group per step with the tracer, and mix
its fresh signals with the cached signals of the other groups before the
matched filter (rf_rendering.py:652-684, 797-800).  The reference keeps the
cache as per-group CPU tensors, copies the rendered group to the host every
step and the other groups back to the device.  Here the cache is one [A, K]
tensor pair on the device; the rows of a group are updated in place and the
mix is an index_select, so no host copy happens on the step path.

Mix order (kept from the reference): the selected group's rows, then the
other groups' rows in ascending group order.  With the selected-rows matched
filter (``faster_matched_filter``) only the selected rows get an adjoint,
which is all the loss differentiates (the cached rows are constants).
"""
from __future__ import annotations

from typing import List, Tuple

import torch

from .matched_filter import base_matched_filter, faster_matched_filter


def generate_series(num_antennas: int, num_groups: int) -> List[torch.Tensor]:
    """Antenna indices of each group (rf_rendering.py:275-276)."""
    return [torch.arange(index, num_antennas, step=num_groups) for index in range(num_groups)]


class ADCBank:
    def __init__(self, num_antennas: int, num_samples: int, bank_size: int,
                 device: torch.device, dtype: torch.dtype = torch.float32):
        if bank_size < 1 or bank_size > num_antennas:
            raise ValueError(f"bank_size must be in [1, {num_antennas}], got {bank_size}")
        self.bank_size = bank_size
        self.groups = [g.to(device) for g in generate_series(num_antennas, bank_size)]
        self.real = torch.zeros((num_antennas, num_samples), device=device, dtype=dtype)
        self.imag = torch.zeros_like(self.real)
        self.chunk_id = 0
        # rows of every group but i, in ascending group order
        self._others = [torch.cat([self.groups[j] for j in range(bank_size) if j != i])
                        if bank_size > 1 else self.groups[0][:0] for i in range(bank_size)]

    @property
    def selected(self) -> torch.Tensor:
        return self.groups[self.chunk_id]

    @property
    def unselected(self) -> torch.Tensor:
        return self._others[self.chunk_id]

    def store(self, group: int, sim_real: torch.Tensor, sim_imag: torch.Tensor) -> None:
        """Cache a group's rendered signals (rows in the group's index order)."""
        idx = self.groups[group]
        self.real.index_copy_(0, idx, sim_real.detach().to(self.real.dtype))
        self.imag.index_copy_(0, idx, sim_imag.detach().to(self.imag.dtype))

    def advance(self) -> None:
        """Select the next group (rf_rendering.py:718-719)."""
        self.chunk_id = (self.chunk_id + 1) % self.bank_size

    def mix(self, sim_real: torch.Tensor, sim_imag: torch.Tensor, t_pos: torch.Tensor,
            r_pos: torch.Tensor) -> Tuple[torch.Tensor, ...]:
        """(mix_real, mix_imag, mix_t, mix_r) for the selected group's fresh
        signals; also refreshes that group's cache rows."""
        self.store(self.chunk_id, sim_real, sim_imag)
        sel, other = self.selected, self.unselected
        return (torch.cat([sim_real, self.real.index_select(0, other).to(sim_real.dtype)]),
                torch.cat([sim_imag, self.imag.index_select(0, other).to(sim_imag.dtype)]),
                torch.cat([t_pos.index_select(0, sel), t_pos.index_select(0, other)]),
                torch.cat([r_pos.index_select(0, sel), r_pos.index_select(0, other)]))

    def matched_filter(self, sim_real: torch.Tensor, sim_imag: torch.Tensor,
                       t_pos: torch.Tensor, r_pos: torch.Tensor, points: torch.Tensor,
                       radar_cfg: dict, selected_rows_only: bool = True):
        """Matched filter of the mixed bank onto points.  With
        ``selected_rows_only`` the adjoint runs for the selected rows alone
        (the same gradients the full adjoint gives those rows)."""
        self.store(self.chunk_id, sim_real, sim_imag)
        sel, other = self.selected, self.unselected
        if not selected_rows_only:
            mr, mi, mt, mrp = self.mix(sim_real, sim_imag, t_pos, r_pos)
            return base_matched_filter(mr, mi, mt, mrp, points, radar_cfg)
        return faster_matched_filter(
            sim_real, sim_imag, self.real.index_select(0, other).to(sim_real.dtype),
            self.imag.index_select(0, other).to(sim_imag.dtype), t_pos.index_select(0, sel),
            r_pos.index_select(0, sel), t_pos.index_select(0, other),
            r_pos.index_select(0, other), points, radar_cfg)
