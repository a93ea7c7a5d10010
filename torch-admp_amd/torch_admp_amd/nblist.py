"""Neighbour list with the reference call surface (src/torch_admp/nblist.py:268-660).

``TorchNeighborList(cutoff)(positions, box) -> pairs (P, 2)`` plus the cached
``get_pairs / get_ds / get_buffer_scales`` and their setters.  The pair search
runs in the HIP library (cell list in fractional space, both orderings, the
reference criterion 0.01 < d < rc); ``ds`` comes from the PairDistance kernel
and stays differentiable in positions and box.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from .ops import PairDistance


def check_cutoff(box: torch.Tensor, cutoff: float) -> None:
    """Assert rc <= half the smallest face height (src/torch_admp/nblist.py:620-660)."""
    b = box.detach().to("cpu", torch.float64).reshape(3, 3)
    normals = torch.stack([torch.cross(b[1], b[2], dim=-1), torch.cross(b[2], b[0], dim=-1),
                           torch.cross(b[0], b[1], dim=-1)])
    unit = normals / torch.norm(normals, dim=1, keepdim=True)
    heights = torch.abs((b * unit).sum(1))
    min_half_height = torch.min(heights) / 2
    assert cutoff <= min_half_height, (
        f"Cutoff {cutoff} is larger than half the minimum height {min_half_height} of the box. "
        "This may lead to unphysical results.")


def build_pairs(positions: torch.Tensor, box: Optional[torch.Tensor], cutoff: float) -> torch.Tensor:
    """Ordered pair list (both orderings) from the HIP cell list.  One stream
    synchronisation returns the pair count."""
    _lib.require_gpu(positions, box)
    lib = _lib.load()
    pos = positions.detach().to(torch.float64).contiguous()
    n = pos.shape[0]
    boxh = _lib.box_array(box)
    nbytes = ctypes.c_size_t()
    _lib.check(lib.admp_nblist_workspace_bytes(n, boxh, float(cutoff), ctypes.byref(nbytes)),
               "admp_nblist_workspace_bytes")
    ws = _lib.workspace(nbytes.value, pos.device)
    stream = _lib.stream_of(pos)
    npairs = ctypes.c_int64()
    _lib.check(lib.admp_nblist_count(_lib.ptr(pos), n, boxh, float(cutoff), _lib.ptr(ws),
                                     nbytes.value, ctypes.byref(npairs), stream),
               "admp_nblist_count")
    pairs = torch.empty(npairs.value, 2, dtype=torch.int64, device=pos.device)
    _lib.check(lib.admp_nblist_fill(_lib.ptr(pos), n, boxh, float(cutoff), _lib.ptr(ws),
                                    nbytes.value, _lib.ptr(pairs), npairs.value, stream),
               "admp_nblist_fill")
    return pairs


class TorchNeighborList(torch.nn.Module):
    """HIP-backed drop-in for the reference TorchNeighborList
    (src/torch_admp/nblist.py:268-333, 553-617)."""

    def __init__(self, cutoff: float) -> None:
        super().__init__()
        self.cutoff = cutoff
        self.pairs = torch.empty(1, dtype=torch.long)
        self.buffer_scales = torch.empty(1, dtype=torch.long)
        self.ds = torch.empty(1, dtype=torch.float64)

    def forward(self, positions: torch.Tensor, box: Optional[torch.Tensor] = None) -> torch.Tensor:
        if box is None:
            pairs = self.forward_obc(positions)
            pbc_flag = False
        else:
            check_cutoff(box, self.cutoff)
            pairs = self.forward_pbc(positions, box)
            pbc_flag = True
        self.pairs = pairs
        self.buffer_scales = self.pairs_buffer_scales(pairs)
        self.ds = self.pairs_ds(positions, pairs, box, pbc_flag)
        return pairs

    def forward_pbc(self, positions: torch.Tensor, box: torch.Tensor) -> torch.Tensor:
        """Periodic pairs with 0.01 < d < rc (src/torch_admp/nblist.py:335-445)."""
        return build_pairs(positions, box, self.cutoff)

    def forward_obc(self, positions: torch.Tensor) -> torch.Tensor:
        """Open-boundary pairs with d < rc, i != j (src/torch_admp/nblist.py:472-490)."""
        return build_pairs(positions, None, self.cutoff)

    @staticmethod
    def pairs_buffer_scales(pairs: torch.Tensor) -> torch.Tensor:
        """1 where i < j else 0 (src/torch_admp/nblist.py:492-515)."""
        return (pairs[:, 0] < pairs[:, 1]).to(torch.long)

    @staticmethod
    def pairs_ds(positions: torch.Tensor, pairs: torch.Tensor, box: Optional[torch.Tensor] = None,
                 pbc_flag: bool = True) -> torch.Tensor:
        """Differentiable pair distances (src/torch_admp/nblist.py:517-551)."""
        if pbc_flag:
            assert box is not None, "Box should be provided for periodic system."
            return PairDistance.apply(positions, box, pairs)
        return PairDistance.apply(positions, None, pairs)

    def set_pairs(self, pairs: torch.Tensor) -> None:
        self.pairs = pairs

    def set_buffer_scales(self, buffer_scales: torch.Tensor) -> None:
        self.buffer_scales = buffer_scales

    def set_ds(self, ds: torch.Tensor) -> None:
        self.ds = ds

    def get_pairs(self) -> torch.Tensor:
        return self.pairs

    def get_buffer_scales(self) -> torch.Tensor:
        return self.buffer_scales

    def get_ds(self) -> torch.Tensor:
        return self.ds
