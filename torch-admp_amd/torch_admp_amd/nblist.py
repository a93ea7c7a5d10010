"""Neighbour list with the reference call surface (src/torch_admp/nblist.py:268-660).

``TorchNeighborList(cutoff)(positions, box) -> pairs (P, 2)`` plus the cached
``get_pairs / get_ds / get_buffer_scales`` and their setters.  The pair search
runs in the HIP library (cell list in fractional space, both orderings, the
reference criterion 0.01 < d < rc); ``ds`` comes from the pair-distance kernel
and stays differentiable in positions and box.
"""

import math
from typing import List, Optional

import torch

from . import _lib
from . import _ops  # noqa: F401  (registers torch.ops.admp)


def _min_half_height(rows: List[float]) -> float:
    """Half the smallest distance between opposite faces of the cell with
    rows a, b, c (row-major floats): V / |b x c| etc."""
    a, b, c = rows[0:3], rows[3:6], rows[6:9]

    def cross(u: List[float], v: List[float]) -> List[float]:
        return [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]

    heights: List[float] = []
    for x, (u, v) in zip((a, b, c), ((b, c), (c, a), (a, b))):
        nrm = cross(u, v)
        length = math.sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2])
        heights.append(abs(x[0] * nrm[0] + x[1] * nrm[1] + x[2] * nrm[2]) / length)
    return min(heights) / 2


def check_cutoff(box: torch.Tensor, cutoff: float) -> None:
    """Assert rc <= half the smallest face height (src/torch_admp/nblist.py:620-660)."""
    if not torch.jit.is_scripting():
        # the same heights from the nine values as Python floats (no tensor ops)
        min_half = _min_half_height(box.detach().reshape(9).to(torch.device("cpu"),
                                                               torch.float64).tolist())
        assert cutoff <= min_half, (
            f"Cutoff {cutoff} is larger than half the minimum height {min_half} of the box. "
            "This may lead to unphysical results.")
        return
    b = box.detach().to(torch.device("cpu"), torch.float64).reshape(3, 3)
    normals = torch.stack([torch.cross(b[1], b[2], dim=-1), torch.cross(b[2], b[0], dim=-1),
                           torch.cross(b[0], b[1], dim=-1)])
    unit = normals / torch.norm(normals, dim=1, keepdim=True)
    heights = torch.abs((b * unit).sum(1))
    min_half_height = float(torch.min(heights)) / 2
    assert cutoff <= min_half_height, (
        f"Cutoff {cutoff} is larger than half the minimum height {min_half_height} of the box. "
        "This may lead to unphysical results.")


def build_pairs(positions: torch.Tensor, box: Optional[torch.Tensor], cutoff: float) -> torch.Tensor:
    """Ordered pair list (both orderings) from the HIP cell list.  One stream
    synchronisation returns the pair count."""
    return torch.ops.admp.neighbor_pairs(positions, box, float(cutoff), None)


class TorchNeighborList(torch.nn.Module):
    """HIP-backed drop-in for the reference TorchNeighborList
    (src/torch_admp/nblist.py:268-333, 553-617); scriptable with torch.jit."""

    def __init__(self, cutoff: float) -> None:
        super().__init__()
        self.cutoff = float(cutoff)
        self.pairs = torch.empty(1, dtype=torch.long)
        self.buffer_scales = torch.empty(1, dtype=torch.long)
        self.ds = torch.empty(1, dtype=torch.float64)
        # row form of the last search: row starts per first atom and the index
        # of each pair's reverse; the distance backward then needs no per-pair
        # atomics
        self.row_ptr = torch.empty(0, dtype=torch.long)
        self.reverse = torch.empty(0, dtype=torch.int32)

    def forward(self, positions: torch.Tensor, box: Optional[torch.Tensor] = None) -> torch.Tensor:
        # one host copy of the box serves the cutoff check, the pair search and
        # the distances (each copy of a device box waits for the stream)
        box_host: Optional[torch.Tensor] = None
        if box is not None:
            box_host = box.detach().to(torch.device("cpu"), torch.float64)
            if not torch.jit.is_scripting():
                _lib.remember_host_box(box, box_host)
            check_cutoff(box_host, self.cutoff)
        # the search pass also writes the buffer scales (i < j)
        pairs, row_ptr, reverse, scales = torch.ops.admp.neighbor_list(positions, box,
                                                                       self.cutoff, box_host)
        self.pairs = pairs
        self.row_ptr = row_ptr
        self.reverse = reverse
        self.buffer_scales = scales
        self.ds = self.pairs_ds(positions, pairs, box, box is not None, box_host, row_ptr,
                                reverse)
        return pairs

    def forward_pbc(self, positions: torch.Tensor, box: torch.Tensor) -> torch.Tensor:
        """Periodic pairs with 0.01 < d < rc (src/torch_admp/nblist.py:335-445)."""
        return torch.ops.admp.neighbor_pairs(positions, box, self.cutoff, None)

    def forward_obc(self, positions: torch.Tensor) -> torch.Tensor:
        """Open-boundary pairs with d < rc, i != j (src/torch_admp/nblist.py:472-490)."""
        return torch.ops.admp.neighbor_pairs(positions, None, self.cutoff, None)

    @staticmethod
    def pairs_buffer_scales(pairs: torch.Tensor) -> torch.Tensor:
        """1 where i < j else 0 (src/torch_admp/nblist.py:492-515)."""
        return (pairs[:, 0] < pairs[:, 1]).to(torch.long)

    @staticmethod
    def pairs_ds(positions: torch.Tensor, pairs: torch.Tensor, box: Optional[torch.Tensor] = None,
                 pbc_flag: bool = True, box_host: Optional[torch.Tensor] = None,
                 row_ptr: Optional[torch.Tensor] = None,
                 reverse: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Differentiable pair distances (src/torch_admp/nblist.py:517-551);
        box_host optionally carries the box's values on the host, row_ptr and
        reverse the row form of these pairs from the search."""
        if pbc_flag:
            assert box is not None, "Box should be provided for periodic system."
            return torch.ops.admp.pair_distance(positions, box, pairs, box_host, row_ptr, reverse)
        return torch.ops.admp.pair_distance(positions, None, pairs, None, row_ptr, reverse)

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


class VerletNeighborList(TorchNeighborList):
    """Neighbour list with a Verlet skin, for MD steps and repeated requests:
    the pair search runs at cutoff + skin and its list is kept while the
    largest displacement since that search stays below skin / 2.

    Same call surface as TorchNeighborList (src/torch_admp/nblist.py:268-617),
    which rebuilds on every call (nblist.py:304-333).  Rows of the kept list
    outside 0.01 < d < cutoff at the current positions get buffer scale 0, so
    every force module gives the energies and forces of a fresh
    TorchNeighborList(cutoff) (the pair set with scale 1 is the same); the
    rows themselves are a superset (both orderings, grouped by first atom).
    A search also runs when the atom count or the box changes, and every call
    searches at cutoff alone when cutoff + skin exceeds half the smallest box
    height.  ``builds`` and ``reuses`` count the two cases.
    """

    def __init__(self, cutoff: float, skin: float = 1.0) -> None:
        super().__init__(cutoff)
        if skin < 0.0:
            raise ValueError(f"skin must be non-negative, got {skin}")
        self.skin = float(skin)
        self._ref: Optional[torch.Tensor] = None
        self._ref_box: Optional[torch.Tensor] = None
        self._list: Optional[torch.Tensor] = None
        self._list_skin = 0.0
        self.builds = 0
        self.reuses = 0

    def _needs_search(self, positions: torch.Tensor, box: Optional[torch.Tensor],
                      box_host: Optional[torch.Tensor]) -> bool:
        if self._list is None or self._ref is None or self._ref.shape != positions.shape:
            return True
        if (box is None) != (self._ref_box is None):
            return True
        if box_host is not None and not torch.equal(box_host, self._ref_box):
            return True
        disp = torch.ops.admp.max_displacement(positions, self._ref, box, box_host)
        return float(disp) >= 0.5 * self._list_skin

    def forward(self, positions: torch.Tensor, box: Optional[torch.Tensor] = None) -> torch.Tensor:
        box_host = None if box is None else box.detach().to(torch.device("cpu"), torch.float64)
        if box is not None and box_host is not None:
            _lib.remember_host_box(box, box_host)
            check_cutoff(box_host, self.cutoff)
        if self._needs_search(positions, box, box_host):
            skin = self.skin
            if box_host is not None and skin > 0.0:
                try:
                    check_cutoff(box_host, self.cutoff + skin)
                except AssertionError:
                    skin = 0.0
            lst, row_ptr, reverse, _ = torch.ops.admp.neighbor_list(positions, box,
                                                                    self.cutoff + skin, box_host)
            self._list = lst
            self.row_ptr = row_ptr
            self.reverse = reverse
            self._list_skin = skin
            self._ref = positions.detach().clone()
            self._ref_box = box_host
            self.builds += 1
        else:
            self.reuses += 1
        pairs = self._list
        self.pairs = pairs
        self.ds = self.pairs_ds(positions, pairs, box, box is not None, box_host, self.row_ptr,
                                self.reverse)
        # i < j and inside 0.01 < d < cutoff (d < cutoff without a box), one kernel
        self.buffer_scales = torch.ops.admp.pair_scales(pairs, self.ds.detach(),
                                                        0.01 if box is not None else -1.0,
                                                        self.cutoff)
        return pairs
