"""Force-module base class: the drop-in boundary of the reference
(src/torch_admp/base_force.py:18-339).

``forward(positions, box, pairs, ds, buffer_scales, params) -> energy (nf,)``
normalises shapes to a leading frame axis, applies the unit coefficients and
calls ``_forward_impl``.  The ValueError messages are the reference's own,
word for word, because callers and tests match on them.  The class is
TorchScript-compatible: subclasses script and save with torch.jit.
"""

from typing import Dict, Optional, Tuple

import torch

from .utils import TorchConstants


class BaseForceModule(torch.nn.Module):
    """Common interface of the force modules (src/torch_admp/base_force.py:18-122)."""

    def __init__(self, units_dict: Optional[Dict] = None) -> None:
        torch.nn.Module.__init__(self)
        self.const_lib = TorchConstants(units_dict)

    def forward(self, positions: torch.Tensor, box: Optional[torch.Tensor], pairs: torch.Tensor,
                ds: torch.Tensor, buffer_scales: torch.Tensor,
                params: Dict[str, torch.Tensor]) -> torch.Tensor:
        _pos, _box, _pairs, _ds, _bs = self.standardize_input_tensor(
            positions, box, pairs, ds, buffer_scales)
        nf = _pos.size(0)
        _params: Dict[str, torch.Tensor] = {}
        for k, v in params.items():
            _params[k] = v.reshape(nf, -1)
        lc = self.const_lib.length_coeff
        ec = self.const_lib.energy_coeff
        if lc != 1.0:
            _pos = _pos * lc
            _ds = _ds * lc
            if _box is not None:
                _box = _box * lc
        energy = self._forward_impl(_pos, _box, _pairs, _ds, _bs, _params)
        if ec != 1.0:
            energy = energy / ec
        return energy

    def _forward_impl(self, positions: torch.Tensor, box: Optional[torch.Tensor],
                      pairs: torch.Tensor, ds: torch.Tensor, buffer_scales: torch.Tensor,
                      params: Dict[str, torch.Tensor]) -> torch.Tensor:
        raise NotImplementedError("_forward_impl")

    @torch.jit.export
    def standardize_input_tensor(
            self, positions: torch.Tensor, box: Optional[torch.Tensor], pairs: torch.Tensor,
            ds: torch.Tensor, buffer_scales: torch.Tensor
    ) -> Tuple[torch.Tensor, Optional[torch.Tensor], torch.Tensor, torch.Tensor, torch.Tensor]:
        """Shape checks of src/torch_admp/base_force.py:175-339 (same messages)."""
        nframes = 1
        if positions.dim() == 3:
            if positions.size(2) != 3:
                raise ValueError(
                    f"positions must have shape (nframes, natoms, 3), got {positions.shape}")
            nframes = positions.size(0)
        elif positions.dim() == 2:
            if positions.size(1) != 3:
                raise ValueError(f"positions must have shape (natoms, 3), got {positions.shape}")
            positions = positions.unsqueeze(0)
        else:
            raise ValueError(f"positions must be 2D or 3D tensor, got {positions.dim()}D")

        if box is not None:
            if box.dim() == 3:
                if box.size(1) != 3 or box.size(2) != 3:
                    raise ValueError(f"box must have shape (nframes, 3, 3), got {box.shape}")
                if box.size(0) != nframes:
                    raise ValueError(
                        f"box is expected to have {nframes} frame(s), got {box.size(0)}")
            elif box.dim() == 2:
                if box.size(0) != 3 or box.size(1) != 3:
                    raise ValueError(f"box must have shape (3, 3), got {box.shape}")
                if nframes != 1:
                    raise ValueError(
                        f"box must include a frame dimension when positions has {nframes} frames")
                box = box.unsqueeze(0)
            else:
                raise ValueError(f"box must be 2D or 3D tensor, got {box.dim()}D")

        if pairs.dim() == 3:
            if pairs.size(2) != 2:
                raise ValueError(
                    f"pairs must have shape (nframes, n_pairs, 2), got {pairs.shape}")
            if pairs.size(0) != nframes:
                raise ValueError(
                    f"pairs is expected to have {nframes} frame(s), got {pairs.size(0)}")
        elif pairs.dim() == 2:
            if pairs.size(1) != 2:
                raise ValueError(f"pairs must have shape (n_pairs, 2), got {pairs.shape}")
            if nframes != 1:
                raise ValueError(
                    f"pairs must include a frame dimension when positions has {nframes} frames")
            pairs = pairs.unsqueeze(0)
        else:
            raise ValueError(f"pairs must be 2D or 3D tensor, got {pairs.dim()}D")
        n_pairs = pairs.size(-2)

        ds = self._per_pair(ds, "ds", nframes, n_pairs)
        buffer_scales = self._per_pair(buffer_scales, "buffer_scales", nframes, n_pairs)
        return positions, box, pairs, ds, buffer_scales

    def _per_pair(self, t: torch.Tensor, name: str, nframes: int, n_pairs: int) -> torch.Tensor:
        """ds / buffer_scales to (nframes, n_pairs) with the reference's checks."""
        if t.dim() == 2:
            if t.size(0) != nframes:
                raise ValueError(f"{name} is expected to have {nframes} frame(s), got {t.size(0)}")
            if t.size(1) != n_pairs:
                raise ValueError(f"{name} is expected to have {n_pairs} pairs(s), got {t.size(1)}")
        elif t.dim() == 1:
            if t.size(0) != n_pairs:
                raise ValueError(f"{name} is expected to have {n_pairs} pairs(s), got {t.size(0)}")
            if nframes != 1:
                raise ValueError(
                    f"{name} must include a frame dimension when positions has {nframes} frames")
            t = t.unsqueeze(0)
        else:
            raise ValueError(f"{name} must be 1D or 2D tensor, got {t.dim()}D")
        return t
