"""Float64 CPU restatement of the reference neighbour list (test oracle).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

=====================  =========================================  ==============================
function               reference                                  pinned by (tests/)
=====================  =========================================  ==============================
wrap_positions         src/torch_admp/nblist.py:447-470           closed form (fractional in [0,1))
nblist_pbc             src/torch_admp/nblist.py:335-445           brute_force_pairs (set equality)
nblist_pbc_cells       src/torch_admp/nblist.py:335-445 (cell     nblist_pbc and brute_force_pairs
                       table, step by step; used as the timed     (set equality)
                       CPU baseline)
nblist_obc             src/torch_admp/nblist.py:472-490           brute_force_pairs
buffer_scales          src/torch_admp/nblist.py:492-515           closed form (i < j)
pbc_shift / pairs_ds   src/torch_admp/spatial.py:15-36,           brute-force minimum image
                       src/torch_admp/nblist.py:517-551
check_cutoff           src/torch_admp/nblist.py:620-660           closed form (face heights)
brute_force_pairs      criteria of nblist.py:439-443,467-470      O(N^2 * 27) enumeration; tiny
                                                                  cases by hand
multi_image_pairs      pair-format convention of                  brute force on tiny boxes,
                       src/torch_admp/nblist.py:199-211           Ewald energy of the result
=====================  =========================================  ==============================
"""

from __future__ import annotations

import itertools
import math

import numpy as np
import torch


def check_cutoff(box: torch.Tensor, cutoff: float) -> None:
    """Assert cutoff <= half the smallest face height (src/torch_admp/nblist.py:620-660)."""
    a1, a2, a3 = box[0], box[1], box[2]
    normals = torch.stack([torch.cross(a2, a3, dim=-1), torch.cross(a3, a1, dim=-1),
                           torch.cross(a1, a2, dim=-1)])
    unit = normals / torch.norm(normals, dim=1, keepdim=True)
    heights = torch.abs((box * unit).sum(1))
    half = torch.min(heights) / 2
    assert cutoff <= half, (
        f"Cutoff {cutoff} is larger than half the minimum height {half} of the box. "
        "This may lead to unphysical results.")


def wrap_positions(pos: torch.Tensor, box: torch.Tensor) -> torch.Tensor:
    """s = (r A + 1e-7) mod 1 - 1e-7; return s h (src/torch_admp/nblist.py:447-470)."""
    eps = 1e-7
    s = (pos @ torch.linalg.inv(box) + eps) % 1.0 - eps
    return s @ box


def nblist_pbc(pos: torch.Tensor, box: torch.Tensor, cutoff: float) -> torch.Tensor:
    """Periodic pair list with both orderings (src/torch_admp/nblist.py:335-445).

    Restated as: pad the wrapped atoms with the image shifts n @ box for
    |n_d| <= ceil(rc |b x c| / V) (nblist.py:348-366), keep (i, j) when
    0.01 < |image_j - wrapped_i| < rc (nblist.py:433-445).  The reference
    finds candidates through a Cartesian cell table; the table only prunes
    candidates, so enumerating all padded images keeps the same set.
    """
    pos = pos.detach()
    box = box.detach()
    w = wrap_positions(pos, box)
    vol = torch.dot(box[0], torch.cross(box[1], box[2], dim=-1))
    pads = []
    for a, b in ((1, 2), (2, 0), (0, 1)):
        pads.append(int(math.ceil(cutoff * torch.cross(box[a], box[b], dim=-1).norm().item()
                                  / vol.item())))
    shifts = torch.cartesian_prod(*[torch.arange(-p, p + 1) for p in pads]).to(box.dtype)
    out = []
    for sh in shifts @ box:                       # one image block at a time, O(N^2) each
        d = torch.norm((w + sh)[None, :, :] - w[:, None, :], dim=-1)
        ii, jj = torch.nonzero((d < cutoff) & (d > 0.01), as_tuple=True)
        out.append(torch.stack((ii, jj), 1))
    pairs = torch.cat(out, 0)
    order = torch.argsort(pairs[:, 0] * pos.shape[0] + pairs[:, 1], stable=True)
    return pairs[order]


def nblist_pbc_cells(pos: torch.Tensor, box: torch.Tensor, cutoff: float) -> torch.Tensor:
    """The reference's own cell-table search (src/torch_admp/nblist.py:335-445),
    restated step by step so its cost can be timed on the CPU:
    image padding, Cartesian cells of edge rc, a stable sort into a dense
    (cells, max occupancy) table, a gather of the 27 neighbour cells of every
    atom, and the 0.01 < d < rc mask."""
    pos = pos.detach()
    box = box.detach()
    w = wrap_positions(pos, box)
    vol = torch.sum(box[0] * torch.cross(box[1], box[2], dim=-1))
    pads = [torch.ceil(cutoff * torch.cross(box[a], box[b], dim=-1).norm() / vol).long().item()
            for a, b in ((1, 2), (2, 0), (0, 1))]
    pad_mat = torch.cartesian_prod(*[torch.arange(-p, p + 1) for p in pads]).to(box.dtype)
    n_img = pad_mat.shape[0]
    padded = (w.unsqueeze(1) + pad_mat @ box).view(-1, 3)
    cpos = torch.floor(padded / cutoff).long()
    corner = torch.min(cpos, dim=0)[0]
    cpos -= corner
    shape = torch.max(cpos, dim=0)[0] + 1
    n_cells = int(torch.prod(shape).item())
    stride = torch.stack([shape[1] * shape[2], shape[2], torch.ones_like(shape[0])])
    cind = torch.sum(cpos * stride, dim=1)
    gind = torch.arange(cind.shape[0]) + 1                       # 1-based padded index
    rind = torch.arange(pos.shape[0]).repeat_interleave(n_img)   # owning atom
    disp = torch.cartesian_prod(*[torch.arange(-1, 2)] * 3)
    atom_cn = ((torch.floor(w / cutoff).long() - corner).unsqueeze(1) + disp)
    atom_cnind = torch.sum(atom_cn * stride, dim=-1)
    cind_sorted, order = torch.sort(cind, stable=True)
    _, counts = torch.unique_consecutive(cind_sorted, return_counts=True)
    max_occ = int(counts.max().item())
    table = torch.zeros((n_cells, max_occ, 2), dtype=torch.long)
    slot = torch.nonzero(torch.arange(max_occ).repeat(counts.shape[0], 1) < counts.unsqueeze(-1))[:, 1]
    table[cind_sorted, slot, 0] = gind[order]
    table[cind_sorted, slot, 1] = rind[order]
    nb = table[atom_cnind]
    pi, nn, jj = torch.where(nb[:, :, :, 0])
    pj = nb[pi, nn, jj, 1]
    pj_pad = nb[pi, nn, jj, 0] - 1
    d = torch.norm(padded[pj_pad] - w[pi], dim=1)
    mask = (d < cutoff) & (d > 0.01)
    return torch.stack((pi, pj), 1)[mask]


def nblist_obc(pos: torch.Tensor, cutoff: float) -> torch.Tensor:
    """Open-boundary pair list from a full distance matrix (src/torch_admp/nblist.py:472-490)."""
    d = torch.cdist(pos.detach(), pos.detach())
    mask = d < cutoff
    mask.fill_diagonal_(False)
    return torch.argwhere(mask)


def buffer_scales(pairs: torch.Tensor) -> torch.Tensor:
    """1 where i < j else 0 (src/torch_admp/nblist.py:492-515)."""
    return (pairs[:, 0] < pairs[:, 1]).to(torch.long)


def pbc_shift(dr: torch.Tensor, box: torch.Tensor) -> torch.Tensor:
    """Minimum-image displacement (src/torch_admp/spatial.py:15-36)."""
    s = dr @ torch.linalg.inv(box)
    s = s - torch.floor(s + 0.5)
    return s @ box


def pairs_ds(pos: torch.Tensor, pairs: torch.Tensor, box=None) -> torch.Tensor:
    """|pbc_shift(r_i - r_j)| or |r_j - r_i| (src/torch_admp/nblist.py:517-551)."""
    ri = pos[pairs[:, 0]]
    rj = pos[pairs[:, 1]]
    dr = rj - ri if box is None else pbc_shift(ri - rj, box)
    return torch.norm(dr, dim=1)


def brute_force_pairs(pos: np.ndarray, box: np.ndarray, cutoff: float, rmin: float = 0.01):
    """Independent numpy enumeration of every (i, j) with rmin < |r_j + n h - r_i| < rc
    over all 27 neighbouring images; returns a set of (i, j) and the list of
    ambiguous pairs whose distance is within 1e-9 of a bound."""
    pos = np.asarray(pos, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64)
    frac = pos @ np.linalg.inv(box)
    frac -= np.floor(frac)
    w = frac @ box
    found, ambiguous = set(), []
    for n in itertools.product((-1, 0, 1), repeat=3):
        sh = np.array(n, dtype=np.float64) @ box
        d = np.linalg.norm(w[None, :, :] + sh - w[:, None, :], axis=-1)
        ii, jj = np.nonzero((d < cutoff) & (d > rmin))
        found.update(zip(ii.tolist(), jj.tolist()))
        amb = np.nonzero((np.abs(d - cutoff) < 1e-9) | (np.abs(d - rmin) < 1e-9))
        ambiguous.extend(zip(amb[0].tolist(), amb[1].tolist()))
    return found, ambiguous


def multi_image_pairs(pos: torch.Tensor, box: torch.Tensor, cutoff: float, rmin: float = 0.01):
    """All periodic images within the cutoff, for boxes smaller than 2 rc.

    Follows the pair-format convention of src/torch_admp/nblist.py:199-211:
    every ordered image pair (i, j, n) is listed, buffer scale 1 for i < j,
    0 for i > j and 0.5 for self images i == j.  Returns pairs (P, 2),
    ds (P,) computed differentiably from ``pos`` and ``box``, and scales (P,).
    """
    with torch.no_grad():
        inv = torch.linalg.inv(box)
        s = pos @ inv
        s = s - torch.floor(s)
        w = s @ box
        heights = []
        vol = torch.dot(box[0], torch.cross(box[1], box[2], dim=-1)).abs()
        for a, b in ((1, 2), (2, 0), (0, 1)):
            heights.append((vol / torch.cross(box[a], box[b], dim=-1).norm()).item())
        nmax = [int(math.ceil(cutoff / h)) + 1 for h in heights]
        shifts = torch.cartesian_prod(*[torch.arange(-m, m + 1) for m in nmax]).to(box.dtype)
        rows = []
        for n in shifts:
            d = torch.norm((w + n @ box)[None] - w[:, None], dim=-1)
            ii, jj = torch.nonzero((d < cutoff) & (d > rmin), as_tuple=True)
            rows.append(torch.cat((ii[:, None], jj[:, None],
                                   n.to(torch.long)[None].expand(ii.shape[0], 3)), 1))
        rows = torch.cat(rows, 0)
        wrap_shift = (torch.floor(pos @ inv)).to(box.dtype)  # w = pos - wrap_shift @ box
    i, j, n = rows[:, 0], rows[:, 1], rows[:, 2:].to(box.dtype)
    # image displacement expressed on the unwrapped positions (differentiable)
    cell = n - wrap_shift[j] + wrap_shift[i]
    dr = pos[j] + cell @ box - pos[i]
    ds = torch.norm(dr, dim=1)
    bs = torch.where(i < j, 1.0, torch.where(i == j, 0.5, 0.0)).to(pos.dtype)
    return rows[:, :2].contiguous(), ds, bs
