"""LAMMPS ``data`` and text ``dump`` files: readers and writers.

The reference's tests read their LAMMPS decks through ``ase.io``
(tests/test_electrode_conp.py:113-300, tests/test_pme_numerical.py:121-349);
ase is not a dependency here.  These functions cover what those decks hold:

* ``read_data`` / ``write_data``: atom styles ``full`` (id mol type q x y z)
  and ``charge`` (id type q x y z), orthogonal or triclinic boxes (the
  ``xy xz yz`` tilt line), Masses;
* ``read_dump`` / ``write_dump``: ``ITEM:`` text dumps with any columns,
  positions from x y z, xu yu zu or scaled xs ys zs, orthogonal or triclinic
  BOX BOUNDS, every frame.

Boxes are returned as (3, 3) row vectors in the LAMMPS convention
a = (lx, 0, 0), b = (xy, ly, 0), c = (xz, yz, lz), with the lower corner as
``origin``; per-atom arrays are ordered by atom id.
"""

from __future__ import annotations

from pathlib import Path
from typing import Dict, List, Optional, Sequence, Union

import numpy as np

PathLike = Union[str, Path]

_DATA_COLUMNS = {"full": ("id", "mol", "type", "q", "x", "y", "z"),
                 "charge": ("id", "type", "q", "x", "y", "z"),
                 "atomic": ("id", "type", "x", "y", "z")}


def _box_rows(lo, hi, tilt):
    xy, xz, yz = tilt
    return np.array([[hi[0] - lo[0], 0.0, 0.0], [xy, hi[1] - lo[1], 0.0],
                     [xz, yz, hi[2] - lo[2]]])


def read_data(path: PathLike, atom_style: Optional[str] = None) -> Dict[str, np.ndarray]:
    """Read a LAMMPS data file.  The atom style comes from the ``Atoms #
    style`` comment, else ``atom_style``, else ``full``.  Returns ids, types,
    charges (zeros for ``atomic``), positions, box, origin, masses (type ->
    mass) and mol (``full`` only)."""
    lines = Path(path).read_text().splitlines()
    n_atoms = None
    lo, hi, tilt = [0.0] * 3, [0.0] * 3, (0.0, 0.0, 0.0)
    sections: Dict[str, List[str]] = {}
    style = atom_style
    k = 1  # the first line is a title
    while k < len(lines):
        raw = lines[k]
        text = raw.split("#")[0].strip()
        k += 1
        if not text:
            continue
        words = text.split()
        if words[-1] == "atoms" and len(words) == 2:
            n_atoms = int(words[0])
        elif words[-2:] in (["xlo", "xhi"], ["ylo", "yhi"], ["zlo", "zhi"]):
            d = "xyz".index(words[-2][0])
            lo[d], hi[d] = float(words[0]), float(words[1])
        elif words[-3:] == ["xy", "xz", "yz"]:
            tilt = tuple(float(w) for w in words[:3])
        elif words[0] in ("Atoms", "Masses", "Velocities", "Bonds", "Angles", "Dihedrals",
                          "Impropers", "Pair", "PairIJ", "Bond", "Angle"):
            name = words[0]
            if name == "Atoms" and "#" in raw and style is None:
                style = raw.split("#", 1)[1].strip().split()[0]
            body = []
            while k < len(lines) and not lines[k].strip():
                k += 1
            while k < len(lines) and lines[k].strip():
                body.append(lines[k].split("#")[0])
                k += 1
            sections[name] = body
    style = style or "full"
    if style not in _DATA_COLUMNS:
        raise ValueError(f"atom style {style!r} is not supported; expected one of "
                         f"{sorted(_DATA_COLUMNS)}")
    cols = _DATA_COLUMNS[style]
    rows = [ln.split() for ln in sections.get("Atoms", [])]
    if n_atoms is not None and len(rows) != n_atoms:
        raise ValueError(f"{path}: the header declares {n_atoms} atoms, the Atoms section "
                         f"holds {len(rows)}")
    table = np.array([[float(v) for v in r[:len(cols)]] for r in rows]).reshape(-1, len(cols))
    order = np.argsort(table[:, cols.index("id")], kind="stable")
    table = table[order]
    out = {"ids": table[:, cols.index("id")].astype(np.int64),
           "types": table[:, cols.index("type")].astype(np.int64),
           "positions": table[:, [cols.index(c) for c in "xyz"]],
           "charges": table[:, cols.index("q")] if "q" in cols else np.zeros(len(table)),
           "box": _box_rows(lo, hi, tilt), "origin": np.array(lo),
           "masses": {int(float(r.split()[0])): float(r.split()[1])
                      for r in sections.get("Masses", [])}}
    if "mol" in cols:
        out["mol"] = table[:, cols.index("mol")].astype(np.int64)
    return out


def write_data(path: PathLike, positions, charges=None, box=None, types=None, masses=None,
               atom_style: str = "full", origin=(0.0, 0.0, 0.0), title: str = "") -> None:
    """Write a LAMMPS data file; box (3, 3) rows in the LAMMPS convention
    (lower-triangular), types default to 1."""
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    q = np.zeros(n) if charges is None else np.asarray(charges, dtype=np.float64).reshape(-1)
    t = np.ones(n, dtype=np.int64) if types is None else np.asarray(types, dtype=np.int64)
    if atom_style not in _DATA_COLUMNS:
        raise ValueError(f"atom style {atom_style!r} is not supported")
    h = np.diag([1.0, 1.0, 1.0]) if box is None else np.asarray(box, dtype=np.float64)
    if h[0, 1] != 0.0 or h[0, 2] != 0.0 or h[1, 2] != 0.0:
        raise ValueError("LAMMPS boxes are lower-triangular (a along x, b in the xy plane)")
    o = np.asarray(origin, dtype=np.float64)
    out = [title or "written by torch_admp_amd", "", f"{n} atoms", f"{int(t.max()) if n else 1} atom types"]
    for d, name in enumerate("xyz"):
        out.append(f"{o[d]:.17g} {o[d] + h[d, d]:.17g} {name}lo {name}hi")
    if h[1, 0] != 0.0 or h[2, 0] != 0.0 or h[2, 1] != 0.0:
        out.append(f"{h[1, 0]:.17g} {h[2, 0]:.17g} {h[2, 1]:.17g} xy xz yz")
    if masses:
        out += ["", "Masses", ""] + [f"{k} {v:.17g}" for k, v in sorted(masses.items())]
    out += ["", f"Atoms # {atom_style}", ""]
    for i in range(n):
        x, y, z = (f"{v:.17g}" for v in pos[i])
        if atom_style == "full":
            out.append(f"{i + 1} 0 {t[i]} {q[i]:.17g} {x} {y} {z}")
        elif atom_style == "charge":
            out.append(f"{i + 1} {t[i]} {q[i]:.17g} {x} {y} {z}")
        else:
            out.append(f"{i + 1} {t[i]} {x} {y} {z}")
    Path(path).write_text("\n".join(out) + "\n")


def _dump_box(lines: List[str], flags: List[str]):
    vals = [[float(v) for v in ln.split()] for ln in lines]
    if "xy" in flags:
        (xlo_b, xhi_b, xy), (ylo_b, yhi_b, xz), (zlo, zhi, yz) = vals
        xlo = xlo_b - min(0.0, xy, xz, xy + xz)
        xhi = xhi_b - max(0.0, xy, xz, xy + xz)
        ylo, yhi = ylo_b - min(0.0, yz), yhi_b - max(0.0, yz)
        return _box_rows((xlo, ylo, zlo), (xhi, yhi, zhi), (xy, xz, yz)), np.array([xlo, ylo, zlo])
    lo = [v[0] for v in vals]
    hi = [v[1] for v in vals]
    return _box_rows(lo, hi, (0.0, 0.0, 0.0)), np.array(lo)


def read_dump(path: PathLike, index: Optional[int] = -1):
    """Read a LAMMPS text dump.  Returns the frame at ``index`` (the last by
    default; None returns every frame) as a dict: timestep, box, origin,
    ids, positions (from x y z, xu yu zu, or scaled xs ys zs), every column
    by its name (forces as fx fy fz, charges as q), atoms ordered by id."""
    lines = Path(path).read_text().splitlines()
    frames = []
    k = 0
    while k < len(lines):
        if not lines[k].startswith("ITEM: TIMESTEP"):
            k += 1
            continue
        frame: Dict[str, np.ndarray] = {"timestep": int(lines[k + 1].split()[0])}
        k += 2
        assert lines[k].startswith("ITEM: NUMBER OF ATOMS"), lines[k]
        n = int(lines[k + 1].split()[0])
        k += 2
        assert lines[k].startswith("ITEM: BOX BOUNDS"), lines[k]
        flags = lines[k].split()[3:]
        frame["box"], frame["origin"] = _dump_box(lines[k + 1:k + 4], flags)
        k += 4
        assert lines[k].startswith("ITEM: ATOMS"), lines[k]
        names = lines[k].split()[2:]
        tokens = np.array([ln.split() for ln in lines[k + 1:k + 1 + n]], dtype=object)
        tokens = tokens.reshape(n, len(names))
        k += 1 + n
        cols = {}
        for c, name in enumerate(names):
            try:
                cols[name] = tokens[:, c].astype(np.float64)
            except ValueError:  # element names and other text columns
                cols[name] = tokens[:, c].astype(str)
        if "id" in cols:
            order = np.argsort(cols["id"], kind="stable")
            cols = {name: v[order] for name, v in cols.items()}
        frame.update(cols)
        if "id" in cols:
            frame["ids"] = cols["id"].astype(np.int64)
        for keys in (("x", "y", "z"), ("xu", "yu", "zu")):
            if all(c in cols for c in keys):
                frame["positions"] = np.stack([cols[c] for c in keys], 1)
                break
        else:
            if all(c in cols for c in ("xs", "ys", "zs")):
                frac = np.stack([cols[c] for c in ("xs", "ys", "zs")], 1)
                frame["positions"] = frac @ frame["box"] + frame["origin"]
        if all(c in cols for c in ("fx", "fy", "fz")):
            frame["forces"] = np.stack([cols[c] for c in ("fx", "fy", "fz")], 1)
        frames.append(frame)
    if not frames:
        raise ValueError(f"{path}: no frames")
    return frames if index is None else frames[index]


def write_dump(path: PathLike, positions, box, forces=None, charges=None, types=None,
               timestep: int = 0, origin=(0.0, 0.0, 0.0), append: bool = False) -> None:
    """Write one frame of a LAMMPS text dump (id type x y z [q] [fx fy fz]);
    box (3, 3) rows, lower-triangular; boundary flags pp."""
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    h = np.asarray(box, dtype=np.float64)
    o = np.asarray(origin, dtype=np.float64)
    t = np.ones(n, dtype=np.int64) if types is None else np.asarray(types, dtype=np.int64)
    lo, hi = o, o + np.diag(h)
    xy, xz, yz = h[1, 0], h[2, 0], h[2, 1]
    out = ["ITEM: TIMESTEP", str(int(timestep)), "ITEM: NUMBER OF ATOMS", str(n)]
    if xy != 0.0 or xz != 0.0 or yz != 0.0:
        out.append("ITEM: BOX BOUNDS xy xz yz pp pp pp")
        out.append(f"{lo[0] + min(0.0, xy, xz, xy + xz):.16e} "
                   f"{hi[0] + max(0.0, xy, xz, xy + xz):.16e} {xy:.16e}")
        out.append(f"{lo[1] + min(0.0, yz):.16e} {hi[1] + max(0.0, yz):.16e} {xz:.16e}")
        out.append(f"{lo[2]:.16e} {hi[2]:.16e} {yz:.16e}")
    else:
        out.append("ITEM: BOX BOUNDS pp pp pp")
        out += [f"{lo[d]:.16e} {hi[d]:.16e}" for d in range(3)]
    names = ["id", "type", "x", "y", "z"]
    cols: List[Sequence] = [np.arange(1, n + 1), t, pos[:, 0], pos[:, 1], pos[:, 2]]
    if charges is not None:
        names.append("q")
        cols.append(np.asarray(charges, dtype=np.float64).reshape(-1))
    if forces is not None:
        f = np.asarray(forces, dtype=np.float64).reshape(-1, 3)
        names += ["fx", "fy", "fz"]
        cols += [f[:, 0], f[:, 1], f[:, 2]]
    out.append("ITEM: ATOMS " + " ".join(names))
    for i in range(n):
        vals = [str(int(cols[0][i])), str(int(cols[1][i]))]
        vals += [f"{float(c[i]):.16e}" for c in cols[2:]]
        out.append(" ".join(vals))
    mode = "a" if append else "w"
    with open(path, mode) as fh:
        fh.write("\n".join(out) + "\n")
