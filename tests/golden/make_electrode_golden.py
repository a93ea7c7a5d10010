"""Convert the reference's LAMMPS `fix electrode/conp|conq` dumps into .npz fixtures.

Reads (text only) tests/data/lmp_con{p,q}_*/dump.lammpstrj of the reference
(one `run 0` frame each: ids, unwrapped positions xu yu zu, forces, charges q)
and writes tests/golden/lmp_electrode_<deck>.npz with, in atom-id order,
positions (N,3), forces (N,3, eV/A), q (N,) (electrolyte charges as set and
electrode charges as LAMMPS solved them) and the box edge lengths (3,) from
the BOX BOUNDS.  The solver settings of every deck (constraints, slab, ffield)
are those of the reference tests (tests/test_electrode_conp.py:108-323,
tests/test_electrode_conq.py:29-223) and are restated in the test file.

Run:  python tests/golden/make_electrode_golden.py [/root/reference]
"""

import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent
DECKS = ["lmp_conp_slab_2d", "lmp_conp_slab_3d", "lmp_conp_interface_2d_pzc",
         "lmp_conp_interface_3d_pzc", "lmp_conp_interface_2d_bias", "lmp_conp_interface_3d_bias",
         "lmp_conq_interface_2d_pzc", "lmp_conq_interface_3d_pzc", "lmp_conq_interface_2d_edl",
         "lmp_conq_interface_3d_edl", "lmp_conq_interface_2d_bias", "lmp_conq_interface_3d_bias"]


def read_dump(path: Path):
    lines = path.read_text().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.startswith("ITEM: BOX BOUNDS"))
    edges = []
    for ln in lines[k + 1:k + 4]:
        lo, hi = (float(v) for v in ln.split()[:2])
        edges.append(hi - lo)
    i = next(j for j, ln in enumerate(lines) if ln.startswith("ITEM: ATOMS"))
    cols = lines[i].split()[2:]
    rows = [ln.split() for ln in lines[i + 1:] if ln.strip()]
    rows.sort(key=lambda r: int(r[0]))

    def take(names):
        idx = [cols.index(c) for c in names]
        return np.array([[float(r[j]) for j in idx] for r in rows])

    return take(["xu", "yu", "zu"]), take(["fx", "fy", "fz"]), take(["q"])[:, 0], np.array(edges)


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    for deck in DECKS:
        pos, f, q, edges = read_dump(ref / "tests" / "data" / deck / "dump.lammpstrj")
        np.savez_compressed(OUT / f"{deck}.npz", positions=pos, forces=f, q=q, edges=edges)
        print(deck, pos.shape[0], edges)


if __name__ == "__main__":
    main()
