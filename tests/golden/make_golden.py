"""Convert the reference's LAMMPS text fixtures into compact .npz golden files.

Reads (text only) from the reference test data directory:
  tests/data/lmp_coul_obc            100 charges, open boundary, coul/cut 4.0
  tests/data/lmp_coul_pbc            same charges, 10 x 10 x 50 A, pppm gewald 0.5, coul/long 4.0
  tests/data/lmp_coul_pbc_slab_corr  same with kspace_modify slab (boundary p p f)
and writes, per case, positions (N,3), charges (N,), box (3,3), the thermo
potential energy (eV) and the dumped forces (N,3, eV/A) in atom-id order.

Run:  python tests/golden/make_golden.py [/root/reference]
"""

import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def read_data(path: Path):
    lines = path.read_text().splitlines()
    n = None
    bounds = {}
    atoms = []
    in_atoms = False
    for ln in lines:
        tok = ln.split()
        if not tok:
            continue
        if len(tok) >= 2 and tok[1] == "atoms":
            n = int(tok[0])
        for ax in ("x", "y", "z"):
            if len(tok) >= 4 and tok[2] == f"{ax}lo" and tok[3] == f"{ax}hi":
                bounds[ax] = (float(tok[0]), float(tok[1]))
        if tok[0] == "Atoms":
            in_atoms = True
            continue
        if in_atoms:
            if not tok[0].lstrip("-").isdigit():
                if atoms:
                    break
                continue
            atoms.append([float(v) for v in tok[:7]])
            if len(atoms) == n:
                break
    a = np.array(sorted(atoms, key=lambda r: r[0]))
    box = np.diag([bounds[ax][1] - bounds[ax][0] for ax in ("x", "y", "z")])
    return a[:, 4:7], a[:, 3], box


def read_dump_forces(path: Path):
    lines = path.read_text().splitlines()
    i = next(k for k, ln in enumerate(lines) if ln.startswith("ITEM: ATOMS"))
    cols = lines[i].split()[2:]
    rows = [ln.split() for ln in lines[i + 1:] if ln.strip()]
    rows.sort(key=lambda r: int(r[0]))
    idx = [cols.index(c) for c in ("fx", "fy", "fz")]
    return np.array([[float(r[j]) for j in idx] for r in rows])


def read_thermo_pe(path: Path):
    vals = [ln.split() for ln in path.read_text().splitlines() if ln.strip() and not ln.startswith("#")]
    return float(vals[-1][1])


def main(ref_root: Path):
    data = ref_root / "tests" / "data"
    for case in ("lmp_coul_obc", "lmp_coul_pbc", "lmp_coul_pbc_slab_corr"):
        d = data / case
        pos, q, box = read_data(d / "system.data")
        f = read_dump_forces(d / "dump.lammpstrj")
        e = read_thermo_pe(d / "thermo.out")
        np.savez_compressed(OUT / f"{case}.npz", positions=pos, charges=q, box=box, energy=e,
                            forces=f)
        print(case, pos.shape, e)


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
