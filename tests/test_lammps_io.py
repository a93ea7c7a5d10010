"""LAMMPS data / dump readers and writers (torch_admp_amd.lammps_io).

* Round trips of write_data / read_data (atom styles full, charge, atomic;
  orthogonal and triclinic boxes) and write_dump / read_dump (several frames).
* A dump written by hand with triclinic BOX BOUNDS and scaled coordinates,
  whose box and Cartesian positions are worked out here from the LAMMPS
  bounding-box convention.
* The golden fixtures under tests/golden were extracted from the reference's
  LAMMPS decks by the independent parsers in tests/golden/make_*.py; when the
  reference tree is present, reading the same decks with lammps_io gives the
  same arrays.
"""

from pathlib import Path

import numpy as np
import pytest

from torch_admp_amd import lammps_io

GOLDEN = Path(__file__).resolve().parent / "golden"
DECKS = Path("/root/reference/tests/data")


@pytest.mark.parametrize("style", ["full", "charge", "atomic"])
@pytest.mark.parametrize("tilt", [False, True])
def test_data_round_trip(tmp_path, style, tilt):
    rng = np.random.default_rng(1)
    box = np.diag([12.5, 13.0, 40.25])
    if tilt:
        box[1, 0], box[2, 0], box[2, 1] = 1.5, -2.25, 0.75
    pos = rng.random((37, 3)) @ box
    q = rng.normal(size=37)
    types = rng.integers(1, 4, 37)
    path = tmp_path / "system.data"
    lammps_io.write_data(path, pos, q, box, types, masses={1: 1.008, 2: 15.999, 3: 196.97},
                         atom_style=style)
    d = lammps_io.read_data(path)
    assert np.array_equal(d["positions"], pos) and np.array_equal(d["box"], box)
    assert np.array_equal(d["types"], types) and list(d["ids"]) == list(range(1, 38))
    assert np.array_equal(d["charges"], q if style != "atomic" else np.zeros(37))
    assert d["masses"] == {1: 1.008, 2: 15.999, 3: 196.97}


def test_dump_round_trip_frames(tmp_path):
    rng = np.random.default_rng(2)
    box = np.array([[10.0, 0.0, 0.0], [2.0, 11.0, 0.0], [-1.0, 0.5, 12.0]])
    path = tmp_path / "dump.lammpstrj"
    frames = []
    for step in range(3):
        pos, f, q = rng.random((20, 3)) @ box, rng.normal(size=(20, 3)), rng.normal(size=20)
        lammps_io.write_dump(path, pos, box, forces=f, charges=q, timestep=10 * step,
                             append=step > 0)
        frames.append((pos, f, q))
    got = lammps_io.read_dump(path, index=None)
    assert [g["timestep"] for g in got] == [0, 10, 20]
    for g, (pos, f, q) in zip(got, frames):
        assert np.allclose(g["box"], box, rtol=0, atol=1e-14)
        assert np.array_equal(g["positions"], pos) and np.array_equal(g["forces"], f)
        assert np.array_equal(g["q"], q)
    assert lammps_io.read_dump(path)["timestep"] == 20


def test_dump_triclinic_bounds_and_scaled_coordinates(tmp_path):
    """Cell a = (10, 0, 0), b = (2, 11, 0), c = (-1, 0.5, 12) at origin
    (1, 2, 3): LAMMPS writes the bounding box xlo_bound = xlo + min(0, xy, xz,
    xy + xz) = 0, xhi_bound = xhi + max(0, xy, xz, xy + xz) = 13, ylo_bound =
    ylo + min(0, yz) = 2, yhi_bound = yhi + max(0, yz) = 13.5."""
    text = "\n".join([
        "ITEM: TIMESTEP", "5", "ITEM: NUMBER OF ATOMS", "2",
        "ITEM: BOX BOUNDS xy xz yz pp pp pp",
        "0.0 13.0 2.0", "2.0 13.5 -1.0", "3.0 15.0 0.5",
        "ITEM: ATOMS id type xs ys zs q",
        "2 1 0.5 0.5 0.5 -1.0", "1 1 0.0 0.0 0.0 1.0"]) + "\n"
    path = tmp_path / "tri.lammpstrj"
    path.write_text(text)
    d = lammps_io.read_dump(path)
    box = np.array([[10.0, 0.0, 0.0], [2.0, 11.0, 0.0], [-1.0, 0.5, 12.0]])
    assert np.allclose(d["box"], box) and np.allclose(d["origin"], [1.0, 2.0, 3.0])
    assert list(d["ids"]) == [1, 2] and list(d["q"]) == [1.0, -1.0]
    assert np.allclose(d["positions"][0], [1.0, 2.0, 3.0])
    assert np.allclose(d["positions"][1], [1.0 + 5.5, 2.0 + 5.75, 3.0 + 6.0])


@pytest.mark.skipif(not DECKS.exists(), reason="the reference tree is not present")
def test_reference_decks_match_golden_fixtures():
    for npz in sorted(GOLDEN.glob("lmp_con*.npz")):
        ref = np.load(npz)
        d = lammps_io.read_dump(DECKS / npz.stem / "dump.lammpstrj")
        assert np.array_equal(d["positions"], ref["positions"]), npz.stem
        assert np.array_equal(d["forces"], ref["forces"]), npz.stem
        assert np.array_equal(d["q"], ref["q"]), npz.stem
        assert np.allclose(np.diag(d["box"]), ref["edges"], rtol=0, atol=1e-12), npz.stem
    for case in ("lmp_coul_obc", "lmp_coul_pbc", "lmp_coul_pbc_slab_corr"):
        ref = np.load(GOLDEN / f"{case}.npz")
        d = lammps_io.read_data(DECKS / case / "system.data")
        assert np.allclose(d["positions"], ref["positions"], rtol=0, atol=1e-12), case
        assert np.array_equal(d["charges"], ref["charges"]), case
        dump = lammps_io.read_dump(DECKS / case / "dump.lammpstrj")
        assert np.allclose(dump["forces"], ref["forces"], rtol=0, atol=1e-12), case
