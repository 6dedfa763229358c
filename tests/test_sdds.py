"""SDDS-lite against a wake file the reference ships with its examples
(tests/golden/wake_Sx_200um_6mm_xband.sdds), and round trips through every
supported data mode."""
import os

import numpy as np
import pytest

from elegant_amd import sdds
from elegant_amd import lattices as Lt

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_reads_the_example_wake_file():
    s = sdds.read(os.path.join(GOLDEN, "wake_Sx_200um_6mm_xband.sdds"))
    assert [c[0] for c in s.columns] == ["z", "W", "t"]
    assert dict((c[0], c[2].get("units")) for c in s.columns)["W"] == "V/C/m"
    pg = s.pages[0]
    assert pg.rows() == 31
    # values as printed in the file
    assert pg.columns["z"][1] == 2.0e-4 and pg.columns["W"][1] == 7.283500000000001e+014
    assert pg.columns["t"][2] == 1.334256380792608e-012
    t, W = sdds.wake_table(os.path.join(GOLDEN, "wake_Sx_200um_6mm_xband.sdds"))
    assert t[0] == 0 and np.all(np.diff(t) > 0)
    # t = z / c for this file
    assert np.allclose(t, pg.columns["z"] / sdds.C_MKS, rtol=1e-12, atol=0)


@pytest.mark.parametrize("mode,colmajor", [("ascii", False), ("binary", False), ("binary", True)])
def test_round_trip(tmp_path, mode, colmajor):
    s = sdds.SDDS()
    s.define_parameter("Step", "long")
    s.define_parameter("Label", "string")
    s.define_parameter("pCentral", "double", units="m$be$nc")
    s.define_column("x", "double", units="m")
    s.define_column("id", "ulong64")
    s.define_column("pass", "long")
    s.define_column("flag", "short")
    rng = np.random.default_rng(3)
    for k in range(3):
        n = 5 + 7 * k
        s.pages.append(sdds.Page({"Step": k + 1, "Label": "page %d" % k, "pCentral": 11741.7 + k},
                                 {"x": rng.normal(size=n), "id": np.arange(n, dtype=np.uint64) + 2 ** 40,
                                  "pass": rng.integers(-1, 1000, n).astype(np.int32),
                                  "flag": rng.integers(-5, 5, n).astype(np.int16)}))
    p = tmp_path / ("f_%s_%d.sdds" % (mode, colmajor))
    sdds.write(str(p), s, mode=mode, column_major=colmajor)
    r = sdds.read(str(p))
    assert [c[0] for c in r.columns] == ["x", "id", "pass", "flag"]
    assert len(r.pages) == 3
    for a, b in zip(s.pages, r.pages):
        assert a.parameters == b.parameters
        for name in a.columns:
            assert np.array_equal(a.columns[name], b.columns[name]), name


def test_phase_space_and_loss_files(tmp_path):
    bl, Po, rows = Lt.storage_ring_workload(n=1000, seed=5)
    s = sdds.phase_space(rows, Po, step=2, charge=1e-9)
    p = tmp_path / "out.sdds"
    sdds.write(str(p), s, mode="binary")
    back = sdds.rows_from_phase_space(sdds.read(str(p)).pages[0], Po)
    assert np.array_equal(back[:, [0, 1, 2, 3, 6]], rows[:, [0, 1, 2, 3, 6]])
    assert np.allclose(back[:, 4], rows[:, 4], rtol=1e-14, atol=1e-20)
    assert np.allclose(back[:, 5], rows[:, 5], rtol=0, atol=1e-15)
    lost = rows[:10].copy()
    lost[:, 4] = np.linspace(1, 10, 10)
    lost[:, 5] = Po * (1 + lost[:, 5])
    lost[:, 7] = np.arange(10)
    q = tmp_path / "lost.sdds"
    sdds.write(str(q), sdds.lost_particles(lost), mode="ascii")
    pg = sdds.read(str(q)).pages[0]
    assert np.array_equal(pg.columns["Pass"], np.arange(10)) and np.array_equal(pg.columns["s"], lost[:, 4])
    assert np.array_equal(pg.columns["p"], lost[:, 5])


def test_impedance_table_from_files(tmp_path):
    """ZREAL/ZIMAG file specifications ("file=fcol+zcol") read into the
    arrays a ZLONGIT table element takes"""
    f = np.arange(17) * 1e9
    zr = np.linspace(0, 4, 17)
    zi = np.concatenate([[0.0], np.linspace(1, 2, 16)])
    p = str(tmp_path / "z.sdds")
    s = sdds.SDDS()
    s.columns = [("f", "double", {"units": "Hz"}), ("ZReal", "double", {"units": "Ohm"}),
                 ("ZImag", "double", {"units": "Ohm"})]
    s.pages = [sdds.Page(parameters={}, columns={"f": f, "ZReal": zr, "ZImag": zi})]
    sdds.write(p, s)
    F, R, I = sdds.impedance_table(p + "=f+ZReal", p + "=f+ZImag")
    assert np.array_equal(F, f) and np.array_equal(R, zr) and np.array_equal(I, zi)
    F, R, I = sdds.impedance_table(real=p + "=f+ZReal")
    assert I is None and np.array_equal(R, zr)
    with pytest.raises(ValueError):
        sdds.impedance_table(real=p)
