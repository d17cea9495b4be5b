"""Plotfile and checkpoint files (CPU only): the plotfile's Header, box headers and data read back
to the written valid cells with the right extents, min/max and offsets; a checkpoint reads back to
the level buffer (ghost cells included) bit for bit, with its steps, dt and time."""
import os
import re

import numpy as np

from compas_amd import Layout, cases, plotfile


def _case():
    c = cases.random_smooth(2, 24, seed=3, max_grid_size=12)
    lay = Layout(2, c["boxes"], 3, c["params"].nstate)
    buf = lay.pack(c["U"])
    # ghost cells hold something other than zero, so that a checkpoint must carry them
    buf[buf == 0.0] = -1.25
    return c, lay, buf


def test_plotfile_reads_back(tmp_path):
    c, lay, buf = _case()
    names = ["a1rho1", "a2rho2", "rho_u", "rho_v", "rho_E", "a1"]
    path = plotfile.write_plotfile(str(tmp_path / "plt00010"), [(c["geom"], lay, buf, None)], names, time=0.125,
                                   steps=[10])
    hdr = open(os.path.join(path, "Header")).read().split("\n")
    assert hdr[0] == "HyperCLaw-V1.1" and int(hdr[1]) == 6 and hdr[2:8] == names
    assert int(hdr[8]) == 2 and float(hdr[9]) == 0.125 and int(hdr[10]) == 0
    assert [float(x) for x in hdr[11].split()] == [0.0, 0.0] and [float(x) for x in hdr[12].split()] == [1.0, 1.0]
    assert hdr[14] == "((0,0) (23,23) (0,0))"
    assert hdr[-2] == "Level_0/Cell"
    cell_h = open(os.path.join(path, "Level_0", "Cell_H")).read()
    offs = [int(x) for x in re.findall(r"FabOnDisk: Cell_D_00000 (\d+)", cell_h)]
    assert len(offs) == lay.nbox
    raw = open(os.path.join(path, "Level_0", "Cell_D_00000"), "rb").read()
    for b, (lo, hi) in enumerate(lay.boxes):
        head_end = raw.index(b"\n", offs[b])
        head = raw[offs[b]:head_end].decode()
        m = re.search(r"\(\((\d+),(\d+)\) \((\d+),(\d+)\) \(0,0\)\) (\d+)$", head)
        assert head.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))")
        assert [int(x) for x in m.groups()] == [lo[0], lo[1], hi[0], hi[1], 6]
        n = 6 * (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1)
        data = np.frombuffer(raw, dtype="<f8", count=n, offset=head_end + 1)
        want = lay.fab_view(buf, b)[:, 3:-3, 3:-3]
        assert np.array_equal(data, want.ravel())
    mins = cell_h.split("\n")[-2 * lay.nbox - 3:-lay.nbox - 2]
    assert float(mins[1].split(",")[0]) == lay.fab_view(buf, 0)[0, 3:-3, 3:-3].min()


def test_checkpoint_round_trip(tmp_path):
    c, lay, buf = _case()
    path = str(tmp_path / "chk00020")
    plotfile.write_checkpoint(path, [(lay, buf)], steps=[20], dts=[1.5e-3], times=[0.03])
    steps, dts, times, bufs = plotfile.read_checkpoint(path, [lay])
    assert steps == [20] and dts == [1.5e-3] and times == [0.03]
    assert np.array_equal(bufs[0], buf)
    assert open(os.path.join(path, "Header")).readline() == "Checkpoint file for COMPAS\n"
