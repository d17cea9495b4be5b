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


def test_checkpoint_files_follow_the_reference_writer(tmp_path):
    """WriteCheckpointFile (source/COMPAS.cpp:2586-2595) writes each level with
    VisMF::Write(dof_new[lev], MultiFabFileFullPrefix(lev, chk, "Level_", "")), i.e. the files
    Level_<l>/_H and Level_<l>/_D_00000, in the NoFabHeader_v1 form: header version 2, one file,
    the component count, the ghost width on its own line, the VALID boxes, the fab offsets, and the
    real format; the data file holds every fab (ghost cells included) with no FAB header lines."""
    c, lay, buf = _case()
    path = str(tmp_path / "chk00030")
    plotfile.write_checkpoint(path, [(lay, buf)], steps=[30], dts=[1e-3], times=[0.5])
    assert sorted(os.listdir(os.path.join(path, "Level_0"))) == ["_D_00000", "_H"]
    h = open(os.path.join(path, "Level_0", "_H")).read().split("\n")
    assert h[0] == "2" and h[1] == "1" and int(h[2]) == 6 and int(h[3]) == 3
    assert h[4] == "(%d 0" % lay.nbox
    for b, (lo, hi) in enumerate(lay.boxes):
        assert h[5 + b] == "((%d,%d) (%d,%d) (0,0))" % (lo[0], lo[1], hi[0], hi[1])
    assert h[5 + lay.nbox] == ")"
    assert int(h[6 + lay.nbox]) == lay.nbox
    fod = h[7 + lay.nbox:7 + 2 * lay.nbox]
    sizes = [lay.fab_view(buf, b).size * 8 for b in range(lay.nbox)]
    assert [int(x.split()[2]) for x in fod] == list(np.cumsum([0] + sizes[:-1]))
    assert all(x.startswith("FabOnDisk: _D_00000 ") for x in fod)
    assert h[-2] == "((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))" and h[-1] == ""
    raw = open(os.path.join(path, "Level_0", "_D_00000"), "rb").read()
    assert len(raw) == sum(sizes)
    assert np.array_equal(np.frombuffer(raw, dtype="<f8"), buf)
