"""Plotfiles and checkpoints in the reference's on-disk formats, from level buffers.

- `write_plotfile`: the AMReX plotfile the reference writes with amrex::WriteMultiLevelPlotfile
  (source/COMPAS.cpp:2483-2512): a directory with a text `Header` ("HyperCLaw-V1.1": variable
  names, dimension, time, levels, physical extent, refinement ratios, domains, steps, cell sizes,
  then per level its boxes in physical coordinates) and per level `Level_<l>/Cell_H` (the VisMF
  header: box list, where each box's data starts, per-box min and max of every component) with the
  data in `Level_<l>/Cell_D_00000` (per box a "FAB" header line, then the valid cells as
  little-endian doubles, component by component, x fastest).
- `write_checkpoint` / `read_checkpoint`: the reference's checkpoint (WriteCheckpointFile,
  :2515-2598): `Header` ("Checkpoint file for COMPAS", finest level, steps, dt and time per level,
  each level's box list), and per level the state with its ghost cells written by VisMF::Write to
  amrex::MultiFabFileFullPrefix(lev, chk, "Level_", "") (:2594; read back at :2744), i.e. the
  files `Level_<l>/_H` and `Level_<l>/_D_00000`, in VisMF's NoFabHeader_v1 form (:2590): header
  version 2, the valid boxes with the ghost width on its own line, no per-box FAB headers and no
  min/max, and the real format last. A checkpoint read back gives the level buffer bit for bit,
  which is what a restart needs.

AMReX itself (26.01, which defines these formats) is not vendored with the reference; the layouts
here follow its published format and are checked by reading them back and against the file names
and header fields the reference's writer implies. Whether AMReX reads them is unpinned.
"""
import os

import numpy as np

LITTLE_ENDIAN_REAL = "(8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8))"


def _prob_hi(g, d):
    return g.prob_lo[d] + (g.domain_hi[d] - g.domain_lo[d] + 1) * g.dx[d]


def _box_str(lo, hi, dim):
    f = lambda v: "(" + ",".join(str(int(x)) for x in v[:dim]) + ")"
    return "(%s %s %s)" % (f(lo), f(hi), f([0] * dim))


def _boxarray_str(boxes, dim, ng=0):
    out = ["(%d 0" % len(boxes)]
    for lo, hi in boxes:
        out.append(_box_str([l - ng for l in lo], [h + ng for h in hi], dim))
    out.append(")")
    return "\n".join(out)


def _valid(layout, buf, b):
    """[n, (k,) j, i] valid cells of box b."""
    ng = layout.ng
    v = layout.fab_view(buf, b)
    return v[(slice(None),) + (slice(ng, -ng if ng else None),) * layout.dim]


def write_plotfile(path, levels, varnames, time=0.0, steps=None, ref_ratios=()):
    """levels: [(geom, layout, buf, components)], coarsest first; components picks the buffer
    components to write (default all); varnames names them. Returns the path."""
    dim = levels[0][1].dim
    nlev = len(levels)
    steps = steps or [0] * nlev
    os.makedirs(path, exist_ok=True)
    geom0 = levels[0][0]
    lines = ["HyperCLaw-V1.1", str(len(varnames))] + list(varnames)
    lines += [str(dim), repr(float(time)), str(nlev - 1)]
    lines.append(" ".join(repr(float(geom0.prob_lo[d])) for d in range(dim)))
    lines.append(" ".join(repr(float(_prob_hi(geom0, d))) for d in range(dim)))
    lines.append(" ".join(str(int(r)) for r in ref_ratios))
    lines.append(" ".join(_box_str(g.domain_lo, g.domain_hi, dim) for g, _, _, _ in levels))
    lines.append(" ".join(str(int(s)) for s in steps))
    for g, _, _, _ in levels:
        lines.append(" ".join(repr(float(g.dx[d])) for d in range(dim)))
    lines += ["0", "0"]                                  # Cartesian, no boundary width
    for lev, (g, lay, buf, comps) in enumerate(levels):
        lines.append("%d %d %s" % (lev, lay.nbox, repr(float(time))))
        lines.append(str(int(steps[lev])))
        for lo, hi in lay.boxes:
            for d in range(dim):
                a = g.prob_lo[d] + lo[d] * g.dx[d]
                b = g.prob_lo[d] + (hi[d] + 1) * g.dx[d]
                lines.append("%s %s" % (repr(float(a)), repr(float(b))))
        lines.append("Level_%d/Cell" % lev)
        _write_multifab(os.path.join(path, "Level_%d" % lev), "Cell", lay, buf, comps, ng=0, fab_headers=True)
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def _write_multifab(dirname, prefix, lay, buf, comps, ng, fab_headers):
    """VisMF::Write of a level: fab_headers -> Version_v1 (per-box FAB header lines, min/max),
    else NoFabHeader_v1 (version 2: raw data, the real format after the FabOnDisk lines). The box
    list is the valid boxes; ng is the ghost width written with each fab."""
    os.makedirs(dirname, exist_ok=True)
    dim = lay.dim
    comps = list(range(lay.ncomp)) if comps is None else list(comps)
    offsets, mins, maxs = [], [], []
    data_name = prefix + "_D_00000"
    with open(os.path.join(dirname, data_name), "wb") as f:
        for b, (lo, hi) in enumerate(lay.boxes):
            v = lay.fab_view(buf, b) if ng else _valid(lay, buf, b)
            v = v[comps]
            offsets.append(f.tell())
            if fab_headers:
                glo = [l - ng for l in lo]
                ghi = [h + ng for h in hi]
                f.write(("FAB (%s)%s %d\n" % (LITTLE_ENDIAN_REAL, _box_str(glo, ghi, dim), len(comps))).encode())
            f.write(np.ascontiguousarray(v, dtype="<f8").tobytes())
            mins.append(v.reshape(len(comps), -1).min(axis=1))
            maxs.append(v.reshape(len(comps), -1).max(axis=1))
    lines = ["1" if fab_headers else "2", "1", str(len(comps)), str(ng), _boxarray_str(lay.boxes, dim),
             str(lay.nbox)]
    lines += ["FabOnDisk: %s %d" % (data_name, o) for o in offsets]
    lines.append("")
    if fab_headers:
        for arr in (mins, maxs):
            lines.append("%d,%d" % (lay.nbox, len(comps)))
            lines += ["".join("%s," % repr(float(x)) for x in row) for row in arr]
    else:
        lines.append("(%s)" % LITTLE_ENDIAN_REAL)
    with open(os.path.join(dirname, prefix + "_H"), "w") as f:
        f.write("\n".join(lines) + "\n")


def write_checkpoint(path, levels, steps, dts, times):
    """levels: [(layout, buf)] coarsest first, buffers with their ghost cells; steps, dts, times per
    level."""
    dim = levels[0][0].dim
    os.makedirs(path, exist_ok=True)
    lines = ["Checkpoint file for COMPAS", str(len(levels) - 1),
             " ".join(str(int(s)) for s in steps) + " ",
             " ".join(repr(float(x)) for x in dts) + " ",
             " ".join(repr(float(x)) for x in times) + " "]
    for lay, _ in levels:
        lines.append(_boxarray_str(lay.boxes, dim))
        lines.append("")
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("\n".join(lines) + "\n")
    for lev, (lay, buf) in enumerate(levels):
        _write_multifab(os.path.join(path, "Level_%d" % lev), "", lay, buf, None, ng=lay.ng, fab_headers=False)


def _parse_boxarray(tokens_text):
    import re
    m = re.match(r"\((\d+) 0", tokens_text.strip())
    n = int(m.group(1))
    boxes = []
    for bm in re.finditer(r"\(\(([-\d,]+)\) \(([-\d,]+)\) \(([-\d,]+)\)\)", tokens_text):
        lo = [int(x) for x in bm.group(1).split(",")]
        hi = [int(x) for x in bm.group(2).split(",")]
        boxes.append((lo, hi))
    assert len(boxes) == n
    return boxes


def read_checkpoint(path, layouts):
    """(steps, dts, times, [level buffers]) of a checkpoint; layouts gives each level's boxes,
    ghost width and components (their box lists must be the checkpoint's)."""
    with open(os.path.join(path, "Header")) as f:
        text = f.read()
    lines = text.split("\n")
    assert lines[0] == "Checkpoint file for COMPAS"
    nlev = int(lines[1]) + 1
    steps = [int(x) for x in lines[2].split()]
    dts = [float(x) for x in lines[3].split()]
    times = [float(x) for x in lines[4].split()]
    bufs = []
    for lev in range(nlev):
        lay = layouts[lev]
        with open(os.path.join(path, "Level_%d" % lev, "_H")) as f:
            hdr = f.read()
        hl = hdr.split("\n")
        if int(hl[0]) != 2 or int(hl[3]) != lay.ng:
            raise ValueError("checkpoint level %d: header version %s, ghost width %s" % (lev, hl[0], hl[3]))
        start = hdr.index("(")
        ba = hdr[start:hdr.index("\n)", start) + 2]      # the box list ends with a line ")"
        boxes = [(lo, hi) for lo, hi in _parse_boxarray(ba)]
        if [(list(a), list(b)) for a, b in boxes] != [(list(a), list(b)) for a, b in lay.boxes]:
            raise ValueError("checkpoint level %d has other boxes than the layout" % lev)
        offs = [int(l.split()[2]) for l in hdr.splitlines() if l.startswith("FabOnDisk:")]
        buf = np.zeros(lay.total)
        raw = np.fromfile(os.path.join(path, "Level_%d" % lev, "_D_00000"), dtype="<f8")
        for b in range(lay.nbox):
            v = lay.fab_view(buf, b)
            start = offs[b] // 8
            v[...] = raw[start:start + v.size].reshape(v.shape)
        bufs.append(buf)
    return steps, dts, times, bufs
