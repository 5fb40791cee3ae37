"""Plotfiles and checkpoints of a level or a two-level hierarchy (Source/mflo_plot.cpp).

Writing goes through the C-ABI (mf_write_plotfile, mf_write_checkpoint,
mf_read_checkpoint and their mf_level_* forms); this module binds the host
forms and reads both kinds of files back into numpy arrays, for
post-processing without AMReX or yt.

Layouts:
  plotfile    <dir>/Header (HyperCLaw-V1.1: names, time, domain, steps, cell
              sizes, grid bounds), <dir>/Level_0/Cell_H (VisMF header: BoxArray,
              "FabOnDisk: <file> <offset>" per box, per-box min/max) and
              <dir>/Level_0/Cell_D_00000 (one FAB per box: text header, then the
              box's doubles, component-major, x fastest)
  checkpoint  <dir>/Header ("Checkpoint file for mflo", finest_level, istep,
              dt, t_new, BoxArray) and the state as a VisMF file
              <dir>/Level_0/<first variable name>_H / _D_00000
"""
import ctypes as C
import os
import re

import numpy as np

from . import abi

_FAB = re.compile(r"^FAB \(\(8, \(64 11 52 0 1 12 0 1023\)\),\(8, \(1 2 3 4 5 6 7 8\)\)\)"
                  r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \((\d+),(\d+),(\d+)\)\) (\d+)$")
_BOX = re.compile(r"^\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \((\d+),(\d+),(\d+)\)\)$")


def _lib():
    from . import runtime
    return runtime.lib()


def _check(st, where):
    if st != abi.MF_OK:
        raise abi.MFError(where, st)


def _geom(prob_lo, prob_hi, n_cell):
    g = abi.Geom()
    for d in range(3):
        g.prob_lo[d] = float(prob_lo[d])
        g.prob_hi[d] = float(prob_hi[d])
        g.n_cell[d] = int(n_cell[d])
        g.is_periodic[d] = 0
    return g


def _boxes(boxes):
    arr = (abi.Box * len(boxes))()
    for q, (lo, hi) in enumerate(boxes):
        for d in range(3):
            arr[q].lo[d] = int(lo[d])
            arr[q].hi[d] = int(hi[d])
    return arr


def _names(varnames):
    if varnames is None:
        return None
    return (C.c_char_p * len(varnames))(*[v.encode() for v in varnames])


def write_plotfile(path, state, boxes, prob_lo, prob_hi, time, istep, varnames=None):
    """state: [nv][nz][ny][nx] float64 over the domain; boxes: [(lo, hi)]"""
    s = np.ascontiguousarray(state, dtype=np.float64)
    nv, nz, ny, nx = s.shape
    _check(_lib().mf_write_plotfile(os.fsencode(path), C.byref(_geom(prob_lo, prob_hi, (nx, ny, nz))), _boxes(boxes),
                                    len(boxes), nv, s.ctypes.data_as(C.c_void_p), float(time), int(istep),
                                    _names(varnames)), "mf_write_plotfile")


def write_checkpoint(path, state, boxes, prob_lo, prob_hi, time, dt, istep, varnames=None):
    s = np.ascontiguousarray(state, dtype=np.float64)
    nv, nz, ny, nx = s.shape
    _check(_lib().mf_write_checkpoint(os.fsencode(path), C.byref(_geom(prob_lo, prob_hi, (nx, ny, nz))),
                                      _boxes(boxes), len(boxes), nv, s.ctypes.data_as(C.c_void_p), float(time),
                                      float(dt), int(istep), _names(varnames)), "mf_write_checkpoint")


def read_checkpoint_c(path, boxes, nv, n_cell, prob_lo=(0, 0, 0), prob_hi=(1, 1, 1), varnames=None, fill=np.nan):
    """mf_read_checkpoint: (state, time, dt, istep); cells outside the boxes hold fill"""
    nx, ny, nz = n_cell
    s = np.full((nv, nz, ny, nx), fill)
    t, dt, ist = C.c_double(), C.c_double(), C.c_int()
    _check(_lib().mf_read_checkpoint(os.fsencode(path), C.byref(_geom(prob_lo, prob_hi, n_cell)), _boxes(boxes),
                                     len(boxes), nv, s.ctypes.data_as(C.c_void_p), C.byref(t), C.byref(dt),
                                     C.byref(ist), _names(varnames)), "mf_read_checkpoint")
    return s, t.value, dt.value, ist.value


# --------------------------------------------------------------------------
# readers (Python)
# --------------------------------------------------------------------------
def _parse_box(line):
    m = _BOX.match(line.strip())
    if not m:
        raise ValueError("not a box: %r" % line)
    v = [int(x) for x in m.groups()]
    return tuple(v[0:3]), tuple(v[3:6])


def _read_boxarray(lines, i):
    m = re.match(r"^\((\d+) (\d+)$", lines[i].strip())
    if not m:
        raise ValueError("not a BoxArray: %r" % lines[i])
    n = int(m.group(1))
    boxes = [_parse_box(lines[i + 1 + q]) for q in range(n)]
    if lines[i + 1 + n].strip() != ")":
        raise ValueError("BoxArray not closed")
    return boxes, i + 2 + n


def read_vismf(prefix, shape, fill=np.nan):
    """a VisMF MultiFab (<prefix>_H and its data files) into [nv][nz][ny][nx];
    returns (data, boxes, per-box minima, per-box maxima)"""
    d = os.path.dirname(prefix)
    lines = open(prefix + "_H").read().split("\n")
    vers, how, ncomp, ngrow = (int(lines[q]) for q in range(4))
    if vers != 1 or ngrow != 0:
        raise ValueError("VisMF version %d with %d ghost cells is not read here" % (vers, ngrow))
    boxes, i = _read_boxarray(lines, 4)
    nfab = int(lines[i])
    fods = []
    for q in range(nfab):
        tag, fname, off = lines[i + 1 + q].split()
        if tag != "FabOnDisk:":
            raise ValueError(lines[i + 1 + q])
        fods.append((fname, int(off)))
    i += 1 + nfab
    while lines[i].strip() == "":
        i += 1
    mm = []
    for _ in range(2):
        nb, nc = (int(x) for x in lines[i].split(","))
        rows = [[float(x) for x in lines[i + 1 + q].rstrip(",").split(",")] for q in range(nb)]
        mm.append(np.array(rows).reshape(nb, nc))
        i += 1 + nb
        while i < len(lines) and lines[i].strip() == "":
            i += 1
    nv = shape[0]
    if ncomp != nv:
        raise ValueError("%d components, %d expected" % (ncomp, nv))
    out = np.full(shape, fill)
    for (lo, hi), (fname, off) in zip(boxes, fods):
        with open(os.path.join(d, fname), "rb") as f:
            f.seek(off)
            head = f.readline().decode().rstrip("\n")
            m = _FAB.match(head)
            if not m:
                raise ValueError("FAB header %r" % head)
            v = [int(x) for x in m.groups()]
            if (tuple(v[0:3]), tuple(v[3:6])) != (lo, hi) or v[9] != ncomp:
                raise ValueError("FAB box or component count differs from the header")
            ext = [hi[q] - lo[q] + 1 for q in range(3)]
            cnt = ncomp * ext[0] * ext[1] * ext[2]
            a = np.frombuffer(f.read(8 * cnt), dtype="<f8")
            if a.size != cnt:
                raise ValueError("short FAB")
        out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = a.reshape(ncomp, ext[2], ext[1], ext[0])
    return out, boxes, mm[0], mm[1]


def read_plotfile(path, fill=np.nan):
    """a plotfile of one or more levels: dict of the Header's fields, the data
    of level 0 [nv][nz][ny][nx] (fill outside the boxes) and "levels", one
    dict per level (domain, level steps, cell sizes, grids, boxes, data)"""
    lines = open(os.path.join(path, "Header")).read().split("\n")
    if lines[0] != "HyperCLaw-V1.1":
        raise ValueError("not a HyperCLaw-V1.1 plotfile")
    nv = int(lines[1])
    names = lines[2:2 + nv]
    i = 2 + nv
    dim = int(lines[i])
    time = float(lines[i + 1])
    finest = int(lines[i + 2])
    if dim != 3:
        raise ValueError("only 3-D plotfiles are read here")
    nl = finest + 1
    prob_lo = [float(x) for x in lines[i + 3].split()]
    prob_hi = [float(x) for x in lines[i + 4].split()]
    ref = [int(x) for x in lines[i + 5].split()]
    doms = re.findall(r"\(\(-?\d+,-?\d+,-?\d+\) \(-?\d+,-?\d+,-?\d+\) \(\d+,\d+,\d+\)\)", lines[i + 6])
    doms = [_parse_box(d) for d in doms]
    steps = [int(x) for x in lines[i + 7].split()]
    dxs = [[float(x) for x in lines[i + 8 + l].split()] for l in range(nl)]
    j = i + 8 + nl
    coord = int(lines[j])
    bwidth = int(lines[j + 1])
    j += 2
    levels = []
    for l in range(nl):
        lev, ngrid, ltime = lines[j].split()
        lsteps = int(lines[j + 1])
        grids = []
        for q in range(int(ngrid)):
            grids.append([tuple(float(x) for x in lines[j + 2 + 3 * q + d].split()) for d in range(3)])
        mfpath = lines[j + 2 + 3 * int(ngrid)]
        j += 3 + 3 * int(ngrid)
        dom = doms[l]
        n = [dom[1][q] - dom[0][q] + 1 for q in range(3)]
        data, boxes, mn, mx = read_vismf(os.path.join(path, mfpath), (nv, n[2], n[1], n[0]), fill)
        levels.append({"level": int(lev), "level_time": float(ltime), "level_steps": lsteps, "domain": dom,
                       "istep": steps[l], "dx": dxs[l], "grids": grids, "boxes": boxes, "min": mn, "max": mx,
                       "data": data})
    L0 = levels[0]
    return {"names": names, "time": time, "finest_level": finest, "prob_lo": prob_lo, "prob_hi": prob_hi,
            "ref_ratio": ref, "domain": L0["domain"], "istep": L0["istep"], "dx": L0["dx"], "coord_sys": coord,
            "bwidth": bwidth, "level": 0, "level_time": L0["level_time"], "level_steps": L0["level_steps"],
            "grids": L0["grids"], "boxes": L0["boxes"], "min": L0["min"], "max": L0["max"], "data": L0["data"],
            "levels": levels}


def read_checkpoint_levels(path, nv, n_cells, varnames=None, fill=np.nan):
    """a checkpoint of one or more levels: dict(finest_level, istep, dt, time
    (one entry per possible level), boxes and data per existing level);
    n_cells: the domain of each level"""
    lines = open(os.path.join(path, "Header")).read().split("\n")
    if lines[0] != "Checkpoint file for mflo":
        raise ValueError("not an mflo checkpoint")
    finest = int(lines[1])
    istep = [int(x) for x in lines[2].split()]
    dt = [float(x) for x in lines[3].split()]
    time = [float(x) for x in lines[4].split()]
    i = 5
    v0 = varnames[0] if varnames else "rho"
    boxes, data = [], []
    for l in range(finest + 1):
        b, i = _read_boxarray(lines, i)
        nx, ny, nz = n_cells[l]
        d, fboxes, _, _ = read_vismf(os.path.join(path, "Level_%d" % l, v0), (nv, nz, ny, nx), fill)
        if fboxes != b:
            raise ValueError("the MultiFab's boxes differ from the Header's")
        boxes.append(b)
        data.append(d)
    return {"finest_level": finest, "istep": istep, "dt": dt, "time": time, "boxes": boxes, "data": data}


def read_checkpoint(path, nv, n_cell, varnames=None, fill=np.nan):
    """a one-level checkpoint: dict(istep, dt, time, boxes, data)"""
    lines = open(os.path.join(path, "Header")).read().split("\n")
    if lines[0] != "Checkpoint file for mflo":
        raise ValueError("not an mflo checkpoint")
    if int(lines[1]) != 0:
        raise ValueError("only single-level checkpoints are read here")
    istep = int(lines[2].split()[0])
    dt = float(lines[3].split()[0])
    time = float(lines[4].split()[0])
    boxes, _ = _read_boxarray(lines, 5)
    v0 = varnames[0] if varnames else "rho"
    nx, ny, nz = n_cell
    data, fboxes, _, _ = read_vismf(os.path.join(path, "Level_0", v0), (nv, nz, ny, nx), fill)
    if fboxes != boxes:
        raise ValueError("the MultiFab's boxes differ from the Header's")
    return {"istep": istep, "dt": dt, "time": time, "boxes": boxes, "data": data}


def output_name(root, istep, mindigits=5):
    """amrex::Concatenate: plt00010, chk00010 (Source/mflo_plot.cpp:27, :71)"""
    return "%s%0*d" % (root, mindigits, istep)
