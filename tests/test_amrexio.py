"""Plotfile and checkpoint files from host data (no device work): the C writer
against the Python reader of mesoflow_amd.amrexio, the checkpoint Header
against the layout of Source/mflo_plot.cpp:101-134, and the read path's
checks. AMReX is not available, so byte compatibility with its own VisMF and
plotfile writers is not checked here (parity unpinned for those layouts); the
fields and their order follow AMReX's documented formats."""
import os

import numpy as np
import pytest

from mesoflow_amd import abi, amrexio

BASE = ["rho", "rhou", "rhov", "rhow", "rhoe", "density", "velx", "vely", "velz", "pressure", "temperature",
        "volfrac"]


def _layout():
    n = (12, 10, 7)
    boxes = [((0, 0, 0), (4, 9, 6)), ((5, 0, 0), (11, 3, 6)), ((5, 4, 0), (11, 9, 2)), ((5, 4, 3), (11, 9, 6))]
    return n, boxes


def _state(nv, n, seed=3):
    r = np.random.default_rng(seed)
    s = r.standard_normal((nv, n[2], n[1], n[0])) * 10.0 ** r.integers(-300, 300, size=(nv, n[2], n[1], n[0]))
    s[0, 0, 0, 0] = -0.0
    s[1, 1, 1, 1] = 5e-324
    return s


def test_plotfile_round_trip(tmp_path):
    n, boxes = _layout()
    nv = 17
    s = _state(nv, n)
    p = str(tmp_path / amrexio.output_name("plt", 40))
    t = 1.2345678901234567e-3
    amrexio.write_plotfile(p, s, boxes, (0.0, -1.0, 0.5), (1.2, 1.0, 1.2), t, 40)
    pf = amrexio.read_plotfile(p)
    assert np.array_equal(pf["data"].view(np.int64), s.view(np.int64))  # bit for bit, -0.0 and denormals included
    assert pf["names"] == BASE + ["S%d" % q for q in range(5)]
    assert pf["time"] == t and pf["level_time"] == t
    assert pf["istep"] == 40 and pf["level_steps"] == 40
    assert pf["domain"] == ((0, 0, 0), (11, 9, 6))
    assert pf["dx"] == [1.2 / 12, 2.0 / 10, 0.7 / 7]
    assert pf["boxes"] == boxes and pf["ref_ratio"] == [] and pf["coord_sys"] == 0
    for q, (lo, hi) in enumerate(boxes):
        for d in range(3):
            assert pf["grids"][q][d] == pytest.approx(((0.0, -1.0, 0.5)[d] + lo[d] * pf["dx"][d],
                                                       (0.0, -1.0, 0.5)[d] + (hi[d] + 1) * pf["dx"][d]), rel=0,
                                                      abs=1e-15)
        blk = s[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].reshape(nv, -1)
        assert np.array_equal(pf["min"][q], blk.min(axis=1))
        assert np.array_equal(pf["max"][q], blk.max(axis=1))
    head = open(os.path.join(p, "Header")).read().split("\n")
    assert head[0] == "HyperCLaw-V1.1" and head[1] == "17" and head[-2] == "Level_0/Cell"
    assert os.path.exists(os.path.join(p, "Level_0", "Cell_D_00000"))


def test_plotfile_names_and_partial_cover(tmp_path):
    n, boxes = _layout()
    names = BASE + ["N2", "O2"]
    s = _state(len(names), n, seed=5)
    p = str(tmp_path / "plt")
    amrexio.write_plotfile(p, s, boxes[:2], (0, 0, 0), (1, 1, 1), 0.0, 0, names)
    pf = amrexio.read_plotfile(p)
    assert pf["names"] == names
    got = pf["data"]
    inside = ~np.isnan(got[0])
    assert inside.sum() == 5 * 10 * 7 + 7 * 4 * 7
    assert np.array_equal(got[:, inside], s[:, inside])


def test_checkpoint_header_layout(tmp_path):
    """mflo_plot.cpp:101-134: title, finest_level, istep, dt, t_new, each list
    entry followed by a space, precision 17, then BoxArray::writeOn"""
    n, boxes = _layout()
    s = _state(13, n)
    p = str(tmp_path / amrexio.output_name("chk", 7))
    t, dt = 0.1 + 0.2, 1.0 / 3.0 * 1e-7
    amrexio.write_checkpoint(p, s, boxes, (0, 0, 0), (1, 1, 1), t, dt, 7)
    want = ("Checkpoint file for mflo\n0\n7 \n%.17g \n%.17g \n(4 0\n" % (dt, t)
            + "".join("((%d,%d,%d) (%d,%d,%d) (0,0,0))\n" % (lo + hi) for lo, hi in boxes) + ")\n")
    assert open(os.path.join(p, "Header")).read() == want
    assert os.path.exists(os.path.join(p, "Level_0", "rho_H"))
    assert os.path.exists(os.path.join(p, "Level_0", "rho_D_00000"))


@pytest.mark.parametrize("dt", [1.0e100, 2.5e-9])
def test_checkpoint_round_trip(tmp_path, dt):
    n, boxes = _layout()
    nv = 22
    s = _state(nv, n, seed=11)
    p = str(tmp_path / "chk")
    t = 7.000000000000001e-4
    amrexio.write_checkpoint(p, s, boxes, (0, 0, 0), (1, 1, 1), t, dt, 123)
    got, tt, dd, ist = amrexio.read_checkpoint_c(p, boxes, nv, n)
    assert np.array_equal(got.view(np.int64), s.view(np.int64))
    assert (tt, dd, ist) == (t, dt, 123)
    py = amrexio.read_checkpoint(p, nv, n)
    assert np.array_equal(py["data"].view(np.int64), s.view(np.int64))
    assert (py["time"], py["dt"], py["istep"], py["boxes"]) == (t, dt, 123, boxes)


def test_checkpoint_first_variable_names_the_files(tmp_path):
    n, boxes = _layout()
    names = ["RHO"] + BASE[1:] + ["A"]
    s = _state(13, n)
    p = str(tmp_path / "chk")
    amrexio.write_checkpoint(p, s, boxes, (0, 0, 0), (1, 1, 1), 0.5, 0.1, 3, names)
    assert os.path.exists(os.path.join(p, "Level_0", "RHO_H"))
    got, *_ = amrexio.read_checkpoint_c(p, boxes, 13, n, varnames=names)
    assert np.array_equal(got, s)
    with pytest.raises(abi.MFError) as e:
        amrexio.read_checkpoint_c(p, boxes, 13, n)  # looks for rho_H
    assert e.value.status == abi.MF_ERR_BAD_ARG


def test_checkpoint_read_checks(tmp_path):
    n, boxes = _layout()
    s = _state(13, n)
    p = str(tmp_path / "chk")
    amrexio.write_checkpoint(p, s, boxes, (0, 0, 0), (1, 1, 1), 0.5, 0.1, 3)
    for bad in (boxes[:3], [((0, 0, 0), (4, 9, 6))] + boxes[1:3] + [((5, 4, 3), (11, 9, 5))]):
        with pytest.raises(abi.MFError) as e:
            amrexio.read_checkpoint_c(p, bad, 13, n)
        assert e.value.status == abi.MF_ERR_BAD_ARG
    with pytest.raises(abi.MFError) as e:
        amrexio.read_checkpoint_c(p, boxes, 14, n)  # component count differs
    assert e.value.status == abi.MF_ERR_BAD_ARG
    with pytest.raises(abi.MFError) as e:
        amrexio.write_checkpoint(str(tmp_path / "x"), s, [((0, 0, 0), (12, 9, 6))], (0, 0, 0), (1, 1, 1), 0, 0, 0)
    assert e.value.status == abi.MF_ERR_BAD_ARG  # box outside the domain
    h = os.path.join(p, "Header")
    text = open(h).read().split("\n")
    text[1] = "1"
    open(h, "w").write("\n".join(text))
    with pytest.raises(abi.MFError) as e:
        amrexio.read_checkpoint_c(p, boxes, 13, n)
    assert e.value.status == abi.MF_ERR_UNSUPPORTED
