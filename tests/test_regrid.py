"""mf_tags_to_boxes (host code of the library, no device needed): the boxes of
a new fine level cover every buffered tag, are disjoint, aligned to the
blocking factor, inside the fine domain and at most max_grid_size per edge;
clustering separates distant tag groups and an aligned block of tags becomes
its own box. The clustering restates AMReX's Berger-Rigoutsos algorithm, which
is not in the tree: agreement with AMReX's box lists is unpinned."""
import numpy as np
import pytest

from mesoflow_amd import abi, amr


def _check(tags, boxes, e, bf, mg, r):
    nz, ny, nx = tags.shape
    cov = np.zeros((nz * r, ny * r, nx * r), np.int32)
    for lo, hi in boxes:
        for d in range(3):
            assert lo[d] % bf == 0 and (hi[d] + 1) % bf == 0
            assert 0 <= lo[d] <= hi[d] < (nx, ny, nz)[d] * r
            assert hi[d] - lo[d] + 1 <= mg
        cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] += 1
    assert cov.max() <= 1, "boxes overlap"
    buf = np.zeros_like(tags, dtype=bool)
    for k, j, i in np.argwhere(tags):
        buf[max(k - e, 0):k + e + 1, max(j - e, 0):j + e + 1, max(i - e, 0):i + e + 1] = True
    fine = np.repeat(np.repeat(np.repeat(buf, r, 0), r, 1), r, 2)
    assert np.all(cov[fine] == 1), "a buffered tag is not covered"
    return cov


@pytest.mark.parametrize("seed", range(6))
def test_random_blobs(seed):
    rng = np.random.default_rng(seed)
    n = (32, 24, 16)
    tags = np.zeros((n[2], n[1], n[0]), np.uint8)
    for _ in range(rng.integers(1, 5)):
        c = [rng.integers(0, n[d]) for d in range(3)]
        rad = rng.integers(1, 5)
        z, y, x = np.ogrid[:n[2], :n[1], :n[0]]
        tags[(x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= rad * rad] = 1
    tags[rng.random(tags.shape) < 0.002] = 1
    for e, bf, mg, eff in ((1, 8, 32, 0.7), (2, 4, 16, 0.9), (0, 2, 64, 0.5)):
        boxes = amr.tags_to_boxes(tags, e, bf, mg, eff, 2)
        cov = _check(tags, boxes, e, bf, mg, 2)
        assert cov.sum() < cov.size or tags.mean() > 0.2


def test_aligned_block_is_one_box_and_distant_groups_split():
    tags = np.zeros((16, 16, 16), np.uint8)
    tags[4:8, 4:8, 4:8] = 1                       # coarse cells 4..7 = fine 8..15
    assert amr.tags_to_boxes(tags, 0, 4, 32, 0.7, 2) == [((8, 8, 8), (15, 15, 15))]
    tags[12:14, 12:14, 0:2] = 1
    boxes = amr.tags_to_boxes(tags, 0, 4, 32, 0.7, 2)
    _check(tags, boxes, 0, 4, 32, 2)
    vol = sum(np.prod([hi[d] - lo[d] + 1 for d in range(3)]) for lo, hi in boxes)
    assert len(boxes) == 2 and vol == 8 ** 3 + 4 ** 3


def test_empty_tags_and_bad_arguments():
    tags = np.zeros((8, 8, 8), np.uint8)
    assert amr.tags_to_boxes(tags, 1, 4, 16, 0.7, 2) == []
    for args in ((1, 3, 16, 0.7, 2), (1, 4, 6, 0.7, 2), (1, 4, 16, 0.0, 2), (1, 1, 16, 0.7, 2), (-1, 4, 16, 0.7, 2)):
        with pytest.raises(abi.MFError) as e:
            amr.tags_to_boxes(tags, *args)
        assert e.value.status == abi.MF_ERR_BAD_ARG
    with pytest.raises(abi.MFError):
        amr.tags_to_boxes(np.zeros((8, 8, 6), np.uint8), 1, 8, 16, 0.7, 2)   # 6 is not a multiple of 8/2


def test_buffer_wraps_across_periodic_seams():
    """AMReX maps buffered tags across periodic boundaries: a tag at i = 0 in a
    periodic direction also buffers the cells at i = n-1 (and likewise at the
    high side); in a non-periodic direction the buffer is clipped."""
    n = 16
    tags = np.zeros((n, n, n), np.uint8)
    tags[8, 8, 0] = 1                      # at the low x face
    tags[8, n - 1, 8] = 1                  # at the high y face
    e, bf, mg, r = 1, 2, 32, 2
    boxes = amr.tags_to_boxes(tags, e, bf, mg, 0.7, r, periodic=(1, 1, 0))
    cov = _check(tags, boxes, e, bf, mg, r)
    fine_cov = lambda i, j, k: cov[k * r, j * r, i * r] == 1
    assert fine_cov(n - 1, 8, 8) and fine_cov(n - 1, 7, 9)        # x image of the first tag's buffer
    assert fine_cov(8, 0, 8) and fine_cov(9, 0, 7)                # y image of the second tag's buffer
    # without periodicity the images are not buffered
    boxes = amr.tags_to_boxes(tags, e, bf, mg, 0.7, r)
    cov = _check(tags, boxes, e, bf, mg, r)
    assert not fine_cov(n - 1, 8, 8) and not fine_cov(8, 0, 8)
    # z is not periodic: a tag at k = 0 is clipped there, not wrapped
    tags2 = np.zeros((n, n, n), np.uint8)
    tags2[0, 4, 4] = 1
    boxes = amr.tags_to_boxes(tags2, e, bf, mg, 0.7, r, periodic=(1, 1, 0))
    cov = _check(tags2, boxes, e, bf, mg, r)
    assert not fine_cov(4, 4, n - 1)
