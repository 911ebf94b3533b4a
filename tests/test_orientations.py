"""Orientations and the scan's on-disk formats (tme/orientations.py,
tme/matching_utils.py:178-252), after tests/test_orientations.py."""

import os
from tempfile import mkstemp

import numpy as np
import pytest

from pytme_amd import Orientations
from pytme_amd.matching_utils import load_pickle, write_pickle
from pytme_amd.orientations import read_star


@pytest.fixture
def orientations():
    rng = np.random.default_rng(0)
    return Orientations(translations=rng.random((100, 3)).astype(np.float32),
                        rotations=rng.random((100, 3)).astype(np.float32),
                        scores=rng.random(100).astype(np.float32), details=rng.random(100).astype(np.float32))


def _tmp(suffix, tmp_path):
    fd, name = mkstemp(suffix=suffix, dir=tmp_path)
    os.close(fd)
    return name


def test_initialization_and_types():
    rng = np.random.default_rng(1)
    t, r, s, d = rng.random((10, 3)), rng.random((10, 3)), rng.random(10), rng.random(10)
    o = Orientations(translations=t.astype(int), rotations=r.astype(int), scores=s.astype(int), details=d.astype(int))
    assert np.issubdtype(o.translations.dtype, np.floating)
    assert np.issubdtype(o.rotations.dtype, np.floating)
    assert np.issubdtype(o.scores.dtype, np.floating)
    with pytest.raises(ValueError):
        Orientations(translations=t, rotations=rng.random(11), scores=s, details=d)
    with pytest.raises(ValueError):
        Orientations(translations=rng.random(10), rotations=r, scores=s, details=d)


@pytest.mark.parametrize("suffix", [".text", ".star", ".tbl"])
def test_round_trip(orientations, suffix, tmp_path):
    name = _tmp(suffix, tmp_path)
    orientations.to_file(name)
    back = Orientations.from_file(name)
    np.testing.assert_allclose(back.translations, orientations.translations)
    np.testing.assert_allclose(back.rotations, orientations.rotations, atol=1e-3)
    if suffix != ".star":
        np.testing.assert_allclose(back.scores, orientations.scores)
    else:
        np.testing.assert_allclose(back.scores.astype(np.float32), orientations.scores)


@pytest.mark.parametrize("fin", [".text", ".star", ".tbl"])
@pytest.mark.parametrize("fout", [".text", ".star", ".tbl"])
def test_format_conversion(orientations, fin, fout, tmp_path):
    a, b = _tmp(fin, tmp_path), _tmp(fout, tmp_path)
    orientations.to_file(a)
    Orientations.from_file(a).to_file(b)
    back = Orientations.from_file(b)
    np.testing.assert_allclose(back.translations, orientations.translations)


def test_unknown_format(orientations, tmp_path):
    with pytest.raises(ValueError):
        orientations.to_file(_tmp(".x", tmp_path), file_format="mrc")
    with pytest.raises(ValueError):
        Orientations.from_file(_tmp(".x", tmp_path), file_format="mrc")


@pytest.mark.parametrize("order", [("x", "y", "z"), ("z", "y", "x"), ("y", "x", "z")])
def test_text_column_order(order, tmp_path):
    name = _tmp(".tsv", tmp_path)
    rows = ((50, 30, 20), (10, 5, 30))
    with open(name, "w") as f:
        f.write("\t".join(order) + "\n")
        for r in rows:
            f.write("\t".join(str(x) for x in r) + "\n")
    o = Orientations.from_file(name)
    perm = [i for _, i in sorted(zip(order, range(3)))]
    np.testing.assert_array_equal(np.asarray(rows, np.float32)[:, perm], o.translations)
    assert o.rotations.shape == (2, 3) and np.all(o.scores == 0) and np.all(o.details == -1)


def test_star_layout_and_relion_file(tmp_path):
    """The written STAR has RELION's particle columns; a RELION-style file
    with comments, a version line and extra columns reads back."""
    o = Orientations(translations=[[1, 2, 3], [4, 5, 6]], rotations=[[10, 20, 30], [40, 50, 60]],
                     scores=[0.5, 0.25], details=[1, 2])
    name = _tmp(".star", tmp_path)
    o.to_file(name, source_path="tomo.mrc", version="# version 30001")
    text = open(name).read().split("\n")
    assert text[0] == "# version 30001"
    blocks = read_star(name)
    cols = blocks["data_particles"]
    assert list(cols) == ["_rlnCoordinateX", "_rlnCoordinateY", "_rlnCoordinateZ", "_rlnAngleRot", "_rlnAngleTilt",
                          "_rlnAnglePsi", "_rlnClassNumber", "_rlnMicrographName", "_pytmeScore"]
    assert cols["_rlnMicrographName"] == ["tomo.mrc", "tomo.mrc"]
    assert cols["_rlnClassNumber"] == ["1", "2"]
    relion = _tmp(".star", tmp_path)
    with open(relion, "w") as f:
        f.write("# RELION\n\ndata_particles\n\nloop_\n_rlnCoordinateX #1\n_rlnCoordinateY #2\n"
                "_rlnCoordinateZ #3\n_rlnMagnification #4\n7 8 9 10000\n1.5 2.5 3.5 10000\n")
    back = Orientations.from_file(relion)
    np.testing.assert_array_equal(back.translations, [[7, 8, 9], [1.5, 2.5, 3.5]])
    np.testing.assert_array_equal(back.rotations, np.zeros((2, 3)))


@pytest.mark.parametrize("drop", [True, False])
@pytest.mark.parametrize("shape", [10, 11, 40, 41, 80, 81])
def test_extraction_slices(shape, drop):
    rng = np.random.default_rng(shape)
    data = rng.random((50, 50, 50))
    t = np.array([[25, 25, 25], [15, 25, 35], [35, 25, 15], [0, 15, 49]])
    o = Orientations(translations=t, rotations=rng.random(t.shape), scores=rng.random(4), details=rng.random(4))
    ext = np.repeat(shape, 3)
    sub, cand, obs = o.get_extraction_slices(data.shape, ext, drop_out_of_box=drop, return_orientations=True)
    assert sub.translations.shape[0] == len(cand) == len(obs)
    assert (cand, obs) == tuple(o.get_extraction_slices(data.shape, ext, drop_out_of_box=drop))[:2] or drop
    out = np.zeros(ext)
    centre = tuple(np.divide(ext, 2).astype(int))
    for i, (c, s) in enumerate(zip(cand, obs)):
        out[c] = data[s]
        assert out[centre] == data[tuple(sub.translations[i].astype(int))]


def test_result_bundle_round_trip(tmp_path):
    """The scan result bundle (scores, offset, rotation indices, mapping,
    metadata) through write_pickle / load_pickle, with a memmap item."""
    rng = np.random.default_rng(2)
    scores = rng.random((20, 21, 22)).astype(np.float32)
    rot = rng.integers(-1, 50, size=scores.shape).astype(np.int32)
    mapping = {i: rng.random((3, 3)).astype(np.float32) for i in range(5)}
    meta = (np.zeros(3), np.zeros(3), 6.8, {"score": "FLC"})
    name = _tmp(".pickle", tmp_path)
    write_pickle((scores, np.array([0, 0, 0]), rot, mapping, meta), name)
    s2, off2, r2, m2, meta2 = load_pickle(name)
    np.testing.assert_array_equal(s2, scores)
    np.testing.assert_array_equal(r2, rot)
    assert m2.keys() == mapping.keys() and all(np.array_equal(m2[k], mapping[k]) for k in mapping)
    assert meta2[2] == 6.8 and meta2[3] == {"score": "FLC"}
    # a memmap item is moved next to the output and reopened
    mm_name = _tmp(".mm", tmp_path)
    mm = np.memmap(mm_name, dtype=np.float32, mode="w+", shape=scores.shape)
    mm[:] = scores
    mm.flush()
    name2 = _tmp(".pickle", tmp_path)
    write_pickle((mm, rot), name2)
    s3, r3 = load_pickle(name2)
    assert isinstance(s3, np.memmap)
    np.testing.assert_array_equal(np.asarray(s3), scores)
    single = _tmp(".pickle", tmp_path)
    write_pickle(scores, single)
    np.testing.assert_array_equal(load_pickle(single), scores)
