"""CCP4/MRC ingestion (tme/density.py:129-545, 784-884): the reference's
own map fixture (tests/data/Raw/em_map.map, copied to tests/golden), round
trips, memory-mapped and partial reads, and split scans reading their box
from the file."""

import gzip
import os
import struct

import numpy as np
import pytest

from pytme_amd import Density

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "em_map.map")


def test_reference_fixture():
    """19 x 14 x 20 float32 map, 5.28 A voxels, ORIGIN (-52.8, -10.56,
    -52.8); the data's min / max / mean are the header's DMIN / DMAX /
    DMEAN written by the tool that made the file."""
    d = Density.from_file(FIXTURE)
    assert d.shape == (19, 14, 20)
    np.testing.assert_allclose(d.sampling_rate, [5.28, 5.28, 5.28], rtol=1e-6)
    np.testing.assert_allclose(d.origin, [-52.8, -10.56, -52.8], rtol=1e-6)
    assert d.data.dtype == np.float32
    assert np.isclose(d.data.min(), d.metadata["min"]) and np.isclose(d.data.max(), d.metadata["max"])
    assert abs(float(d.data.mean()) - d.metadata["mean"]) < 1e-5
    mm = Density.from_file(FIXTURE, use_memmap=True)
    assert isinstance(mm.data, np.memmap)
    np.testing.assert_array_equal(np.asarray(mm.data), d.data)
    box = (slice(3, 17), slice(2, 9), slice(5, 20))
    np.testing.assert_array_equal(Density.from_file(FIXTURE, subset=box).data, d.data[box])


@pytest.mark.parametrize("zipped", [False, True])
@pytest.mark.parametrize("use_memmap", [False, True])
@pytest.mark.parametrize("subset", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.int16, np.float16])
def test_round_trip(tmp_path, zipped, use_memmap, subset, dtype):
    rng = np.random.default_rng(0)
    data = (rng.random((50, 40, 30)) * 100).astype(dtype)
    base = Density(data, origin=(10.0, -5.0, 2.5), sampling_rate=(2.5, 2.5, 2.5))
    name = str(tmp_path / ("m.mrc.gz" if zipped else "m.mrc"))
    base.to_file(name[:-3] if zipped else name, gzip=zipped)
    box = (slice(0, 22), slice(31, 40), slice(12, 25)) if subset else None
    if zipped:
        with pytest.warns(UserWarning) if use_memmap else _nullcontext():
            back = Density.from_file(name, use_memmap=use_memmap, subset=box)
    else:
        back = Density.from_file(name, use_memmap=use_memmap, subset=box)
    want = data[box] if subset else data
    np.testing.assert_array_equal(np.asarray(back.data), want)
    np.testing.assert_allclose(back.origin, base.origin)
    np.testing.assert_allclose(back.sampling_rate, base.sampling_rate)


class _nullcontext:
    def __enter__(self):
        return None

    def __exit__(self, *a):
        return False


def test_subset_errors_and_fill(tmp_path):
    name = str(tmp_path / "m.mrc")
    data = np.random.default_rng(3).random((20, 21, 22)).astype(np.float32)
    Density(data).to_file(name)
    for bad in [(slice(-1, 10), slice(5, 10), slice(5, 10)), (slice(5, 100), slice(5, 10), slice(5, 10))]:
        with pytest.raises(ValueError):
            Density.from_file(name, subset=bad)
    # a short subset is completed with whole axes, as tme/density.py:300-303
    np.testing.assert_array_equal(Density.from_file(name, subset=(slice(2, 9),)).data, data[2:9])


def test_non_standard_axis_order(tmp_path):
    """MAPC/MAPR/MAPS (3, 2, 1): the file's fastest axis is z; the map comes
    back in the standard order."""
    name = str(tmp_path / "m.mrc")
    rng = np.random.default_rng(1)
    data = rng.random((6, 7, 8)).astype(np.float32)  # returned axes (x, y, z)
    Density(data).to_file(name)
    raw = bytearray(open(name, "rb").read())
    # rewrite as columns along z: stored array (x, y, z) in C order, header nx=8? no:
    # the file's sections are then x, rows y, columns z
    stored = np.ascontiguousarray(data)  # (x, y, z) with z fastest
    struct.pack_into("<3i", raw, 0, 8, 7, 6)  # NX (columns = z), NY, NZ (sections = x)
    struct.pack_into("<3i", raw, 28, 8, 7, 6)
    struct.pack_into("<3i", raw, 64, 3, 2, 1)
    raw[1024:] = stored.tobytes()
    open(name, "wb").write(bytes(raw))
    with pytest.warns(UserWarning):
        back = Density.from_file(name)
    np.testing.assert_array_equal(back.data, data)


def test_header_fields(tmp_path):
    name = str(tmp_path / "m.mrc")
    data = np.arange(4 * 5 * 6, dtype=np.float32).reshape(4, 5, 6)
    Density(data, origin=(4.0, 8.0, 12.0), sampling_rate=2.0).to_file(name)
    raw = open(name, "rb").read()
    i4 = np.frombuffer(raw[:1024], "<i4")
    f4 = np.frombuffer(raw[:1024], "<f4")
    assert tuple(i4[0:4]) == (4, 5, 6, 2)  # NX = data axis 0
    assert tuple(i4[4:7]) == (2, 4, 6)  # start = origin / sampling
    assert tuple(i4[16:19]) == (1, 2, 3)
    assert raw[208:212] == b"MAP "
    np.testing.assert_allclose(f4[10:13], (8.0, 10.0, 12.0))
    np.testing.assert_allclose(f4[19:22], (0, 119, 59.5))


def test_split_scan_reads_boxes_from_the_file(tmp_path):
    """MatchingData over a memory-mapped map: each split's padded box is
    read from the file and equals the in-memory subset."""
    from pytme_amd.matching_data import MatchingData

    rng = np.random.default_rng(2)
    data = rng.random((40, 36, 30)).astype(np.float32)
    name = str(tmp_path / "t.mrc")
    Density(data).to_file(name)
    mm = Density.from_file(name, use_memmap=True)
    md_file = MatchingData(mm, np.ones((8, 8, 8), np.float32))
    md_mem = MatchingData(data, np.ones((8, 8, 8), np.float32))
    for box, pad in [((slice(0, 20), slice(0, 36), slice(0, 30)), (7, 7, 7)),
                     ((slice(20, 40), slice(10, 30), slice(5, 25)), (7, 6, 5))]:
        got = md_file.subset_array(md_file._target, box, np.asarray(pad))
        want = md_mem.subset_array(data, box, np.asarray(pad))
        np.testing.assert_array_equal(got, want)
