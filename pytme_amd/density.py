"""Tomogram ingestion: CCP4/MRC maps with memory-mapped and partial reads.

Re-creates the part of ``tme.density.Density`` the scan's callers use
(tme/density.py:44-545, 784-884): ``from_file`` with ``subset`` and
``use_memmap``, ``to_file``, and the data / origin / sampling_rate /
metadata attributes.  The MRC2014 header is parsed and written here
directly (the reference goes through ``mrcfile``, which this image lacks).
Data are returned in the reference's axis order: the file's (z, y, x)
sections transposed to (x, y, z), with origin and sampling rate in the same
order.  A subset is read row by row from the file, so a split of a tomogram
larger than host memory never loads the rest.
"""

import gzip as gzip_module
import warnings
from io import BytesIO
from typing import Dict, Tuple

import numpy as np

__all__ = ["Density"]

#: MRC2014 data modes (mode -> dtype), tme/density.py reads them via mrcfile
_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}
_DTYPE_MODE = {np.dtype(v): k for k, v in _MODES.items()}


def _is_gzipped(filename: str) -> bool:
    with open(filename, "rb") as f:
        return f.read(2) == b"\x1f\x8b"


def _read_header(raw: bytes) -> Dict:
    """Fields of the 1024-byte MRC2014 header (little or big endian by MACHST)."""
    if len(raw) < 1024:
        raise ValueError("file shorter than an MRC header")
    machst = raw[212:214]
    endian = ">" if machst[0] == 0x11 else "<"
    i4 = np.frombuffer(raw[:1024], dtype=f"{endian}i4")
    f4 = np.frombuffer(raw[:1024], dtype=f"{endian}f4")
    mode = int(i4[3])
    if mode not in _MODES:
        raise ValueError(f"unsupported MRC mode {mode}")
    if raw[208:212] not in (b"MAP ", b"MAP\x00", b"\x00\x00\x00\x00"):
        raise ValueError("not an MRC file (MAP word missing)")
    nx, ny, nz = (int(x) for x in i4[0:3])
    if min(nx, ny, nz) <= 0:
        raise ValueError("MRC header with non-positive dimensions")
    mx, my, mz = (int(x) for x in i4[7:10])
    cell = f4[10:13].astype(np.float64)
    grid = np.array([mx, my, mz], dtype=np.float64)
    voxel = np.divide(cell, grid, out=np.zeros(3), where=grid > 0)
    return {
        "endian": endian,
        "shape_zyx": (nz, ny, nx),
        "dtype": np.dtype(_MODES[mode]).newbyteorder(endian),
        "start_zyx": np.array([int(i4[6]), int(i4[5]), int(i4[4])]),
        "crs": tuple(int(x) - 1 for x in i4[16:19]),
        "stats": (float(f4[19]), float(f4[20]), float(f4[21]), float(f4[54])),
        "nsymbt": int(i4[23]),
        "origin_xyz": f4[49:52].astype(np.float64),
        "voxel_xyz": voxel,
    }


class Density:
    """A map with its origin and sampling rate (tme/density.py:44-127).

    ``data`` is a 2-D or 3-D array, ``origin`` and ``sampling_rate`` have
    one entry per axis (a scalar sampling rate is broadcast), ``metadata``
    is a dict; anything else raises ``ValueError``.
    """

    def __init__(self, data, origin=None, sampling_rate=None, metadata: Dict = None):
        origin = np.zeros(np.ndim(data)) if origin is None else origin
        sampling_rate = np.ones(np.ndim(data)) if sampling_rate is None else sampling_rate
        metadata = {} if metadata is None else metadata
        origin, sampling_rate = np.asarray(origin), np.asarray(sampling_rate)
        if sampling_rate.size == 1:
            sampling_rate = np.repeat(sampling_rate.reshape(-1), np.ndim(data))
        if origin.size != np.ndim(data) or sampling_rate.size != np.ndim(data):
            raise ValueError("origin and sampling_rate need one entry per data axis")
        if not isinstance(metadata, dict):
            raise ValueError("metadata has to be a dict")
        self.data, self.origin, self.sampling_rate, self.metadata = data, origin, sampling_rate, metadata

    def __repr__(self):
        return (f"Density object at {hex(id(self))}\nOrigin: {tuple(round(float(x), 3) for x in self.origin)}, "
                f"Sampling Rate: {tuple(round(float(x), 3) for x in self.sampling_rate)}, Shape: {self.shape}")

    @property
    def shape(self) -> Tuple[int]:
        return tuple(self.data.shape)

    def copy(self) -> "Density":
        return Density(np.array(self.data, copy=True), self.origin.copy(), self.sampling_rate.copy(),
                       dict(self.metadata))

    # ----------------------------------------------------------------- input
    @classmethod
    def from_file(cls, filename: str, subset: Tuple[slice] = None, use_memmap: bool = False) -> "Density":
        """Read a CCP4/MRC map (gzip allowed; tme/density.py:129-222).
        ``subset`` (slices in the returned axis order) reads only that box;
        ``use_memmap`` maps the file instead of loading it (ignored for
        gzip)."""
        data, origin, sampling_rate, meta = cls._load_mrc(filename, subset=subset, use_memmap=use_memmap)
        return cls(data=data, origin=origin, sampling_rate=sampling_rate, metadata=meta)

    @classmethod
    def _load_mrc(cls, filename: str, subset: Tuple[slice] = None, use_memmap: bool = False):
        """tme/density.py:225-336: origin from the header's ORIGIN, or from the
        start indices times the voxel size when ORIGIN is zero; non-standard
        MAPC/MAPR/MAPS are transposed to the standard order."""
        zipped = _is_gzipped(filename)
        opener = gzip_module.open if zipped else open
        with opener(filename, "rb") as f:
            hdr = _read_header(f.read(1024))
        data_shape = hdr["shape_zyx"]
        crs = hdr["crs"]
        if sorted(crs) != [0, 1, 2]:
            raise ValueError(f"Malformatted CRS array in {filename}")
        origin = hdr["origin_xyz"][::-1].copy()
        sampling_rate = hdr["voxel_xyz"][::-1].copy()
        if np.allclose(origin, 0) and not np.allclose(hdr["start_zyx"], 0):
            origin = np.multiply(hdr["start_zyx"], sampling_rate)
        dmin, dmax, dmean, rms = hdr["stats"]
        meta = {"min": dmin, "max": dmax, "mean": dmean, "std": rms}
        header_size = 1024 + hdr["nsymbt"]
        if zipped and use_memmap:
            warnings.warn(f"Cannot open gzipped file {filename} as memmap.")
            use_memmap = False
        if subset is not None:
            # missing axes are whole axes of the returned map (the reference
            # fills them with the file-order length of the same position,
            # tme/density.py:300-303, which is wrong for non-cubic maps)
            ret_shape = tuple(data_shape[c] for c in crs)[::-1]
            subset = tuple(subset[i] if i < len(subset) else slice(0, ret_shape[i]) for i in range(3))
            subset = tuple(subset[i] for i in crs[::-1])
            if tuple(s.stop - s.start for s in subset) == tuple(data_shape):
                return cls._load_mrc(filename, subset=None, use_memmap=use_memmap)
            data = cls._read_binary_subset(filename, subset, data_shape, hdr["dtype"], header_size)
        elif use_memmap:
            data = np.memmap(filename, dtype=hdr["dtype"], mode="r", offset=header_size, shape=data_shape)
        else:
            with opener(filename, "rb") as f:
                f.seek(header_size)
                nbytes = int(np.prod(data_shape)) * hdr["dtype"].itemsize
                data = np.frombuffer(bytearray(f.read(nbytes)), dtype=hdr["dtype"]).reshape(data_shape)
        if tuple(crs) != (0, 1, 2):
            warnings.warn("Non standard MAPC, MAPR, MAPS, adapting data and origin.")
            data = np.transpose(data, crs)
            origin = np.take(origin, crs)
        return data.T, origin[::-1], sampling_rate[::-1], meta

    @staticmethod
    def _validate_slices(slices: Tuple[slice], shape: Tuple[int]):
        """tme/density.py:439-474."""
        if len(slices) != len(shape):
            raise ValueError(f"Expected length of slices : {len(shape)}, got : {len(slices)}")
        if any(s.stop > n or s.start > n for s, n in zip(slices, shape)):
            raise ValueError(f"Subset exceeds data dimensions ({shape}).")
        if any(s.stop < 0 or s.start < 0 for s in slices):
            raise ValueError("Subsets have to be non-negative.")

    @classmethod
    def _read_binary_subset(cls, filename: str, slices: Tuple[slice], data_shape: Tuple[int], dtype,
                            header_size: int) -> np.ndarray:
        """The (z, y, x) box ``slices`` read row by row at its file offsets
        (tme/density.py:476-545); gzip files are decompressed first."""
        if len(data_shape) != 3:
            raise NotImplementedError("Only 3-dimensional data can be subsetted.")
        cls._validate_slices(slices, data_shape)
        item = dtype.itemsize
        out = np.empty([s.stop - s.start for s in slices], dtype=dtype)
        row_bytes = (slices[2].stop - slices[2].start) * item
        full_row = data_shape[2] * item
        x_off = slices[2].start * item
        zipped = _is_gzipped(filename)
        with (gzip_module.open if zipped else open)(filename, "rb") as f:
            src = BytesIO(f.read()) if zipped else f
            for z in range(slices[0].start, slices[0].stop):
                base = header_size + z * data_shape[1] * full_row
                for y in range(slices[1].start, slices[1].stop):
                    src.seek(base + y * full_row + x_off)
                    out[z - slices[0].start, y - slices[1].start] = np.frombuffer(src.read(row_bytes), dtype=dtype)
        return out

    # ---------------------------------------------------------------- output
    def to_file(self, filename: str, gzip: bool = False) -> None:
        """Write a CCP4/MRC map (tme/density.py:784-883): data in (z, y, x)
        sections, start indices rint(origin / sampling_rate), MAPC/MAPR/MAPS
        1, 2, 3, ORIGIN, cell = shape x sampling rate, min/max/mean/rms.
        Unsupported dtypes are written as float32 (with a warning)."""
        if gzip and not filename.endswith(".gz"):
            filename = f"{filename}.gz"
        data = np.asarray(self.data)
        if data.ndim == 2:
            data = data[..., None]
        data = np.swapaxes(data, 0, 2)
        if data.dtype not in _DTYPE_MODE:
            warnings.warn("Current data type not supported by MRC format. Defaulting to float32.")
            data = data.astype(np.float32)
        data = np.ascontiguousarray(data.astype(data.dtype.newbyteorder("<")))
        nz, ny, nx = data.shape
        sr = np.resize(np.asarray(self.sampling_rate, dtype=np.float64), 3)
        origin = np.resize(np.asarray(self.origin, dtype=np.float64), 3)
        i4 = np.zeros(256, dtype="<i4")
        f4 = i4.view("<f4")
        i4[0:3] = (nx, ny, nz)
        i4[3] = _DTYPE_MODE[np.dtype(data.dtype.newbyteorder("="))]
        i4[4:7] = np.rint(np.divide(origin, sr, out=np.zeros(3), where=sr != 0)).astype(int)
        i4[7:10] = (nx, ny, nz)
        f4[10:13] = np.multiply((nx, ny, nz), sr)
        f4[13:16] = 90.0
        i4[16:19] = (1, 2, 3)
        stats = data.astype(np.float64)
        f4[19:22] = (stats.min(), stats.max(), stats.mean()) if stats.size else (0, 0, 0)
        i4[22] = 1
        i4[27] = 20140
        f4[49:52] = origin
        f4[54] = stats.std() if stats.size else 0
        raw = bytearray(i4.tobytes())
        raw[208:212] = b"MAP "
        raw[212:216] = b"\x44\x44\x00\x00"
        opener = gzip_module.open if gzip else open
        with opener(filename, "wb") as f:
            f.write(bytes(raw))
            f.write(data.tobytes())
