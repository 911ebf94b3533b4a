"""Peak orientations and their on-disk formats.

Re-creates ``tme.orientations.Orientations`` (tme/orientations.py:23-611):
translations (n, d), zyz Euler angles in degrees (n, d), scores (n,) and
details (n,), written to and read from pyTME's tab-separated text format,
RELION STAR files and Dynamo tables, and the extraction boxes around each
orientation.  The column layouts and header names are the reference's so
files interchange with pyTME's ``postprocess.py`` and RELION.
"""

import re
from dataclasses import dataclass
from string import ascii_lowercase, ascii_uppercase
from typing import Dict, List, Tuple

import numpy as np

from .matching_utils import compute_extraction_box

__all__ = ["Orientations", "read_star"]

#: axis names of the text format (tme/orientations.py:19)
NAMES = ["x", "y", "z", *ascii_lowercase[:-3], *ascii_uppercase]

_STAR_COLUMNS = ["_rlnCoordinateX", "_rlnCoordinateY", "_rlnCoordinateZ", "_rlnAngleRot", "_rlnAngleTilt",
                 "_rlnAnglePsi", "_rlnClassNumber"]


def read_star(filename: str, delimiter: str = None) -> Dict[str, Dict[str, List[str]]]:
    """Blocks of a STAR file as {block name: {column label: values}}
    (tme/parser.py:562-608): ``data_`` lines open a block, ``_label`` lines
    name its columns, other non-comment lines are rows split on
    ``delimiter`` (whitespace by default).  Values stay strings."""
    try:
        with open(filename, mode="r") as f:
            text = f.read()
    except UnicodeDecodeError:
        with open(filename, mode="r", encoding="utf-16") as f:
            text = f.read()
    comment = re.compile(r"\s*#.*")
    blocks, labels, rows, name = {}, {}, {}, None
    for line in text.split("\n"):
        if not line or line[0] == "#":
            continue
        if line.startswith("data") and not line.startswith("_"):
            name = line
            labels.setdefault(name, [])
            rows.setdefault(name, [])
            continue
        if name is None:
            continue
        if line.startswith("_"):
            labels[name].append(comment.sub("", line))
            continue
        if line.startswith("loop"):
            continue
        fields = line.split(delimiter)
        if fields:
            rows[name].append(fields)
    for name in labels:
        cols = list(zip(*rows[name])) if rows[name] else [() for _ in labels[name]]
        blocks[name] = {label: list(col) for label, col in zip(labels[name], cols)}
    return blocks


@dataclass
class Orientations:
    """Orientations of matched template copies (tme/orientations.py:23-103).

    Attributes are cast on construction: translations, rotations and scores
    to float32; the first dimensions must agree and translations and
    rotations must be two-dimensional (``ValueError`` otherwise).
    """

    translations: np.ndarray
    rotations: np.ndarray
    scores: np.ndarray
    details: np.ndarray

    def __post_init__(self):
        self.translations = np.array(self.translations).astype(np.float32)
        self.rotations = np.array(self.rotations).astype(np.float32)
        self.scores = np.array(self.scores).astype(np.float32)
        self.details = np.array(self.details)
        n = {self.translations.shape[0], self.rotations.shape[0], self.scores.shape[0], self.details.shape[0]}
        if len(n) != 1:
            raise ValueError("The first dimension of all parameters needs to be of equal length.")
        if self.translations.ndim != 2:
            raise ValueError("Expected two dimensional translations parameter.")
        if self.rotations.ndim != 2:
            raise ValueError("Expected two dimensional rotations parameter.")

    def __iter__(self):
        yield from zip(self.translations, self.rotations, self.scores, self.details)

    def __getitem__(self, indices) -> "Orientations":
        indices = np.asarray(indices)
        return self.__class__(translations=self.translations[indices].copy(),
                              rotations=self.rotations[indices].copy(), scores=self.scores[indices].copy(),
                              details=self.details[indices].copy())

    def copy(self) -> "Orientations":
        return self[np.arange(self.scores.size)]

    # ---------------------------------------------------------------- writing
    def to_file(self, filename: str, file_format: str = None, **kwargs) -> None:
        """Write as "text" (default), "star" (``.star``) or "dynamo" (``.tbl``)
        (tme/orientations.py:152-198)."""
        writers = {"text": self._to_text, "star": self._to_star, "dynamo": self._to_dynamo_tbl}
        if file_format is None:
            file_format = "text"
            if filename.lower().endswith(".star"):
                file_format = "star"
            elif filename.lower().endswith(".tbl"):
                file_format = "dynamo"
        writer = writers.get(file_format)
        if writer is None:
            raise ValueError(f"{file_format} not implemented. Supported are {','.join(writers.keys())}.")
        return writer(filename=filename, **kwargs)

    def _to_text(self, filename: str, **kwargs) -> None:
        """Header x, y, z, euler_x, euler_y, euler_z, score, detail; one
        tab-separated row per orientation (:200-230)."""
        d_t, d_r = self.translations.shape[1], self.rotations.shape[1]
        header = [*NAMES[:d_t], *[f"euler_{x}" for x in NAMES[:d_r]], "score", "detail"]
        with open(filename, mode="w", encoding="utf-8") as f:
            f.write("\t".join(header) + "\n")
            for t, r, s, d in self:
                f.write("\t".join(str(x) for x in (*t, *r, s, d)) + "\n")

    def _to_dynamo_tbl(self, filename: str, name_prefix: str = None, sampling_rate: float = 1.0,
                       subtomogram_size: int = 0, **kwargs) -> None:
        """38-column Dynamo table (:232-301): tag, aligned flag, shifts,
        angles (columns 7-9), scores, wedge (-60, 60), coordinates (24-26),
        pixel size."""
        with open(filename, mode="w", encoding="utf-8") as f:
            for i, (t, r, s, d) in enumerate(self):
                row = [i, 1, 0, 0, 0, 0, *r, self.scores[i], self.scores[i], 0, 0, -90, 90, -60, 60,
                       0, 0, 0, 0, 0, 0, *t, 0, 0, 0, 0, 0, 0, 0, 0, sampling_rate, 3, 0, 0]
                f.write(" ".join(str(x) for x in row) + "\n")

    def _to_star(self, filename: str, source_path: str = None, version: str = None, **kwargs) -> None:
        """RELION particle table (:303-353): coordinates, rot/tilt/psi, class
        number (= detail), optional micrograph name, pyTME score."""
        header = ["data_particles", "", "loop_", *_STAR_COLUMNS]
        if source_path is not None:
            header.append("_rlnMicrographName")
        header.append("_pytmeScore")
        with open(filename, mode="w", encoding="utf-8") as f:
            if version is not None:
                f.write(f"{version.strip()}\n\n")
            f.write("\n".join(header) + "\n")
            for t, r, s, d in self:
                row = [*(str(x) for x in t), *(str(x) for x in r), str(d)]
                if source_path is not None:
                    row.append(source_path)
                row.append(str(s))
                f.write("\t".join(row) + "\n")

    # ---------------------------------------------------------------- reading
    @classmethod
    def from_file(cls, filename: str, file_format: str = None, **kwargs) -> "Orientations":
        """Read "text", "star" or "tbl" (by extension by default;
        tme/orientations.py:355-416)."""
        readers = {"text": cls._from_text, "star": cls._from_star, "tbl": cls._from_tbl}
        if file_format is None:
            file_format = "text"
            if filename.lower().endswith(".star"):
                file_format = "star"
            elif filename.lower().endswith(".tbl"):
                file_format = "tbl"
        reader = readers.get(file_format)
        if reader is None:
            raise ValueError(f"{file_format} not implemented. Supported are {','.join(readers.keys())}.")
        translations, rotations, scores, details, *_ = reader(filename=filename, **kwargs)
        return cls(translations=translations, rotations=rotations, scores=scores, details=details)

    @staticmethod
    def _from_text(filename: str, **kwargs):
        """Columns named by axis are translations, ``euler_*`` columns are
        angles, the last two are score and detail; a file of translations
        only gets zero angles, zero scores and details -1; columns are put
        in sorted name order (:418-490).  For a file of translations only
        the reference's angle reordering indexes with an empty tuple and
        leaves (n, 0) angles; here they stay (n, d) zeros."""
        with open(filename, mode="r", encoding="utf-8") as f:
            rows = [x.strip().split("\t") for x in f.read().split("\n")]
        header = rows.pop(0)
        t_cols = [i for i, x in enumerate(header) if x in NAMES]
        r_cols = [i for i, x in enumerate(header) if "euler" in x]
        translation, rotation, score, detail = [], [], [], []
        for row in rows:
            if len(row) <= 1:
                continue
            translation.append(tuple(row[i] for i in t_cols))
            rotation.append(tuple(row[i] for i in r_cols))
            score.append(row[-2])
            detail.append(row[-1])
        translation = np.vstack(translation)
        rotation = np.vstack(rotation)
        score, detail = np.array(score), np.array(detail)
        if translation.shape[1] == len(header):
            rotation = np.zeros(translation.shape, dtype=np.float32)
            score = np.zeros(translation.shape[0], dtype=np.float32)
            detail = np.zeros(translation.shape[0], dtype=np.float32) - 1
        if rotation.size == 0 and translation.shape[0] != 0:
            rotation = np.zeros(translation.shape, dtype=np.float32)
        names = [x for x in header if x in NAMES]
        translation = translation[..., sorted(range(len(names)), key=lambda i: names[i])]
        enames = [x for x in header if "euler" in x and x.replace("euler_", "") in NAMES]
        if enames and rotation.shape[1] == len(enames):
            rotation = rotation[..., sorted(range(len(enames)), key=lambda i: enames[i])]
        return translation, rotation, score, detail

    @classmethod
    def _from_star(cls, filename: str, delimiter: str = None, **kwargs):
        """The ``data_particles`` (or ``data_``) block (:492-523): coordinates,
        rot/tilt/psi (zero when absent), ``_pytmeScore`` (zero when absent),
        details zero.  Rows split on whitespace unless a ``delimiter`` is
        given (the reference splits on tabs, which its own writer uses)."""
        blocks = read_star(filename, delimiter=delimiter)
        ret = blocks.get("data_particles", blocks.get("data_"))
        if ret is None:
            raise ValueError(f"No data_particles section found in {filename}.")
        translation = np.vstack((ret["_rlnCoordinateX"], ret["_rlnCoordinateY"],
                                 ret["_rlnCoordinateZ"])).astype(np.float32).T
        zero = np.zeros(translation.shape[0], dtype=np.float32)
        angles = [ret.get(x, zero) for x in ("_rlnAngleRot", "_rlnAngleTilt", "_rlnAnglePsi")]
        rotation = np.vstack(angles).astype(np.float32).T
        default = np.zeros(translation.shape[0])
        scores = ret.get("_pytmeScore", default)
        return translation, rotation, scores, default

    @staticmethod
    def _from_tbl(filename: str, **kwargs):
        """Dynamo table of 38 columns (:525-546)."""
        with open(filename, mode="r", encoding="utf-8") as f:
            rows = [x.strip().split(" ") for x in f.read().split("\n") if len(x.strip())]
        if len(rows[0]) != 38:
            raise ValueError("Expected tbl file to have 38 columns generated by _to_tbl.")
        translations = np.array([(r[23], r[24], r[25]) for r in rows])
        rotations = np.array([(r[6], r[7], r[8]) for r in rows])
        scores = np.array([r[9] for r in rows])
        details = np.array([-1 for _ in rows])
        return translations, rotations, scores, details

    # ------------------------------------------------------------- extraction
    def get_extraction_slices(self, target_shape: Tuple[int], extraction_shape: Tuple[int],
                              drop_out_of_box: bool = False, return_orientations: bool = False):
        """Slices of the boxes of ``extraction_shape`` around each
        translation: into the output box and into the target
        (tme/orientations.py:548-611)."""
        obs_beg, obs_end, cand_beg, cand_end, keep = compute_extraction_box(
            self.translations.astype(int), extraction_shape=extraction_shape, original_shape=target_shape)
        subset = self
        if drop_out_of_box:
            subset = self[keep]
            cand_beg, cand_end = cand_beg[keep], cand_end[keep]
            obs_beg, obs_end = obs_beg[keep], obs_end[keep]
        cand_beg, cand_end = cand_beg.astype(int), cand_end.astype(int)
        obs_beg, obs_end = obs_beg.astype(int), obs_end.astype(int)
        candidate = [tuple(slice(a, b) for a, b in zip(r0, r1)) for r0, r1 in zip(cand_beg, cand_end)]
        observation = [tuple(slice(a, b) for a, b in zip(r0, r1)) for r0, r1 in zip(obs_beg, obs_end)]
        if return_orientations:
            return subset, candidate, observation
        return candidate, observation
