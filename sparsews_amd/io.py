"""Matrix Market and FROSTT ``.tns`` readers and writers that keep values.

The reference's loaders (utils/dataloader.h) read a Matrix Market file into
CSR and a FROSTT tensor into CSF:

- ``read_mtx_csr_keep_value`` (dataloader.h:752-827) takes real, general,
  coordinate files; pattern entries become 1.0.  Entries are sorted by
  (row, column), and duplicate coordinates are kept as separate entries.
- ``read_mtx_csr`` (dataloader.h:53-190) also mirrors symmetric files, but
  replaces every value with ``rand()`` (:165, :177).
- ``read_tns_csf`` (dataloader.h:636-724) builds the CSF through SPLATT,
  which is not available here.

``read_mtx`` keeps every value.  It mirrors the off-diagonal entries of
symmetric, skew-symmetric and Hermitian-real files.  Duplicate coordinates
are summed by default (``duplicates="sum"``, the CSR that SpGEMM inputs
assume); ``duplicates="keep"`` keeps them in file order, as the reference
does.  ``read_tns`` reads a FROSTT file (1-based indices, one nonzero per
line, ``#`` comments) into a sorted 3-order COO.  ``formats.csf3_from_coo``
then builds the CSF in any mode order.

Files are parsed with pandas' C reader with correctly rounded values, so
million-line files load in seconds and every value is the file's.
"""
from __future__ import annotations

import io
import os

import numpy as np

from .gen import COO3, CSR

_FIELDS = ("real", "integer", "pattern", "double")
_SYMMETRIES = ("general", "symmetric", "skew-symmetric", "hermitian")


def _read_table(f, ncols: int, comment: str = "%") -> np.ndarray:
    import pandas as pd
    # round_trip: correctly rounded values (the default fast parser can be
    # off by one unit in the last place)
    df = pd.read_csv(f, sep=r"\s+", header=None, comment=comment, engine="c", dtype=np.float64,
                     usecols=range(ncols), float_precision="round_trip")
    return df.to_numpy(dtype=np.float64)


def _compress(nrows: int, ncols: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray,
              duplicates: str) -> CSR:
    order = np.lexsort((cols, rows))  # stable: duplicates stay in file order
    rows, cols, vals = rows[order], cols[order], vals[order]
    if duplicates == "sum" and len(rows):
        head = np.ones(len(rows), dtype=bool)
        head[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
        seg = np.cumsum(head) - 1
        summed = np.zeros(int(head.sum()))
        np.add.at(summed, seg, vals)  # in file order within a coordinate
        rows, cols, vals = rows[head], cols[head], summed
    elif duplicates != "keep":
        raise ValueError("duplicates must be 'sum' or 'keep'")
    pos = np.zeros(nrows + 1, dtype=np.int64)
    np.add.at(pos, rows + 1, 1)
    pos = np.cumsum(pos)
    pos = pos.astype(np.int32) if pos[-1] < 2**31 - 1 else pos
    return CSR(nrows, ncols, pos, cols.astype(np.int32), vals.astype(np.float64))


def read_mtx(path, duplicates: str = "sum") -> CSR:
    """CSR of a Matrix Market coordinate file, values kept (pattern -> 1.0)."""
    with open(path, "r") as f:
        banner = f.readline().split()
        if len(banner) < 5 or banner[0].lower() != "%%matrixmarket":
            raise ValueError(f"{path}: not a Matrix Market file")
        obj, fmt, field, sym = (b.lower() for b in banner[1:5])
        if obj != "matrix" or fmt != "coordinate":
            raise ValueError(f"{path}: only sparse (coordinate) matrices are read")
        if field not in _FIELDS:
            raise ValueError(f"{path}: field '{field}' is not supported (real, integer, pattern)")
        if sym not in _SYMMETRIES:
            raise ValueError(f"{path}: symmetry '{sym}' is not supported")
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        nrows, ncols, nnz = (int(t) for t in line.split()[:3])
        width = 2 if field == "pattern" else 3
        body = f.read()
    table = _read_table(io.StringIO(body), width) if nnz else np.zeros((0, width))
    if len(table) != nnz:
        raise ValueError(f"{path}: {len(table)} entries, the size line says {nnz}")
    rows = table[:, 0].astype(np.int64) - 1
    cols = table[:, 1].astype(np.int64) - 1
    vals = np.ones(nnz) if field == "pattern" else table[:, 2].copy()
    if nnz and (rows.min() < 0 or cols.min() < 0 or rows.max() >= nrows or cols.max() >= ncols):
        raise ValueError(f"{path}: an index lies outside the {nrows} x {ncols} matrix")
    if sym != "general":
        off = rows != cols
        mirror = -vals[off] if sym == "skew-symmetric" else vals[off]
        rows, cols, vals = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, mirror]))
    return _compress(nrows, ncols, rows, cols, vals, duplicates)


def write_mtx(path, A: CSR, comment: str = "") -> None:
    """A as a general real Matrix Market coordinate file (values in %.17g)."""
    rows = np.repeat(np.arange(A.nrows, dtype=np.int64), np.diff(np.asarray(A.pos, dtype=np.int64)))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        for c in comment.splitlines():
            f.write(f"% {c}\n")
        f.write(f"{A.nrows} {A.ncols} {len(A.crd)}\n")
        data = np.column_stack([rows + 1, np.asarray(A.crd, dtype=np.int64) + 1])
        for (r, c), v in zip(data, A.vals):
            f.write(f"{r} {c} {float(v):.17g}\n")


def read_tns(path, dims=None) -> COO3:
    """A 3-order FROSTT tensor (1-based indices, value last) as a COO3 sorted
    lexicographically; dims default to the largest index of each mode."""
    if os.path.getsize(path) == 0:
        raise ValueError(f"{path}: empty file")
    table = _read_table(path, 4, comment="#")
    if table.ndim != 2 or table.shape[1] != 4:
        raise ValueError(f"{path}: expected 3 indices and a value per line")
    idx = table[:, :3].astype(np.int64) - 1
    vals = table[:, 3].copy()
    if len(idx) and idx.min() < 0:
        raise ValueError(f"{path}: indices are 1-based")
    if dims is None:
        dims = tuple(int(m) + 1 for m in idx.max(axis=0)) if len(idx) else (0, 0, 0)
    dims = tuple(int(d) for d in dims)
    if len(idx) and np.any(idx.max(axis=0) >= np.asarray(dims)):
        raise ValueError(f"{path}: an index lies outside the dimensions {dims}")
    if max(dims) >= 2**31:
        raise ValueError(f"{path}: a dimension exceeds the int32 index range")
    order = np.lexsort((idx[:, 2], idx[:, 1], idx[:, 0]))
    idx, vals = idx[order], vals[order]
    if len(idx) > 1 and np.any(np.all(idx[1:] == idx[:-1], axis=1)):
        raise ValueError(f"{path}: repeated coordinates")
    return COO3(dims, idx.astype(np.int32), vals)


def write_tns(path, X: COO3) -> None:
    """X as a FROSTT file: 1-based indices, values in %.17g."""
    with open(path, "w") as f:
        for (i, j, k), v in zip(np.asarray(X.idx, dtype=np.int64) + 1, X.vals):
            f.write(f"{i} {j} {k} {float(v):.17g}\n")
