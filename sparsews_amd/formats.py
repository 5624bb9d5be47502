"""Host-side level-format builders (numpy).

These mirror the reference's format constructors:

* ``DCSR`` / ``DCSC`` ("compressed-compressed", CC_to_taco_tensor,
  utils/dataloader.h:504-569): the outer level lists only non-empty rows
  (``l1_crd``), the inner level is the usual pos/crd pair over those rows.
* ``CSF3``: a 3-level compressed sparse fiber tree for a given mode order,
  the layout the reference loads through SPLATT (read_tns_csf,
  utils/dataloader.h:636-724).  The reference ships no CSF builder of its
  own; this one is built from the sorted COO.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .gen import CSR, COO3


@dataclass
class DCSR:
    nrows: int
    ncols: int
    l1_crd: np.ndarray   # int32 ids of the non-empty outer slices
    l2_pos: np.ndarray   # int32, len(l1_crd) + 1
    l2_crd: np.ndarray   # int32
    vals: np.ndarray     # float64


def dcsr_from_csr(a: CSR) -> DCSR:
    """Drop empty rows from the outer level (dense_to_compressed)."""
    lengths = a.row_lengths()
    keep = np.nonzero(lengths > 0)[0].astype(np.int32)
    pos = np.zeros(len(keep) + 1, dtype=np.int64)
    np.cumsum(lengths[keep], out=pos[1:])
    return DCSR(a.nrows, a.ncols, keep, pos.astype(np.int32), a.crd.astype(np.int32), a.vals)


def dcsc_from_csr(a: CSR) -> DCSR:
    """A in DCSC: the outer level runs over the columns of ``a``."""
    return dcsr_from_csr(a.transpose())


@dataclass
class CSF3:
    """3-level CSF; ``order`` lists the tensor modes from root to leaf."""
    dims: tuple          # tensor dimensions in mode order 0,1,2 (not level order)
    order: tuple
    l1_crd: np.ndarray   # root coordinates (compressed root: pos = [0, len])
    l2_pos: np.ndarray
    l2_crd: np.ndarray
    l3_pos: np.ndarray
    l3_crd: np.ndarray
    vals: np.ndarray

    @property
    def nnz(self) -> int:
        return len(self.vals)


def csf3_from_coo(x: COO3, order) -> CSF3:
    a, b, c = (int(o) for o in order)
    i0 = x.idx[:, a].astype(np.int64)
    i1 = x.idx[:, b].astype(np.int64)
    i2 = x.idx[:, c].astype(np.int64)
    perm = np.lexsort((i2, i1, i0))
    i0, i1, i2 = i0[perm], i1[perm], i2[perm]
    vals = x.vals[perm]
    n = len(vals)
    # level-2 fibers: distinct (i0, i1)
    new2 = np.ones(n, dtype=bool)
    if n > 1:
        new2[1:] = (i0[1:] != i0[:-1]) | (i1[1:] != i1[:-1])
    f2 = np.nonzero(new2)[0]
    l3_pos = np.append(f2, n).astype(np.int32)
    l3_crd = i2.astype(np.int32)
    # level-1 roots: distinct i0 over the fibers
    r0 = i0[f2]
    new1 = np.ones(len(f2), dtype=bool)
    if len(f2) > 1:
        new1[1:] = r0[1:] != r0[:-1]
    f1 = np.nonzero(new1)[0]
    l2_pos = np.append(f1, len(f2)).astype(np.int32)
    l2_crd = i1[f2].astype(np.int32)
    l1_crd = r0[f1].astype(np.int32)
    return CSF3(tuple(x.dims), (a, b, c), l1_crd, l2_pos, l2_crd, l3_pos, l3_crd, vals)


def partial_dcsr(m: np.ndarray, keep_rows, col_density: float, seed: int) -> DCSR:
    """DCSR of the rows of m where keep_rows is true, each storing a random
    subset of its columns (probability col_density each, at least one): a
    hypersparse factor with missing rows and partial rows."""
    rng = np.random.default_rng(seed)
    rows = np.nonzero(np.asarray(keep_rows))[0].astype(np.int32)
    pos, crd = [0], []
    for r in rows:
        cols = np.nonzero(rng.random(m.shape[1]) < col_density)[0]
        if len(cols) == 0:
            cols = np.array([int(rng.integers(m.shape[1]))])
        crd.append(cols)
        pos.append(pos[-1] + len(cols))
    crd = np.concatenate(crd).astype(np.int32) if crd else np.zeros(0, np.int32)
    vals = np.concatenate([m[r, c] for r, c in zip(rows, np.split(crd, pos[1:-1]))]) if len(rows) else np.zeros(0)
    return DCSR(m.shape[0], m.shape[1], rows, np.asarray(pos, dtype=np.int32), crd, np.asarray(vals, np.float64))


def dcsr_dense_and_mask(d: DCSR):
    """(dense matrix, zero where not stored; boolean mask of stored entries)."""
    D = np.zeros((d.nrows, d.ncols))
    S = np.zeros((d.nrows, d.ncols), dtype=bool)
    for t, r in enumerate(d.l1_crd):
        c = d.l2_crd[d.l2_pos[t]:d.l2_pos[t + 1]]
        D[r, c] = d.vals[d.l2_pos[t]:d.l2_pos[t + 1]]
        S[r, c] = True
    return D, S


def dense_as_dcsr(m: np.ndarray) -> DCSR:
    """A dense matrix stored as a fully populated DCSR (the reference loads
    factor matrices this way, bench_hash_mttkrp.cpp:18-25)."""
    r, c = m.shape
    return DCSR(r, c, np.arange(r, dtype=np.int32),
                (np.arange(r + 1, dtype=np.int64) * c).astype(np.int32),
                np.tile(np.arange(c, dtype=np.int32), r), np.ascontiguousarray(m, dtype=np.float64).ravel())
