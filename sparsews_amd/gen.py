"""Seeded synthetic inputs for the sparse-workspace kernels.

Every generator is deterministic in (seed, stream): it draws from
``numpy.random.default_rng([seed, stream])`` (PCG64).  The reference reads
SuiteSparse / FROSTT files instead (utils/dataloader.h:53-186, :636-724);
the synthetic shapes below are the workloads this package is measured on:

* ``uniform_csr``   -- m x n CSR, exactly ``r`` distinct uniform columns per
  row (10,000 x 10,000 with r = 8 is the row-wise hash SpGEMM workload).
* ``zipf_csr``      -- row lengths from a truncated Zipf law, columns uniform
  distinct (200,000 x 200,000, exponent 1.8, cap 20,000 is the outer-product
  workload, A = B^T).
* ``zipf_coo3``     -- 3-order COO tensor with per-mode Zipf (or uniform)
  indices, deduplicated to an exact nonzero count (MTTKRP / TTM workloads).
* ``request_stream`` -- the batched/served SpGEMM request mix.

Values are FP64 uniform in [-1, 1] (exact zeros redrawn) unless stated.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def rng_for(seed: int, stream: int) -> np.random.Generator:
    return np.random.default_rng([int(seed), int(stream)])


def uniform_values(rng: np.random.Generator, n: int, lo: float = -1.0, hi: float = 1.0,
                   nonzero: bool = True) -> np.ndarray:
    v = rng.uniform(lo, hi, n)
    if nonzero:
        bad = v == 0.0
        while bad.any():
            v[bad] = rng.uniform(lo, hi, int(bad.sum()))
            bad = v == 0.0
    return v


@dataclass
class CSR:
    """CSR matrix: int32 row pointers (int64 when nnz >= 2^31), int32 columns, FP64 values."""
    nrows: int
    ncols: int
    pos: np.ndarray
    crd: np.ndarray
    vals: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.pos[-1])

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.pos.astype(np.int64))

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.crd, self.pos), shape=(self.nrows, self.ncols))

    def transpose(self) -> "CSR":
        """CSR of the transpose (equivalently: this matrix in CSC)."""
        order = np.lexsort((np.repeat(np.arange(self.nrows, dtype=np.int64), self.row_lengths()),
                            self.crd))
        rows = np.repeat(np.arange(self.nrows, dtype=np.int32), self.row_lengths())
        cnt = np.bincount(self.crd, minlength=self.ncols).astype(np.int64)
        pos = np.zeros(self.ncols + 1, dtype=np.int64)
        np.cumsum(cnt, out=pos[1:])
        return CSR(self.ncols, self.nrows, pos.astype(self.pos.dtype),
                   rows[order].astype(np.int32), self.vals[order].copy())


def _pos_from_lengths(lengths: np.ndarray) -> np.ndarray:
    pos = np.zeros(len(lengths) + 1, dtype=np.int64)
    np.cumsum(lengths, out=pos[1:])
    if pos[-1] < 2**31 - 1:
        return pos.astype(np.int32)
    return pos


def distinct_sorted_columns(rng: np.random.Generator, lengths: np.ndarray, n: int) -> np.ndarray:
    """For each row r, ``lengths[r]`` distinct columns uniform in [0, n), sorted.

    Draw all columns, sort by (row, column), redraw the duplicates, repeat
    until no row holds a duplicate.
    """
    lengths = np.asarray(lengths, dtype=np.int64)
    if np.any(lengths > n):
        raise ValueError("row length exceeds the number of columns")
    total = int(lengths.sum())
    rows = np.repeat(np.arange(len(lengths), dtype=np.int64), lengths)
    cols = rng.integers(0, n, total, dtype=np.int64)
    while True:
        key = rows * n + cols
        key.sort()
        dup = np.zeros(total, dtype=bool)
        if total > 1:
            dup[1:] = key[1:] == key[:-1]
        cols = key % n
        nd = int(dup.sum())
        if nd == 0:
            return cols.astype(np.int32)
        cols[dup] = rng.integers(0, n, nd, dtype=np.int64)


def uniform_csr(m: int, n: int, r: int, seed: int, stream: int) -> CSR:
    """m x n CSR with exactly r distinct uniform columns per row, FP64 U[-1,1]."""
    rng = rng_for(seed, stream)
    lengths = np.full(m, r, dtype=np.int64)
    cols = distinct_sorted_columns(rng, lengths, n)
    vals = uniform_values(rng, len(cols))
    return CSR(m, n, _pos_from_lengths(lengths), cols, vals)


def zipf_lengths(rng: np.random.Generator, count: int, s: float, lo: int, hi: int) -> np.ndarray:
    """Truncated discrete Zipf on [lo, hi]: P(L = l) proportional to l^-s (inverse CDF)."""
    support = np.arange(lo, hi + 1, dtype=np.float64)
    pmf = support ** (-s)
    cdf = np.cumsum(pmf)
    cdf /= cdf[-1]
    u = rng.random(count)
    return (np.searchsorted(cdf, u, side="right") + lo).clip(lo, hi).astype(np.int64)


def zipf_csr(n: int, s: float, cap: int, seed: int, stream: int, lo: int = 1) -> CSR:
    """n x n CSR with Zipf(s) row lengths on [lo, cap], uniform distinct columns."""
    rng = rng_for(seed, stream)
    lengths = zipf_lengths(rng, n, s, lo, min(cap, n))
    cols = distinct_sorted_columns(rng, lengths, n)
    vals = uniform_values(rng, len(cols))
    return CSR(n, n, _pos_from_lengths(lengths), cols, vals)


def zipf_index(rng: np.random.Generator, count: int, dim: int, s: float,
               perm: np.ndarray) -> np.ndarray:
    """Indices in [0, dim): Zipf(s) ranks 1..dim mapped through a permutation."""
    ranks = zipf_lengths(rng, count, s, 1, dim) - 1
    return perm[ranks]


@dataclass
class COO3:
    dims: tuple
    idx: np.ndarray   # (nnz, 3) int32, sorted lexicographically
    vals: np.ndarray  # (nnz,) float64

    @property
    def nnz(self) -> int:
        return len(self.vals)


_GEN = None


def _genlib():
    global _GEN
    if _GEN is None:
        import ctypes as C
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libspws_gen.so")
        if not os.path.exists(path):
            from .build import build_gen
            build_gen()
        L = C.CDLL(path)
        L.spwsgen_coo3.restype = C.c_int64
        L.spwsgen_coo3.argtypes = [C.c_int64] * 3 + [C.c_double] * 3 + [C.c_int64, C.c_uint64, C.c_uint64,
                                   C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int]
        L.spwsgen_row_products.restype = None
        L.spwsgen_row_products.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_int]
        L.spwsgen_csr.restype = C.c_int
        L.spwsgen_csr.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64,
                                  C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_int]
        _GEN = L
    return _GEN


def zipf_coo3(dims, nnz: int, exps, seed: int, stream: int, lo: float, hi: float, threads: int = 0) -> COO3:
    """3-order COO with exactly ``nnz`` distinct coordinates (native generator,
    sparsews_amd/native/gen_native.c).

    Mode m draws Zipf(exps[m]) ranks mapped by a seeded permutation of
    [0, dims[m]); ``exps[m] = None`` draws uniformly.  The coordinates are the
    first ``nnz`` distinct triples in draw order, sorted lexicographically;
    values are U[lo, hi).  The result does not depend on the thread count.
    """
    d0, d1, d2 = (int(x) for x in dims)
    idx = np.empty((int(nnz), 3), dtype=np.int32)
    vals = np.empty(int(nnz), dtype=np.float64)
    s = [float(e) if e is not None else 0.0 for e in exps]
    got = _genlib().spwsgen_coo3(d0, d1, d2, s[0], s[1], s[2], int(nnz), int(seed), int(stream),
                                 float(lo), float(hi), idx.ctypes.data, vals.ctypes.data, int(threads))
    if got != nnz:
        raise ValueError("the index space cannot hold that many distinct coordinates")
    return COO3((d0, d1, d2), idx, vals)


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------

def rowwise_pair(m=10_000, n=10_000, r=8, seed=0):
    """Independent A, B for the row-wise SpGEMM (A stream 0, B stream 1)."""
    return uniform_csr(m, n, r, seed, 0), uniform_csr(n, n, r, seed, 1)


def outer_operand(n=200_000, s=1.8, cap=20_000, seed=1):
    """B for the outer-product SpGEMM; A = B^T, so CSC(A) shares B's arrays."""
    return zipf_csr(n, s, cap, seed, 2)


def mttkrp_tensor(dims=(50_000, 40_000, 30_000), nnz=20_000_000, R=32, seed=0):
    rng = rng_for(seed, 11)
    X = zipf_coo3(dims, nnz, (1.5, 1.5, 1.5), seed, 10, 0.0, 1.0)
    Bf = rng.random((dims[1], R))
    Cf = rng.random((dims[2], R))
    return X, Bf, Cf


def ttm_tensor(dims=(100_000, 80_000, 2_000), nnz=10_000_000, R=16, seed=0):
    rng = rng_for(seed, 21)
    X = zipf_coo3(dims, nnz, (1.5, 1.5, None), seed, 20, -1.0, 1.0)
    U = rng.uniform(-1.0, 1.0, (dims[2], R))
    return X, U


@dataclass
class Request:
    rid: int
    n: int
    kind: str         # "uniform" or "powerlaw"
    arrival: float    # seconds from stream start
    seed: int


def request_stream(count=4096, lo=256, hi=65_536, rate=2000.0, seed=0):
    """Request descriptors: size log-uniform in [lo, hi], half uniform-row and
    half power-law-row operands, Poisson arrivals at ``rate`` per second."""
    rng = rng_for(seed, 30)
    sizes = np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), count))).astype(np.int64)
    kinds = np.where(rng.random(count) < 0.5, "uniform", "powerlaw")
    gaps = rng.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    return [Request(i, int(sizes[i]), str(kinds[i]), float(arrivals[i]), seed * 100003 + i)
            for i in range(count)]


def native_csr(n_rows: int, n_cols: int, lengths, seed: int, stream: int, threads: int = 0) -> CSR:
    """CSR with the given row lengths, distinct uniform sorted columns and
    FP64 U[-1,1] values (native generator, sparsews_amd/native/gen_native.c)."""
    lengths = np.ascontiguousarray(lengths, dtype=np.int32)
    pos = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(lengths, out=pos[1:])
    crd = np.empty(int(pos[-1]), dtype=np.int32)
    vals = np.empty(int(pos[-1]), dtype=np.float64)
    rc = _genlib().spwsgen_csr(n_rows, n_cols, lengths.ctypes.data, pos.ctypes.data, int(seed), int(stream),
                               -1.0, 1.0, crd.ctypes.data, vals.ctypes.data, int(threads))
    if rc != 0:
        raise ValueError("a row length exceeds the number of columns")
    return CSR(n_rows, n_cols, pos.astype(np.int32) if pos[-1] < 2**31 - 1 else pos, crd, vals)


def request_operands(req: Request):
    """A_b, B_b for one request: independent square n x n CSR matrices.  Row
    lengths are uniform in 2..64 ("uniform" requests) or Zipf(1.8) on
    [1, min(4096, n)] ("powerlaw" requests), capped at n."""
    mats = []
    for stream in (0, 1):
        r2 = rng_for(req.seed, 40 + stream)
        if req.kind == "uniform":
            lengths = r2.integers(2, 65, req.n).clip(max=req.n)
        else:
            lengths = zipf_lengths(r2, req.n, 1.8, 1, min(4096, req.n))
        mats.append(native_csr(req.n, req.n, lengths, req.seed, 50 + stream))
    return mats[0], mats[1]


def row_products(A: CSR, B: CSR, threads: int = 0):
    """(sum of row upper bounds, largest row upper bound) of A*B."""
    import ctypes as C
    t, m = C.c_int64(), C.c_int64()
    ap = np.ascontiguousarray(A.pos, dtype=np.int32)
    ac = np.ascontiguousarray(A.crd, dtype=np.int32)
    bp = np.ascontiguousarray(B.pos, dtype=np.int32)
    _genlib().spwsgen_row_products(A.nrows, ap.ctypes.data, ac.ctypes.data, bp.ctypes.data, C.addressof(t),
                                   C.addressof(m), int(threads))
    return int(t.value), int(m.value)
