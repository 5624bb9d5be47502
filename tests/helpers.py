"""Shared test helpers: small seeded matrices and independent references."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import scipy.sparse as sp

from sparsews_amd.gen import CSR, distinct_sorted_columns, rng_for, uniform_values, zipf_lengths


def random_csr(m, n, lengths, seed, stream=0):
    rng = rng_for(seed, stream)
    lengths = np.asarray(lengths, dtype=np.int64)
    cols = distinct_sorted_columns(rng, lengths, n)
    vals = uniform_values(rng, len(cols))
    pos = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(lengths, out=pos[1:])
    return CSR(m, n, pos.astype(np.int32), cols, vals)


def uniform_rows(m, n, lo, hi, seed, stream=0):
    rng = rng_for(seed, 100 + stream)
    return random_csr(m, n, rng.integers(lo, hi + 1, m).clip(max=n), seed, stream)


def powerlaw_rows(m, n, s, cap, seed, stream=0, lo=0):
    rng = rng_for(seed, 200 + stream)
    lengths = zipf_lengths(rng, m, s, max(lo, 1), min(cap, n))
    if lo == 0:
        lengths[rng.random(m) < 0.1] = 0  # some empty rows
    return random_csr(m, n, lengths, seed, stream)


def pattern_product(A: CSR, B: CSR):
    """Pattern of A*B and the number of products per entry, from integer
    counts (independent of value cancellation).  Returns (pos, crd, count)."""
    Ab = sp.csr_matrix((np.ones(A.nnz, dtype=np.int64), A.crd, A.pos), shape=(A.nrows, A.ncols))
    Bb = sp.csr_matrix((np.ones(B.nnz, dtype=np.int64), B.crd, B.pos), shape=(B.nrows, B.ncols))
    P = (Ab @ Bb).tocsr()
    P.sort_indices()
    return P.indptr.astype(np.int64), P.indices.astype(np.int32), P.data


def scipy_product(A: CSR, B: CSR):
    P = (A.to_scipy() @ B.to_scipy()).tocsr()
    P.sort_indices()
    return P


def dense(A: CSR) -> np.ndarray:
    return A.to_scipy().toarray()


def assert_same_structure(C: CSR, pos, crd):
    assert C.nnz == len(crd), (C.nnz, len(crd))
    np.testing.assert_array_equal(np.asarray(C.pos, dtype=np.int64), np.asarray(pos, dtype=np.int64))
    np.testing.assert_array_equal(np.asarray(C.crd, dtype=np.int64), np.asarray(crd, dtype=np.int64))


# --------------------------------------------------------------------------- workspace records

def random_records(rng, n, rows, cols, empty_frac=0.0):
    """n workspace records with distinct (crd0, crd1) in [0,rows)x[0,cols),
    a fraction of them empty (crd0 = -1, crd1 arbitrary)."""
    from oracle import WSPACE_DTYPE
    keys = rng.choice(rows * cols, size=n, replace=False)
    w = np.zeros(n, dtype=WSPACE_DTYPE)
    w["crd"][:, 0] = keys // cols
    w["crd"][:, 1] = keys % cols
    w["val"] = rng.uniform(-1.0, 1.0, n)
    empty = rng.random(n) < empty_frac
    w["crd"][empty, 0] = -1
    w["crd"][empty, 1] = rng.integers(-5, 5, int(empty.sum()))
    return w


def sorted_coo(rng, n, rows, cols):
    keys = np.sort(rng.choice(rows * cols, size=n, replace=False))
    return (keys // cols).astype(np.int32), (keys % cols).astype(np.int32), rng.uniform(-1.0, 1.0, n)


# --------------------------------------------------------------------------- drop-in header drivers

COMPAT_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "compat")


def build_compat_driver(name, outdir):
    """Compile tests/compat/<name>.cpp (a reference-style driver over the
    drop-in headers) against libspws with g++; returns the binary path."""
    from sparsews_amd import build as B
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.dirname(B.LIB)
    exe = os.path.join(str(outdir), name)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(repo, "include"), "-o", exe,
           os.path.join(COMPAT_DIR, name + ".cpp"), "-L", libdir, "-lspws", "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def write_frames(path, arrays):
    with open(path, "wb") as f:
        for a in arrays:
            b = np.ascontiguousarray(a).tobytes()
            f.write(np.int64(len(b)).tobytes())
            f.write(b)


def read_frames(path, dtypes):
    out = []
    with open(path, "rb") as f:
        for dt in dtypes:
            n = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
            out.append(np.frombuffer(f.read(n), dtype=dt).copy())
    return out
