"""Sparse inputs from files: Matrix Market to BSR, and a BSR archive format.

The reference builds every sparse input with its in-program generator
(src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:372-397) and has no file reader.  This
module reads Matrix Market coordinate files (real, integer or pattern; general, symmetric
or skew-symmetric) into the block-sparse row layout the SpGEMM / SpMM entry points take:
row pointers over block rows, ascending block column indices, and bs x bs row-major
blocks in which absent entries are zero.  `save_bsr` / `load_bsr` keep a BSR matrix in a
NumPy archive (no pickled objects).
"""
from __future__ import annotations

import gzip
import os

import numpy as np


def _open_text(path):
    path = os.fspath(path)
    if path.endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path, "r")


def read_matrix_market(path):
    """(nrows, ncols, rows, cols, vals) of a Matrix Market coordinate file, 0-based, with
    symmetric / skew-symmetric storage expanded and duplicate entries kept (they are
    summed when the blocks are built)."""
    with _open_text(path) as f:
        header = f.readline().strip().split()
        if len(header) != 5 or header[0].lower() != "%%matrixmarket":
            raise ValueError(f"{path}: not a Matrix Market file")
        obj, fmt, field, sym = (h.lower() for h in header[1:])
        if obj != "matrix" or fmt != "coordinate":
            raise ValueError(f"{path}: only 'matrix coordinate' files are supported")
        if field not in ("real", "integer", "pattern", "double"):
            raise ValueError(f"{path}: field '{field}' is not supported")
        if sym not in ("general", "symmetric", "skew-symmetric"):
            raise ValueError(f"{path}: symmetry '{sym}' is not supported")
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: missing size line")
        nrows, ncols, nnz = (int(x) for x in line.split())
        ncol_fields = 2 if field == "pattern" else 3
        data = np.loadtxt(f, ndmin=2, comments="%", max_rows=nnz) if nnz else np.zeros((0, ncol_fields))
    if data.shape[0] != nnz or data.shape[1] < ncol_fields:
        raise ValueError(f"{path}: expected {nnz} entries with {ncol_fields} fields")
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = np.ones(nnz) if field == "pattern" else data[:, 2].astype(np.float64)
    if np.any(rows < 0) or np.any(rows >= nrows) or np.any(cols < 0) or np.any(cols >= ncols):
        raise ValueError(f"{path}: entry index out of range")
    if sym != "general":
        off = rows != cols
        sign = -1.0 if sym == "skew-symmetric" else 1.0
        rows, cols, vals = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, sign * vals[off]]))
    return nrows, ncols, rows, cols, vals


def coo_to_bsr(nrows, ncols, rows, cols, vals, bs):
    """BSR of a coordinate matrix padded to whole blocks: (mb, nb, rowptr int32,
    colidx int32 ascending per block row, blocks float64 (nnzb, bs, bs)).  Duplicates add."""
    mb, nb = -(-nrows // bs), -(-ncols // bs)
    rows, cols = np.asarray(rows, np.int64), np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.float64)
    br, bc = rows // bs, cols // bs
    key = br * nb + bc
    uniq, inv = np.unique(key, return_inverse=True)  # sorted: block row, then block column
    blocks = np.zeros((len(uniq), bs, bs), dtype=np.float64)
    np.add.at(blocks, (inv, rows % bs, cols % bs), vals)
    ubr = (uniq // nb).astype(np.int64)
    rowptr = np.zeros(mb + 1, dtype=np.int32)
    np.add.at(rowptr, ubr + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    colidx = (uniq % nb).astype(np.int32)
    return mb, nb, rowptr, colidx, blocks


def load_matrix_market_bsr(path, bs):
    """Matrix Market file -> BSR tuple as `coo_to_bsr` returns it."""
    return coo_to_bsr(*read_matrix_market(path), bs)


def save_bsr(path, mb, nb, rowptr, colidx, blocks):
    np.savez(path, shape=np.array([mb, nb, blocks.shape[1] if blocks.ndim == 3 else 0]),
             rowptr=np.asarray(rowptr, np.int32), colidx=np.asarray(colidx, np.int32),
             blocks=np.asarray(blocks))


def load_bsr(path):
    with np.load(path, allow_pickle=False) as z:
        mb, nb, bs = (int(x) for x in z["shape"])
        rowptr, colidx, blocks = z["rowptr"], z["colidx"], z["blocks"]
    if rowptr.shape != (mb + 1,) or rowptr[0] != 0 or np.any(np.diff(rowptr) < 0):
        raise ValueError(f"{path}: bad row pointer")
    if len(colidx) != rowptr[-1] or blocks.shape != (len(colidx), bs, bs):
        raise ValueError(f"{path}: arrays disagree with the row pointer")
    if len(colidx) and (colidx.min() < 0 or colidx.max() >= nb):
        raise ValueError(f"{path}: block column out of range")
    for i in range(mb):
        if np.any(np.diff(colidx[rowptr[i]:rowptr[i + 1]]) <= 0):
            raise ValueError(f"{path}: block columns of row {i} are not ascending")
    return mb, nb, rowptr, colidx, blocks
