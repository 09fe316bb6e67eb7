"""Writes the small golden fixtures under tests/golden/ (run from the repo root).

* ramp_kat.json: samples of the ramp known-answer product (A[i] = i, B[i] = i, B stored
  N x K; the reference's FP64 self-check inputs, src/block_gemm/KAMI/1d/fp64/
  block_gemm_1d_double_mma.cu:224-236) from the closed form
  C_ij = ijK^3 + (i+j)K^2(K-1)/2 + (K-1)K(2K-1)/6, evaluated in exact integers.
* spgemm_ref_default.npz: the reference's default SpGEMM input (generate_random_bsr_matrix
  with srand(0), 8 x 8 blocks of 16 x 16, sparsity 0.5; src/sparse/spgemm/1d/
  block_spgemm_1d_half_mma.cu:372-397, :411-424) and the output structure of A * A from a
  dense boolean product computed here with numpy (not with the oracle's SPA).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def ramp_exact(i, j, K):
    return i * j * K**3 + (i + j) * K**2 * (K - 1) // 2 + (K - 1) * K * (2 * K - 1) // 6


def main():
    kat = []
    for s in (16, 32, 64, 128):
        samples = []
        for (i, j) in [(0, 0), (0, s - 1), (s - 1, 0), (s - 1, s - 1), (s // 2, s // 3), (1, 2)]:
            samples.append([i, j, ramp_exact(i, j, s)])
        kat.append({"size": s, "samples": samples})
    with open(os.path.join(HERE, "ramp_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)

    rp, ci, vals = oracle.bsr_generate(8, 8, 0.5, 16)
    mask = np.zeros((8, 8), dtype=bool)
    for i in range(8):
        mask[i, ci[rp[i]:rp[i + 1]]] = True
    prod = (mask.astype(np.int64) @ mask.astype(np.int64)) > 0
    rowptrC = np.concatenate([[0], np.cumsum(prod.sum(1))]).astype(np.int32)
    colidxC = np.nonzero(prod)[1].astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "spgemm_ref_default.npz"), rowptr=rp, colidx=ci,
                        vals=vals, rowptrC=rowptrC, colidxC=colidxC)


if __name__ == "__main__":
    main()
