"""The reference programs' BSR input, regenerated with the C library's rand() (the
SpGEMM measurement workload's matrix), against the nnzCub column of the reference's own
logs: logs/sparse/gh200_spgemm_1d_fp16.csv holds, per matrix size 16 b, the block
products nnzCub of the matrix its generator made (srand(0); block kept when
rand() / RAND_MAX > 0.5; values rand() % 3 + 1)."""
import numpy as np
import pytest

from kami_amd import workloads as W


@pytest.mark.parametrize("blocks,nnzcub", [(1, 1), (2, 4), (4, 8), (8, 149), (12, 356)])
def test_generator_reproduces_the_reference_logs_nnzcub(blocks, nnzcub):
    rp, ci, vals = W.reference_bsr(blocks, 16)
    assert rp[0] == 0 and rp[-1] == len(ci) and np.all(np.diff(rp) >= 0)
    for i in range(blocks):  # ascending columns within each block row
        assert np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0)
    assert sum(int(rp[c + 1] - rp[c]) for c in ci) == nnzcub
    assert vals.shape == (len(ci), 16, 16) and set(np.unique(vals)) <= {1.0, 2.0, 3.0}
