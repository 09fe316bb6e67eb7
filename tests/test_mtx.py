"""Matrix Market -> BSR reader and the BSR archive (kami_amd.mtx), pinned against
scipy.io.mmread / scipy.sparse on the same files."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from kami_amd import mtx


def bsr_dense(mb, nb, rowptr, colidx, blocks):
    bs = blocks.shape[1]
    D = np.zeros((mb * bs, nb * bs))
    for i in range(mb):
        for q in range(rowptr[i], rowptr[i + 1]):
            D[i * bs:(i + 1) * bs, colidx[q] * bs:(colidx[q] + 1) * bs] += blocks[q]
    return D


@pytest.mark.parametrize("shape,density,bs,sym", [((100, 70), 0.05, 16, None),
                                                   ((64, 64), 0.2, 32, "symmetric"),
                                                   ((33, 129), 0.01, 16, None),
                                                   ((48, 48), 0.1, 16, "skew-symmetric")])
def test_matrix_market_to_bsr_matches_scipy(tmp_path, shape, density, bs, sym):
    rng = np.random.default_rng(7)
    M = sp.random(*shape, density=density, random_state=rng, format="coo")
    if sym == "symmetric":
        M = sp.coo_matrix(M + M.T)
    elif sym == "skew-symmetric":
        M = sp.coo_matrix(M - M.T)
    f = tmp_path / "m.mtx"
    scipy.io.mmwrite(str(f), M, symmetry=sym or "general")
    ref = scipy.io.mmread(str(f)).toarray()
    mb, nb, rp, ci, blocks = mtx.load_matrix_market_bsr(f, bs)
    assert (mb, nb) == (-(-shape[0] // bs), -(-shape[1] // bs))
    D = bsr_dense(mb, nb, rp, ci, blocks)
    np.testing.assert_array_equal(D[: shape[0], : shape[1]], ref)
    assert np.all(D[shape[0]:] == 0) and np.all(D[:, shape[1]:] == 0)
    for i in range(mb):
        assert np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0)
    # no block is stored without an entry of the file in it
    r, c = M.row // bs, M.col // bs
    if sym:
        r, c = np.concatenate([r, c]), np.concatenate([c, r])
    assert len(ci) == len(set(zip(r.tolist(), c.tolist())))


def test_pattern_file_and_duplicates(tmp_path):
    f = tmp_path / "p.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n% comment\n"
                 "5 40 4\n1 1\n5 40\n3 17\n3 17\n")
    mb, nb, rp, ci, blocks = mtx.load_matrix_market_bsr(f, 16)
    assert (mb, nb) == (1, 3)
    np.testing.assert_array_equal(rp, [0, 3])
    np.testing.assert_array_equal(ci, [0, 1, 2])
    assert blocks[0, 0, 0] == 1 and blocks[1, 2, 0] == 2 and blocks[2, 4, 7] == 1
    assert blocks.sum() == 4


def test_bad_files_are_refused(tmp_path):
    cases = {"a.mtx": "not a header\n",
             "b.mtx": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
             "c.mtx": "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
             "d.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",
             "e.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n"}
    for name, body in cases.items():
        f = tmp_path / name
        f.write_text(body)
        with pytest.raises(ValueError):
            mtx.read_matrix_market(f)


def test_bsr_archive_round_trip_and_validation(tmp_path):
    rng = np.random.default_rng(3)
    M = sp.random(90, 90, density=0.03, random_state=rng, format="coo")
    mb, nb, rp, ci, blocks = mtx.coo_to_bsr(90, 90, M.row, M.col, M.data, 16)
    f = tmp_path / "a.npz"
    mtx.save_bsr(f, mb, nb, rp, ci, blocks)
    out = mtx.load_bsr(f)
    assert out[:2] == (mb, nb)
    for x, y in zip(out[2:], (rp, ci, blocks)):
        np.testing.assert_array_equal(x, y)
    mtx.save_bsr(f, mb, nb, rp, np.full_like(ci, nb), blocks)  # column out of range
    with pytest.raises(ValueError):
        mtx.load_bsr(f)
    two = np.zeros((2, 16, 16))
    mtx.save_bsr(f, 1, 4, np.array([0, 2]), np.array([3, 1]), two)  # not ascending
    with pytest.raises(ValueError):
        mtx.load_bsr(f)
    mtx.save_bsr(f, 1, 4, np.array([0, 3]), np.array([0, 1]), two)  # rowptr vs arrays
    with pytest.raises(ValueError):
        mtx.load_bsr(f)
