"""Tuning choices saved to and loaded from a file (kami_tune_save / kami_tune_load).

CPU only: a loaded block-GEMM choice changes what KAMI_PART_AUTO resolves to (the query
needs no device), a save writes back exactly what was loaded, and malformed files are
refused without applying anything.
"""
import pytest

import kami_amd as K
from kami_amd import _abi


@pytest.fixture(autouse=True)
def _clean():
    K.lib().kami_tune_clear()
    yield
    K.lib().kami_tune_clear()


def load(path):
    return K.lib().kami_tune_load(str(path).encode())


def save(path):
    return K.lib().kami_tune_save(str(path).encode())


def test_loaded_block_choice_drives_auto(tmp_path):
    before = K.query(64, 64, 64, K.PREC_F64)
    assert before["partition"] != 3                           # the measured prior is not 3D
    f = tmp_path / "t.txt"
    f.write_text("kami-tune 1\nblock 0 64 64 64 3 2 1\n")
    assert load(f) == _abi.OK
    after = K.query(64, 64, 64, K.PREC_F64)
    assert (after["partition"], after["ranks"]) == (3, 2)
    # an explicit partition still wins over the recorded choice
    assert K.query(64, 64, 64, K.PREC_F64, partition="2d")["partition"] == 2


def test_save_round_trip(tmp_path):
    src = tmp_path / "in.txt"
    src.write_text("kami-tune 1\n# comment\nblock 1 128 128 128 2 2 2\n"
                   "block 0 32 32 32 1 2 1\nbatched 0 62 3\nbatched 1 0 1\n")
    assert load(src) == _abi.OK
    out = tmp_path / "out.txt"
    assert save(out) == _abi.OK
    lines = out.read_text().splitlines()
    assert lines[0] == "kami-tune 1"
    assert sorted(lines[1:]) == sorted(["block 1 128 128 128 2 2 2", "block 0 32 32 32 1 2 1",
                                        "batched 0 62 3", "batched 1 0 1"])
    cfg = K.query(128, 128, 128, K.PREC_F16_F32)
    assert (cfg["partition"], cfg["ranks"], cfg["pipeline"]) == (2, 2, 2)


@pytest.mark.parametrize("body", [
    "block 0 64 64 64 2 2 1\n",                      # no header
    "kami-tune 2\nblock 0 64 64 64 2 2 1\n",         # unknown version
    "kami-tune 1\nblock 0 96 96 96 2 2 1\n",         # no compiled instance of that shape
    "kami-tune 1\nblock 0 64 64 64 3 7 1\n",         # no such wave grid
    "kami-tune 1\nblock 3 128 128 128 2 4 1\n",      # split precision has no tuned partition
    "kami-tune 1\nbatched 0 125 0\n",                # bucket out of range
    "kami-tune 1\nbatched 0 3 6\n",                  # variant out of range
    "kami-tune 1\nbatched 3 3 1\n",                  # batched has no split precision
    "kami-tune 1\nsomething else\n",
])
def test_malformed_file_is_refused_and_nothing_applied(tmp_path, body):
    f = tmp_path / "bad.txt"
    f.write_text(body)
    assert load(f) == _abi.EINVAL
    out = tmp_path / "out.txt"
    assert save(out) == _abi.OK
    assert out.read_text().splitlines() == ["kami-tune 1"]


def test_missing_file_and_null_path(tmp_path):
    assert load(tmp_path / "absent.txt") == _abi.EINVAL
    assert K.lib().kami_tune_load(None) == _abi.EINVAL
    assert K.lib().kami_tune_save(None) == _abi.EINVAL
