"""The command-line driver (kami_amd/kami_bench, host code over include/kami.h).

CPU: it is built and prints its usage.  GPU: each subcommand runs a small case of the
reference's host programs and must report "Validation successful!" and a
"[hemeng_log]" line in the reference's column layout.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "kami_amd", "kami_bench")


def run(*args, timeout=120):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def test_cli_is_built_and_prints_usage():
    assert os.access(CLI, os.X_OK), "build kami_amd/csrc first"
    r = run("--help")
    assert r.returncode == 0
    for cmd in ("block", "batched", "spgemm", "spmm"):
        assert f"kami_bench {cmd}" in r.stdout


def test_cli_rejects_unknown_subcommand():
    r = run("nonsense")
    assert r.returncode == 2


CASES = [
    (["block", "--prec", "f64", "--m", "64", "--part", "2d", "--count", "256"],
     r"\[hemeng_log\],2d,64,64,64,\d+,\d+,[0-9.e+-]+,\d+"),
    (["block", "--prec", "f64", "--m", "32", "--part", "1d", "--count", "64", "--iters", "10"],
     r"\[hemeng_log\],1d,32,32,32,"),
    (["block", "--prec", "f16", "--m", "128", "--part", "3d", "--count", "128"],
     r"\[hemeng_log\],3d,128,128,128,"),
    (["block", "--prec", "bf16", "--m", "64", "--count", "64"], r"\[hemeng_log\],\dd,64,64,64,"),
    (["block", "--prec", "split16", "--m", "128", "--count", "16"], r"\[hemeng_log\],2d,128,128,128,"),
    (["block", "--prec", "f64", "--m", "128", "--part", "2d", "--pipeline", "2", "--count", "600"],
     r"\[hemeng_log\],2d,128,128,128,"),
    (["batched", "--prec", "f64", "--m", "64", "--count", "300"], r"\[hemeng_log\],1d,64,64,64,300,"),
    (["batched", "--prec", "f16", "--m", "48", "--n", "80", "--k", "112", "--count", "200", "--graph", "0"],
     r"\[hemeng_log\],1d,48,80,112,200,"),
    (["spgemm", "--prec", "f16", "--blocks", "8", "--bs", "16"], r"\[hemeng_log\],2d,\d+,128,128,128,"),
    (["spgemm", "--prec", "f16", "--blocks", "8", "--bs", "16", "--iters", "1000"],
     r"\[hemeng_log\],2d,\d+,128,128,128,"),
    # the reference's measurement layout: the same matrix many times (block-diagonal)
    (["spgemm", "--prec", "f16", "--blocks", "8", "--bs", "16", "--iters", "10", "--copies", "512"],
     r"\[hemeng_log\],2d,149,128,128,128,"),
    (["spgemm", "--prec", "f64", "--blocks", "24", "--bs", "32", "--sparsity", "0.8"],
     r"\[hemeng_log\],2d,\d+,768,768,768,"),
    (["spmm", "--prec", "f16", "--blocks", "8", "--bs", "16"], r"\[hemeng_log\],2d,128,4,"),
    (["spmm", "--prec", "f64", "--blocks", "12", "--bs", "32", "--n", "80"],
     r"\[hemeng_log\],2d,384,4,"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("args,pattern", CASES, ids=[" ".join(c[0][:4]) for c in CASES])
def test_cli_case_validates(args, pattern):
    print(f"\n[case] kami_bench {' '.join(args)}", flush=True)
    r = run(*args)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "Validation successful!" in r.stdout
    assert re.search(pattern, r.stdout), r.stdout
