"""Argument handling of the drop-in command lines, on the CPU.

`main_BAA n k [read_from_file] [a] [save_iterations] [fp64|split16]`
(main_BAA.cu:134-170) and `generate_bit_transition_cache n`
(generate_bit_transition_cache.cpp:46-52) reject bad arguments before any
GPU call, so these cases run without a GPU.
"""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bdc-upper-bounds-gpu_amd", "bin")


def run(tool, args, cwd):
    path = os.path.join(BIN, tool)
    if not os.path.exists(path):
        pytest.skip(f"{path} is not built (run __graft_entry__.build())")
    return subprocess.run([path, *args], cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize(
    "args,message",
    [
        ([], "Usage"),
        (["28"], "Usage"),
        (["33", "16"], "n must be in [2, 32]"),
        (["1", "1"], "n must be in [2, 32]"),
        (["28", "0"], "k must be in [1, n]"),
        (["28", "29"], "k must be in [1, n]"),
        (["32", "31"], "k <= 30"),
        (["28", "14", "0", "0.005", "100", "fp32"], "precision must be fp64 or split16"),
    ],
)
def test_main_baa_rejects_arguments(tmp_path, args, message):
    r = run("main_BAA", args, tmp_path)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert message in r.stderr, r.stderr
    # nothing was written: the checks come before the context and the files
    assert not (tmp_path / "distributions").exists()


def test_generator_usage(tmp_path):
    r = run("generate_bit_transition_cache", [], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run("generate_bit_transition_cache", ["1", "2"], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_bench_rejects_several_gpus():
    """bench.py runs on one GPU; --gpus above 1 exits with a message (and a
    JSON line saying so) before any GPU work."""
    import json
    import sys
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(repo, "bench.py"), "--gpus", "2", "--steps", "1", "--warmup", "0"],
                       cwd=repo, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
