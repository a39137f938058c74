"""CPU checks of bench.py's own logic: the streamed coverage-text checker
used at full size, and the one-GPU contract of the command line."""
import os
import subprocess
import sys

import numpy as np
import pytest

from util import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("chunk", [1, 2, 3, 7, 64, 1 << 20])
def test_mapper_text_checks_across_chunk_seams(tmp_path, chunk):
    rng = np.random.default_rng(chunk)
    names = [b"chr1", b"chrX", b"c3"]
    lens = [1000, 37, 5000]
    covs = [rng.integers(0, 32, n) for n in lens]
    covs[2][:] = rng.choice([9, 10, 31, 0], lens[2])
    path = tmp_path / "map.txt"
    with open(path, "wb") as f:
        for nm, c in zip(names, covs):
            f.write(nm + b"\n" + b",".join(b"%d" % v for v in c) + b"\n")
    got = bench.mapper_text_checks(str(path), names, lens, chunk=chunk)
    assert got == [(int(c.sum()), n - 1) for c, n in zip(covs, lens)]


def test_mapper_text_checks_rejects_a_wrong_id(tmp_path):
    path = tmp_path / "map.txt"
    path.write_bytes(b"chr2\n1,2,3\n")
    assert bench.mapper_text_checks(str(path), [b"chr1"], [3]) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
