"""CPU checks of bench.py's host logic: the one-GPU contract and the seeded batched-queue request list."""
import math
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_batch_requests_are_seeded_and_in_range():
    import bench

    a, b = bench.c3_requests(), bench.c3_requests()
    assert a == b
    assert len(a) == bench.C3["requests"]
    lo_res, hi_res = bench.C3["res"]
    lo_spp, hi_spp = bench.C3["spp"]
    kinds = set()
    for q in a:
        kinds.add(q["kind"])
        assert 0 <= q["kind"] < len(bench.C3["volumes"])
        assert lo_res <= q["side"] <= hi_res
        assert lo_spp <= q["spp"] <= hi_spp
        assert abs(math.sqrt(sum(c * c for c in q["camera"])) - 0.8) < 1e-9
        assert abs(math.sqrt(sum(c * c for c in q["sun"])) - 1.0) < 1e-9 and q["sun"][1] > 0.0
    assert kinds == set(range(len(bench.C3["volumes"])))

