"""One-off calls from several host threads at once (ctypes releases the GIL).

The library shares a device memory cache, the pinned staging buffers and a
host copy pool between calls, and the tree passes share a free list of
scratch arrays. Four threads compute four different matrices at the same time:
FP32 matrices of 2,560 samples (26 MB, copied through the staging buffers),
FP64, and Faith PD. Each result must equal the same call made alone, bit for
bit, on every one of three rounds.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


def test_concurrent_one_off_calls_equal_calls_alone(ssu, synth):
    jobs = []
    for q, (method, kw) in enumerate([("unweighted", {}), ("weighted_normalized", {}),
                                      ("generalized", {"alpha": 0.5}), ("weighted_unnormalized_fp64", {})]):
        tips = 1500 + 500 * q
        tree = synth.random_tree(tips, 300 + q)
        table = synth.random_table(tips, 2560, 0.03, 310 + q, 1.1)
        jobs.append((ssu.Inputs(*tree, *table), method, kw))

    def run(job):
        inp, method, kw = job
        st, m = ssu.matrix_inmem(inp, method, **kw)
        assert st == ssu.OKAY, ssu.last_error()
        s2, faith, _ = ssu.faith_pd_inmem(inp)
        assert s2 == ssu.OKAY, ssu.last_error()
        return m, faith

    say("alone")
    alone = [run(j) for j in jobs]
    with ThreadPoolExecutor(len(jobs)) as ex:
        for rep in range(3):
            say("together", rep)
            together = list(ex.map(run, jobs))
            for (m0, f0), (m1, f1), (_, method, _) in zip(alone, together, jobs):
                assert np.array_equal(m0, m1), method
                assert np.array_equal(f0, f1), method
