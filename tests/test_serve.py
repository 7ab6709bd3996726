"""The request loop: every job's result is the same alone and in the mix."""
import threading

import numpy as np
import pytest

from casphy_amd import synth


@pytest.mark.gpu
def test_mixed_jobs_equal_standalone(hip):
    rng = np.random.default_rng(60)
    jobs = []
    for k in range(40):
        kind = ("bit", "indel", "transfer")[k % 3]
        n = int(rng.choice([3, 40, 300, 1500, 3000]))
        if kind == "bit":
            jobs.append(("bit", synth.binary_matrix(seed=100 + k, n=n, nsite=256, p=0.02)))
        elif kind == "indel":
            jobs.append(("indel", synth.indel_matrix(seed=100 + k, depth=int(rng.integers(2, 11)), nsite=50)))
        else:
            ntip = int(rng.choice([40, 500, 2000]))
            parent, leaves = synth.yule_parent(ntip, rng)
            Y = synth.clade_bitsets(parent, leaves, ntip)
            X = synth.clade_bitsets(synth.spr(parent, leaves, 5, rng), leaves, ntip)
            jobs.append(("transfer", (Y, X, ntip)))
    srv = hip.Server(hbm_budget_bytes=256 << 20, workers=3)
    ids = [None] * len(jobs)

    def submit(idx):
        for k in idx:
            kind, data = jobs[k]
            ids[k] = srv.submit_transfer(*data) if kind == "transfer" else srv.submit_join(data, kind)

    th = [threading.Thread(target=submit, args=(range(t, len(jobs), 4),)) for t in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k, (kind, data) in enumerate(jobs):
        got, lat = srv.wait(ids[k])
        assert lat >= 0
        if kind == "bit":
            want, _, _ = hip.join_bit(data)
            assert np.array_equal(got, want)
        elif kind == "indel":
            want, _, _ = hip.join_indel(data)
            assert np.array_equal(got, want)
        else:
            want = hip.transfer_dis(*data)
            for g, w in zip(got, want):
                assert np.array_equal(g, w)
    st = srv.stats()
    assert st["completed"] == len(jobs) and st["p99_ms"] >= st["p50_ms"] > 0
    assert st["launches"] < len(jobs)  # small joins were batched
    srv.close()


@pytest.mark.gpu
def test_server_errors_are_per_job(hip):
    srv = hip.Server(workers=1)
    with pytest.raises(hip.CasphyError):
        srv.submit_join(np.zeros((1, 4), np.uint32))  # fewer than two cells
    j = srv.submit_join(np.ones((5, 2), np.uint32))
    e, _ = srv.wait(j)
    assert e.shape == (8, 2)
    srv.close()
