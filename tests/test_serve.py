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


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_serving_mix_mid_size_equal_standalone(hip):
    """The serving mix at mid size: 200 jobs of 1k..30k cells -- binary and
    indel tree builds over rows of resident matrices (the config-2 and config-1
    site models) and transfer distances from trees at the job's size --
    submitted from four threads into three workers under a 512 MiB budget,
    with the resident matrices dropped before the results are read; every
    job's output equals the same computation alone."""
    bmat = synth.binary_matrix(seed=51, n=30000, nsite=4096)
    imat = synth.indel_matrix(seed=52, depth=15, nsite=200)
    jobs = synth.serving_jobs(53, 200, 1000, 30000, bmat.shape[0], imat.shape[0])
    srv = hip.Server(hbm_budget_bytes=512 << 20, workers=3)
    mids = {"bit": srv.put_matrix(bmat, "bit"), "indel": srv.put_matrix(imat, "indel")}
    ids = [None] * len(jobs)

    def submit(idx):
        for k in idx:
            kind, data = jobs[k]
            ids[k] = srv.submit_transfer_trees(*data) if kind == "trees" else srv.submit_join_rows(mids[kind], data)

    th = [threading.Thread(target=submit, args=(range(t, len(jobs), 4),)) for t in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for mid in mids.values():
        srv.drop_matrix(mid)  # submitted jobs keep their matrix
    for k, (kind, data) in enumerate(jobs):
        got, lat = srv.wait(ids[k])
        if kind == "trees":
            want = hip.transfer_trees(*data)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), k
            for g, w in zip(got[2] + got[3], want[2] + want[3]):
                assert np.array_equal(g, w), k
        else:
            want, _, _ = (hip.join_bit if kind == "bit" else hip.join_indel)((bmat if kind == "bit" else imat)[data])
            assert np.array_equal(got, want), (k, kind, data.size)
    st = srv.stats()
    assert st["completed"] == len(jobs) and st["p99_ms"] >= st["p50_ms"] > 0
    with pytest.raises(hip.CasphyError):
        srv.submit_join_rows(mids["bit"], np.arange(4, dtype=np.uint32))  # dropped
    srv.close()


@pytest.mark.gpu
def test_server_resident_rows_errors(hip):
    srv = hip.Server(workers=2)
    m = srv.put_matrix(np.ones((10, 3), np.uint32), "bit")
    with pytest.raises(hip.CasphyError, match="out of range"):
        srv.submit_join_rows(m, np.array([0, 10], np.uint32))
    with pytest.raises(hip.CasphyError):
        srv.submit_join_rows(m, np.array([3], np.uint32))  # fewer than two cells
    with pytest.raises(hip.CasphyError, match="unknown matrix"):
        srv.submit_join_rows(m + 1000, np.array([0, 1], np.uint32))
    e, _ = srv.wait(srv.submit_join_rows(m, np.array([9, 2, 5], np.uint32)))
    assert e.shape == (4, 2)
    srv.close()
