"""Benchmark of the CasPhy hot path on one MI355X.

Headline workload (the DNA configuration): 100,000 sequences x 1,000 columns
(ACGT with 2% gaps, evolved down a Yule tree under Jukes-Cantor, seed 3) in
fasta2bin byte codes, resident in HBM.  One step is the autoSplit radius
search of R/CasPhy_functions.R:1964-2025 as R runs it for DNA
(radius_initial 100, step 10, nReplicate 1,024, nNeighborBound 9.5) over the
resident cells:

  1. tryRadiusDNA for 1,024 seeded query rows at the radii 100, 110, ...,
     64 radii per device round, until a radius r passes the rule
     log10(mean count) <= 9.5 - log10(n), chosen on the device;
  2. in the round where one passes, radiusNeighborDNA lists at r for all
     100,000 sequences (upper-triangle pair tiles -> neighbour bitmask ->
     counts -> scan -> on-device compaction; 2.6e9 ids, more than 2^31).
Each round is one device pass with one host synchronisation
(casphy_cells_autosplit); before timing, the result is checked against the
host-driven loop (tryRadius counts back to the host, host rule,
radiusNeighbor) and sampled rows against the CPU oracle.

value = cell-pair scores per second: (rounds x 1,024 x n + n(n+1)/2) scores
per step.  The sequences are resident before timing; per round only the radii
and the query list (4 KB) go up and the choice and list length (16 B) come
back.  The other configurations are reported under "extra".
"""
from __future__ import annotations

import argparse
import functools
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# the DNA configuration (headline)
N3, NCOL3, R = 100000, 1000, 1024
DNA_RADIUS_INITIAL, DNA_STEP = 100, 10   # autoSplit defaults, R/CasPhy_functions.R:1964-1966
AUTOSPLIT_BOUND = 9.5  # nNeighborBound default of the autoSplit loop, R/CasPhy_functions.R:1967
# the base-editor configuration
N, NSITE = 10000, 4096
RADII = np.arange(1, 17, dtype=np.int32)
CHECK_REPLAYS = 50  # repeated device rounds checked against the first before timing (base-editor round)
VALU_SPEC_TOPS = 256 * 64 * 2.4e9 / 1e12  # MI355X: 256 CUs x 64 lanes x 2.4 GHz, int32 lane-ops (data sheet)
MFMA_I8_PEAK_TOPS = 5000.0                # MI355X dense int8 matrix-core peak, 5.0 POP/s
MFMA_FP4_PEAK_TOPS = 10100.0              # MI355X dense fp4 matrix-core peak, 10.1 PFLOP/s
HBM_PEAK_GBS = 8000.0
CPU_THREADS_MAX = 16                      # host threads of the CPU baselines, at most


@functools.lru_cache(maxsize=None)
def host_info() -> dict:
    """Core count and CPU model of the host the CPU baselines run on."""
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"nproc": aff, "cpu_count": os.cpu_count(), "model": model}


def cpu_threads() -> int:
    return max(1, min(CPU_THREADS_MAX, host_info()["nproc"]))


def cpu_rows_timed(fn, nrows: int, chunk: int, threads: int, budget_s: float, check=None):
    """Runs fn(i0, i1) over consecutive row chunks on `threads` host threads
    (ctypes calls release the GIL) until `budget_s` has passed; check(i0,
    result) compares each chunk with the GPU.  Returns (rows done, seconds)."""
    from concurrent.futures import ThreadPoolExecutor
    rows_done, nxt = 0, 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < budget_s and nxt < nrows:
            jobs = []
            for _ in range(threads):
                if nxt >= nrows:
                    break
                i1 = min(nxt + chunk, nrows)
                jobs.append((nxt, i1, ex.submit(fn, nxt, i1)))
                nxt = i1
            for i0, i1, f in jobs:
                res = f.result()
                if check is not None:
                    check(i0, i1, res)
                rows_done += i1 - i0
    return rows_done, time.perf_counter() - t0


def cpu_radius_baseline(kind: str, mat: np.ndarray, radius: int, counts: np.ndarray, idx: np.ndarray,
                        offsets: np.ndarray, budget_s: float = 8.0, stride: int = 1) -> dict:
    """The reference's CPU radius path (radiusWeightedParallel /
    radiusWeightedDNAParallel restated in the oracle: per query row, one pass
    over every target) on bounded samples of query rows, on 1 host thread and
    on all of this process's host threads; each chunk's counts and ids are
    compared with the GPU lists.  Rows are taken every `stride` rows."""
    from oracle import oracle
    n = mat.shape[0]
    rows = np.arange(0, n, stride)

    def run(i0, i1):
        out = []
        for k in range(i0, i1):
            r = int(rows[k])
            out.append((r, oracle.radius_pass(kind, mat, radius, (r, r + 1))))
        return out

    def check(i0, i1, res):
        for r, (c, ids) in res:
            assert c[0] == counts[r], "CPU/GPU count mismatch"
            assert np.array_equal(ids, idx[offsets[r]:offsets[r + 1]]), "CPU/GPU list mismatch"

    out = {}
    for threads in (1, cpu_threads()):
        done, el = cpu_rows_timed(run, rows.size, 2, threads, budget_s, check)
        out[threads] = (done * n / el, done, el)
    t1, tall = 1, cpu_threads()
    return {"value": out[tall][0], "unit": "cell-pair scores/s", "cores": tall, "kind": "port",
            "value_1_thread": out[t1][0], "host": host_info(),
            "sample": f"{out[tall][1]} query rows x {n} targets at radius {radius} on {tall} threads "
                      f"({out[tall][2]:.1f} s), {out[t1][1]} rows on 1 thread ({out[t1][2]:.1f} s); "
                      f"every row's count and ids equal the GPU lists"}


def autosplit_radius(counts: np.ndarray, n: int) -> int:
    """Host form of the autoSplit stopping rule over the tried radii 1..16."""
    bound = AUTOSPLIT_BOUND - np.log10(n)
    for k, r in enumerate(RADII):
        if np.log10(max(counts[k].mean(), 1e-30)) <= bound:
            return int(r)
    return -1


ROUND_RADII = 64  # radii per device round (counting passes of 16 over one score matrix)


def dna_search(cells, q, n: int, bound: float, start: int = DNA_RADIUS_INITIAL, step: int = DNA_STEP,
               limit: int = NCOL3):
    """R's autoSplit radius search on the device: rounds of 64 radii
    (start, start + step, ...) until one passes; returns (radius, ids, rounds)."""
    rounds = 0
    while True:
        radii = np.arange(start, start + ROUND_RADII * step, step, dtype=np.int32)
        r, total = cells.autosplit(radii, q, bound)
        rounds += 1
        if r >= 0 or radii[-1] >= limit:
            return r, total, rounds
        start += ROUND_RADII * step


def dna_search_host(cells, q, n: int, bound: float, start: int = DNA_RADIUS_INITIAL, step: int = DNA_STEP,
                    limit: int = NCOL3):
    """The same search as R runs it with the executables: one tryRadius pass
    per radius, the rule on the host, then the radiusNeighbor lists."""
    r = start
    tries = 0
    while r <= limit:
        cnt = cells.try_radii([r], q)[0]
        tries += 1
        if np.log10(max(cnt.mean(), 1e-30)) <= bound:
            return r, cells.radius(r), tries
        r += step
    return -1, 0, tries


@functools.lru_cache(maxsize=None)
def ceilings() -> dict:
    """Measured instruction-rate ceilings (casphy_amd/calib/calib.hip): the
    32-bit integer VALU issue rate (8 independent xor / and / popcount-add
    chains per lane, 8 waves per SIMD) and the fp4 matrix-core rate
    (v_mfma_scale_f32_32x32x64_f8f6f4 on register operands, 4 accumulators per
    wave, 4 waves per SIMD), with the same operands every instruction and with
    operands that change before every instruction as a tile kernel's do; the
    best of 3 timed launches after one warm-up."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "casphy_amd", "libcasphy_calib.so"))
    hip = C.CDLL("libamdhip64.so")
    vblocks, mblocks = 256 * 8, 256 * 4
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), C.c_size_t(max(vblocks, mblocks) * 256 * 4)) == 0
    ms, work = C.c_float(), C.c_uint64()
    best = {}
    try:
        for name, fn, blocks, iters in (("valu", lib.casphy_calib_valu, vblocks, 10000),
                                        ("mfma_fp4", lib.casphy_calib_mfma_fp4, mblocks, 100000),
                                        ("mfma_fp4_data", lib.casphy_calib_mfma_fp4_data, mblocks, 100000)):
            rates = []
            for k in range(4):
                assert fn(buf, C.c_uint32(blocks), C.c_uint32(iters), None, C.byref(ms), C.byref(work)) == 0
                if k:
                    rates.append(work.value / (ms.value * 1e-3) / 1e12)
            best[name] = max(rates)
    finally:
        hip.hipFree(buf)
    return {"valu_tlane_ops_per_s": best["valu"], "valu_spec_tlane_ops_per_s": VALU_SPEC_TOPS,
            "mfma_fp4_tops": best["mfma_fp4"], "mfma_fp4_switching_operands_tops": best["mfma_fp4_data"],
            "mfma_fp4_spec_tops": MFMA_FP4_PEAK_TOPS,
            "kernels": "k_valu_rate (192 VALU instructions per loop trip), k_mfma_fp4_rate<false / true>"}


def calibration_copy(nbytes: int = 256 << 20) -> None:
    """One device-to-device copy of `nbytes` (the HIP runtime's copy kernel),
    the known-size reference against which FETCH_SIZE / WRITE_SIZE counters of
    the pair-tile kernel are scaled."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


# kernel behind each matrix-core tile variant with the radius-mask epilogue
VARIANT_KERNEL = {"128x1": "k_mfma_tiles<0, 128, 1, 2>", "128x2": "k_mfma_tiles<0, 128, 2, 2>",
                  "128x4": "k_mfma_tiles<0, 128, 4, 2>", "64x2": "k_mfma_tiles<0, 64, 2, 4>",
                  "64x4": "k_mfma_tiles<0, 64, 4, 4>", "64x8": "k_mfma_tiles<0, 64, 8, 4>",
                  "reg": "k_mfma_reg<0>", "wide": "k_mfma_wide<0>", "wide8": "k_mfma_wide8<0>",
                  "fp4": "k_mfma_fp4<0>", "fp4b": "k_mfma_fp4b<0>", "fp4d": "k_mfma_fp4d<0>",
                  "fp4w": "k_mfma_fp4w<0>", "fp4x": "k_mfma_fp4x<0>",
                  "fp4y": "k_mfma_fp4y<0>", "fp4p": "k_mfma_fp4p<0>"}
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "k_mfma_tiles_mask_10000x4096_pmc.md")
TRAFFIC_TABLE_DNA = os.path.join(ROOT, "profiles", "k_mfma_mask_dna_100000x1000_pmc.md")


def measured_traffic(kernel: str, table: str = TRAFFIC_TABLE):
    """HBM bytes per launch of `kernel` on a workload, from the committed
    summary of two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, each scaled
    by what the counter reads for a 256 MiB copy); None when not measured."""
    if not os.path.exists(table):
        return None
    for line in open(table):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0] == kernel:
            return int(cells[-1].replace(",", ""))  # last column: HBM bytes per launch
    return None


def timed(fn, reps=1):
    t0 = time.perf_counter()
    out = None
    for _ in range(reps):
        out = fn()
    return (time.perf_counter() - t0) / reps, out


def mask_roofline(cp, tile_ms: float, pairs: int, ops_per_pair: int, shape: tuple, traffic_table: str,
                  valu: bool = False, valu_ops_per_pair: int = 0) -> dict:
    """Roofline of a radius-mask launch: achieved ops over its hipEvent
    duration against the peak of the form the tile autotuner chose for the
    shape class (epilogue 0, triangle, log2 rows, sites)."""
    tile_s = tile_ms / 1000
    if valu:
        roof = {"bound": "valu", "achieved": valu_ops_per_pair * pairs / tile_s / 1e12,
                "peak": ceilings()["valu_tlane_ops_per_s"], "peak_kind": "measured VALU issue ceiling",
                "unit": "Tlane-op/s", "kernel": "k_tiles<DNA2, MASK>", "variant": "valu"}
    else:
        tuning = cp.mfma_tuning()
        chosen = [e["chosen"] for e in tuning if e.get("epilogue") == 0 and e.get("triangle") == 1 and
                  (e.get("log2_rows_a"), e.get("sites")) == shape]
        forced = os.environ.get("CASPHY_MFMA_VARIANT")
        variant = forced if forced in VARIANT_KERNEL else (chosen[0] if chosen else "64x2")
        fp4 = variant.startswith("fp4")
        roof = {"bound": "mfma", "achieved": ops_per_pair * pairs / tile_s / 1e12,
                "peak": MFMA_FP4_PEAK_TOPS if fp4 else MFMA_I8_PEAK_TOPS, "unit": "TOP/s",
                "kernel": VARIANT_KERNEL[variant], "variant": variant}
    roof["frac"] = roof["achieved"] / roof["peak"]
    if not valu and roof["variant"].startswith("fp4"):
        roof["frac_of_measured_ceiling"] = roof["achieved"] / ceilings()["mfma_fp4_tops"]
        roof["frac_of_switching_operand_ceiling"] = (roof["achieved"] /
                                                     ceilings()["mfma_fp4_switching_operands_tops"])
    roof["traffic"] = measured_traffic(roof["kernel"], traffic_table)
    roof["kernel_ms"] = tile_ms
    roof["algorithmic"] = (f"{valu_ops_per_pair if valu else ops_per_pair} ops per cell pair x {pairs} pairs "
                           f"per launch")
    return roof


def base_editor_round(cp, synth, with_cpu: bool, steps: int = 100, warmup: int = 10) -> dict:
    """The base-editor configuration (10,000 cells x 4,096 sites, seed 2): one
    device autoSplit round (tryRadius of 1,024 queries at radii 1..16, the
    radius choice, radiusNeighbor lists) per step on the resident .bmat, after
    checking it against the host-driven loop and 50 graph replays; the mask
    tiles' roofline from their hipEvents; the CPU path on 1 and all threads."""
    mat = synth.binary_matrix(seed=2, n=N, nsite=NSITE, p=0.005, missing=0.05)
    q = synth.query_rows(2, N, R)
    cells = cp.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
    bound = AUTOSPLIT_BOUND - np.log10(N)
    counts = cells.try_radii(RADII, q)
    r_host = autosplit_radius(counts, N)
    t_host = cells.radius(r_host)
    c_host, i_host = cells.fetch()
    r_dev, t_dev, c_dev = cells.autosplit(RADII, q, bound, counts=True)
    c_d, i_d = cells.fetch()
    assert r_dev == r_host and t_dev == t_host and np.array_equal(c_dev, counts)
    assert np.array_equal(c_d, c_host) and np.array_equal(i_d, i_host)
    for _ in range(CHECK_REPLAYS):
        assert cells.autosplit(RADII, q, bound) == (r_dev, t_dev)
    for _ in range(warmup):
        cells.autosplit(RADII, q, bound)
    tile_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        cells.autosplit(RADII, q, bound)
        tile_ms.append(cells.last_ms()[0])
    el = time.perf_counter() - t0
    pairs = N * (N + 1) // 2 + R * N
    out = {"cells": N, "sites": NSITE, "queries": R, "radius": r_dev, "neighbour_ids": int(t_dev),
           "steps": steps, "ms_per_round": 1e3 * el / steps, "cell_pair_scores_per_s": pairs * steps / el,
           "roofline": mask_roofline(cp, float(np.mean(tile_ms)), N * (N + 1) // 2, 2 * NSITE, (14, NSITE),
                                     TRAFFIC_TABLE)}
    if with_cpu:
        off = np.zeros(N + 1, np.int64)
        off[1:] = np.cumsum(c_d, dtype=np.int64)
        out["cpu_baseline"] = cpu_radius_baseline("bit", mat, r_dev, c_d, i_d, off, budget_s=6.0)
    del c_host, i_host, c_d, i_d
    # tryRadius: one launch per radius (the autoSplit loop) vs all 16 radii at once
    plain_s, per = timed(lambda: [cells.try_radii([r], q) for r in RADII], 3)
    batch_s, allr = timed(lambda: cells.try_radii(RADII, q), 3)
    assert all(np.array_equal(per[k][0], allr[k]) for k in range(len(RADII)))
    out["tryradius_16_radii"] = {"plain_ms": plain_s * 1e3, "batched_ms": batch_s * 1e3, "speedup": plain_s / batch_s}
    # binary tree build of the 10,000-cell matrix
    cp.join_bit(mat)  # first call of the shape class runs the tile autotuner
    join_s, _ = timed(lambda: cp.join_bit(mat), 1)
    out["binary_join"] = {"plain_jobs_per_s": 1 / join_s, "ms_per_job": join_s * 1e3}
    out["nodekeep"] = node_keep_step(cp, mat, with_cpu)
    out["tipmerge"], member = tip_merge_step(cp, cells, q, with_cpu)
    out["split_layer"] = split_layer_step(cp, mat, cells, member, out["tipmerge"]["radius"])
    cells.close()
    return out


SPLIT_MIN = 10  # regularize.split's min_size default (R/CasPhy_functions.R:2896-2910)


def split_layer_step(cp, mat, cells, member, radius: int) -> dict:
    """The next split layer (run.split per group, R/CasPhy_functions.R:2896-2937):
    every tipMerge cluster of at least 10 cells is a group whose autoSplit
    search starts at the parent's radius (16 radii, step 20, nNeighborBound
    9.5, up to 1,024 query rows per group).  Plain: the way R runs it, one
    sub-matrix per group made resident and searched with its own device round;
    batched: one group search over the resident parent matrix.  Every group's
    radius and lists are compared between the two."""
    rng = np.random.default_rng(8)
    labels = np.unique(member)
    groups = [np.flatnonzero(member == lab).astype(np.uint32) for lab in labels]
    groups = [g for g in groups if g.size >= SPLIT_MIN]
    queries = [rng.integers(0, g.size, min(g.size, R)).astype(np.uint32) for g in groups]
    radii = (radius + 20 * np.arange(16)).astype(np.int32)
    cells.split_groups(groups, radii, queries)
    batch_s, (rad, nidx) = timed(lambda: cells.split_groups(groups, radii, queries), 5)
    batched = [cells.group_fetch(g) for g in range(len(groups))]

    def plain():
        res = []
        for g, rows in enumerate(groups):
            sub = cp.Cells(np.ascontiguousarray(mat[rows].T), "bit", wordmajor=True)
            r, total = sub.autosplit(radii, queries[g], 9.5 - float(np.log10(rows.size)))
            res.append((r, sub.fetch() if total else None))
            sub.close()
        return res
    plain()
    plain_s, res = timed(plain, 1)
    for g in range(len(groups)):
        assert res[g][0] == rad[g], "plain / batched split layer radius mismatch"
        if res[g][1] is not None:
            assert np.array_equal(res[g][1][0], batched[g][0]) and np.array_equal(res[g][1][1], batched[g][1])
    sizes = np.array([g.size for g in groups])
    return {"groups": len(groups), "cells": int(sizes.sum()), "largest": int(sizes.max()),
            "median": float(np.median(sizes)), "neighbour_ids": int(nidx.sum()),
            "plain_ms": plain_s * 1e3, "batched_ms": batch_s * 1e3, "speedup": plain_s / batch_s,
            "plain_form": "one resident sub-matrix and one device autoSplit round per group",
            "bit_exact_between_forms": True}


def indel_config(cp, synth, with_cpu: bool) -> dict:
    """The indel configuration (1,024 cells x 200 sites, seed 1): the full
    shared-mutation score matrix, one tree build, and 256 bootstrap
    replicates (site resampling) in one persistent launch, from replicate
    matrices built on the host and from the column indices alone (the
    replicates built on the device); the CPU path (greedyJoinEdge2 and
    countApoFloat restated in the oracle) on 1 thread and on all threads
    (replicates across threads, as R's fork cluster)."""
    base = synth.indel_matrix(seed=1)
    n, ncol = base.shape
    cp.score_indel_matrix(base)
    sc_s, S = timed(lambda: cp.score_indel_matrix(base), 5)
    rng = np.random.default_rng(11)
    cols = rng.integers(0, ncol, (256, ncol)).astype(np.uint32)
    reps = [np.ascontiguousarray(base[:, c]) for c in cols]
    cp.join_indel(reps[0])
    plain_s, e_plain = timed(lambda: cp.join_indel(reps[0]), 3)
    batch_s, e_batch = timed(lambda: cp.join_indel_batch(reps), 1)
    cp.join_indel_bootstrap(base, cols[:4])
    boot_s, e_boot = timed(lambda: cp.join_indel_bootstrap(base, cols), 3)
    assert np.array_equal(e_batch[0], e_plain[0])
    assert all(np.array_equal(a, b) for a, b in zip(e_boot, e_batch)), "bootstrap / batch mismatch"
    out = {"cells": n, "sites": ncol,
           "score_matrix": {"ms": sc_s * 1e3, "cell_pair_scores_per_s": n * n / sc_s},
           "plain_jobs_per_s": 1 / plain_s, "batched_jobs_per_s": len(reps) / batch_s, "batched_jobs": len(reps),
           "speedup": plain_s * len(reps) / batch_s,
           "bootstrap_jobs_per_s": len(reps) / boot_s, "bootstrap_speedup": plain_s * len(reps) / boot_s,
           "bootstrap_form": "the base matrix and 256 x 200 column indices in, replicates built on the device"}
    if with_cpu:
        from concurrent.futures import ThreadPoolExecutor

        from oracle import oracle

        def weights(m):
            return np.ones((max(1, int(m.max())), m.shape[1]), np.float32)
        t = time.perf_counter()
        S2 = oracle.score_indel_matrix(base, weights(base))
        sc_cpu = time.perf_counter() - t
        assert np.array_equal(S2.astype(np.int32), S), "CPU/GPU indel score mismatch"
        k1 = 0
        t = time.perf_counter()
        while time.perf_counter() - t < 4.0 and k1 < len(reps):
            e2, _, _ = oracle.join_indel(reps[k1], weights(reps[k1]))
            assert np.array_equal(e2, e_batch[k1]), "CPU/GPU indel join mismatch"
            k1 += 1
        one = k1 / (time.perf_counter() - t)
        th = cpu_threads()
        kk = min(len(reps), 4 * th)
        t = time.perf_counter()
        with ThreadPoolExecutor(th) as ex:
            res = list(ex.map(lambda m: oracle.join_indel(m, weights(m))[0], reps[:kk]))
        allc = kk / (time.perf_counter() - t)
        assert all(np.array_equal(res[k], e_batch[k]) for k in range(kk)), "CPU/GPU indel join mismatch"
        out["cpu_baseline"] = {"value": allc, "unit": "tree-build jobs/s", "cores": th, "kind": "port",
                               "value_1_thread": one, "host": host_info(),
                               "score_matrix_ms_1_thread": sc_cpu * 1e3,
                               "sample": f"{kk} bootstrap joins on {th} threads, {k1} on 1 thread "
                                         f"(greedyJoinEdge2 restated), edges equal the GPU's"}
    return out


def dna_plain_vs_batched(cp, dna, dq, bound: float, batched: dict) -> dict:
    """The DNA configuration as the executables would run it: VALU tiles on
    the 2-bit + mask form, one tryRadius launch per radius with the rule on
    the host, then the radiusNeighbor lists (plain), against the headline's
    autotuned device rounds (batched); the same radius and list length."""
    os.environ["CASPHY_DNA_FORM"] = "valu"
    try:
        cells = cp.Cells(dna, "dna")
        dna_search_host(cells, dq, dna.shape[0], bound)
        el, (r, total, tries) = timed(lambda: dna_search_host(cells, dq, dna.shape[0], bound), 1)
        tile_ms, _ = cells.last_ms()
        cells.close()
    finally:
        os.environ.pop("CASPHY_DNA_FORM", None)
    assert r == batched["radius"] and total == batched["neighbour_ids"], "plain/batched DNA search mismatch"
    n = dna.shape[0]
    pairs = tries * R * n + n * (n + 1) // 2
    return {"plain": {"form": "VALU 2-bit + mask tiles, one tryRadius launch per radius", "tryradius_launches": tries,
                      "ms_per_search": el * 1e3, "mask_tiles_ms": tile_ms, "cell_pair_scores_per_s": pairs / el},
            "batched_autotuned": {"form": f"device rounds of 64 radii, {batched['form']}",
                                  "ms_per_search": batched["ms_per_step"],
                                  "mask_tiles_ms": batched["mask_tiles_ms"]},
            "speedup": el * 1e3 / batched["ms_per_step"]}


def extras(cp, synth, dna, dq, headline: dict, with_cpu: bool = True) -> dict:
    """The other configurations, plain vs batched, each a device-resident or
    host-API pass as a user would run it; every batched result is compared
    with its plain counterpart."""
    ex = {}
    bound = AUTOSPLIT_BOUND - np.log10(dna.shape[0])
    ex["dna_100000x1000_plain_vs_batched"] = dna_plain_vs_batched(cp, dna, dq, bound, headline)
    print("# DNA plain search done", file=sys.stderr, flush=True)
    ex["dna_vote_root_10000_tips"] = vote_root_step(cp, synth, dna, with_cpu)
    ex["dna_fasta_encode_100000x1000"] = fasta_step(cp, dna, with_cpu)
    print("# DNA format edge done", file=sys.stderr, flush=True)
    ex["base_editor_10000x4096"] = base_editor_round(cp, synth, with_cpu)
    print("# base-editor round done", file=sys.stderr, flush=True)
    ex["indel_1024x200"] = indel_config(cp, synth, with_cpu)
    print("# indel configuration done", file=sys.stderr, flush=True)
    ex.update(tbe_step(cp, synth, with_cpu))
    print("# TBE done", file=sys.stderr, flush=True)
    ex["scan_100000x81920_bit"] = scan_form(cp, synth)
    print("# scan form done", file=sys.stderr, flush=True)
    ex["serving_mix"] = serving_mix(cp, synth)
    print("# serving done", file=sys.stderr, flush=True)
    return ex


TBE_TIPS, TBE_REPS = 65536, 1000


def tbe_step(cp, synth, with_cpu: bool = True, nrep: int = TBE_REPS) -> dict:
    """The batched transfer configuration: a 65,536-tip Yule reference tree
    (seed 4) against 1,000 bootstrap trees, each ceil(5% of the internal
    edges) random SPR moves away from it, as R edge matrices.  Batched: every
    replicate through casphy_transfer_trees in one call (replicates streamed
    through the device in chunks, tree walks, no bitsets).  Plain: one
    replicate per call on the bitset pair tiles (the per-replicate
    TransferDisStream form), on a few replicates, whose minima must equal the
    batched ones.  CPU: the oracle's brute force over bitset pairs on sampled
    reference rows, 1 and all threads."""
    ex = {}
    rng = np.random.default_rng(4)
    ntip = TBE_TIPS
    parent, leaves = synth.yule_parent(ntip, rng)
    ref_e, ref_n, _ = synth.ape_edges(parent, leaves)
    moves = int(np.ceil(0.05 * (ntip - 3)))
    t = time.perf_counter()
    reps = synth.spr_edges_batch(parent, leaves, moves, nrep, 4, threads=cpu_threads())
    gen_s = time.perf_counter() - t
    edges, nnodes = [e for e, _ in reps], [n for _, n in reps]
    cp.transfer_trees(ref_e, ref_n, edges[:2], nnodes[:2], ntip)  # first use: code load, pools
    batch_s, (ymin, ysize, xmin, xsize) = timed(lambda: cp.transfer_trees(ref_e, ref_n, edges, nnodes, ntip), 1)
    print(f"# TBE batched: {nrep} replicates in {batch_s:.2f} s", file=sys.stderr, flush=True)
    ny = ymin.shape[1]
    pairs = sum(ny * (n - 1) for n in nnodes)
    # plain: one replicate per call on the bitset tiles
    nplain = 3
    os.environ["CASPHY_TBE"] = "bitsets"
    try:
        cp.transfer_trees(ref_e, ref_n, edges[:1], nnodes[:1], ntip)
        plain_s, pl = timed(lambda: [cp.transfer_trees(ref_e, ref_n, [edges[b]], [nnodes[b]], ntip)
                                     for b in range(nplain)], 1)
    finally:
        os.environ.pop("CASPHY_TBE", None)
    for b in range(nplain):
        assert np.array_equal(pl[b][0][0], ymin[b]) and np.array_equal(pl[b][2][0], xmin[b]), "plain/batched TBE"
    out = {"tips": ntip, "replicates": nrep, "spr_moves": moves, "clade_rows": ny,
           "batched_jobs_per_s": nrep / batch_s, "batched_s": batch_s,
           "plain_jobs_per_s": nplain / plain_s, "plain_form": "bitset pair tiles (fp4 matrix cores), one call per "
                                                                "replicate",
           "speedup": (plain_s / nplain) / (batch_s / nrep), "clade_pairs_per_s": pairs / batch_s,
           "replicates_generated_s": gen_s,
           "mean_ymin_zero": float((ymin == 0).mean())}
    if with_cpu:
        from oracle import oracle
        Y, _ = oracle.clade_bitsets(ref_e, ref_n, ntip)
        X, _ = oracle.clade_bitsets(edges[0], nnodes[0], ntip)
        rows = np.random.default_rng(41).permutation(ny)

        def run(i0, i1):
            sel = rows[i0:i1]
            return sel, oracle.transfer(Y[sel], X, ntip)

        def check(i0, i1, res):
            sel, (y2, ys2, _, _) = res
            assert np.array_equal(y2, ymin[0][sel]) and np.array_equal(ys2, ysize[sel]), "CPU/GPU TBE mismatch"
        res = {}
        for th in (1, cpu_threads()):
            done, el = cpu_rows_timed(run, ny, 2, th, 5.0, check)
            res[th] = (done * X.shape[0] / el, done, el)
        th = cpu_threads()
        out["cpu_baseline"] = {"value": res[th][0] / (ny * X.shape[0]), "unit": "TBE jobs/s", "cores": th,
                               "kind": "port", "value_1_thread": res[1][0] / (ny * X.shape[0]),
                               "clade_pairs_per_s": res[th][0], "host": host_info(),
                               "sample": f"{res[th][1]} reference rows x {X.shape[0]} replicate rows on {th} threads "
                                         f"({res[th][2]:.1f} s), {res[1][1]} rows on 1 thread ({res[1][2]:.1f} s), "
                                         f"brute force over 2,048-word bitsets; every row equals the GPU's"}
    ex["tbe_65536_tips_1000_replicates"] = out
    del reps, edges, ymin, xmin
    # from Newick text: read (tips matched by label), trees in, transfer minima
    # and TBE support in one pass per job set
    nrng = np.random.default_rng(15)
    ref_text = synth.newick(parent, leaves, rng=nrng)
    rep_text = "\n".join(synth.newick(synth.spr(parent, leaves, moves, nrng), leaves, rng=nrng) for _ in range(4))

    def from_newick():
        rt = cp.read_newick(ref_text)[0]
        return cp.transfer_support(rt, cp.read_newick(rep_text, tip_labels=rt[3]))
    from_newick()
    nwk_s, (sup, _, _) = timed(from_newick, 1)
    ex["tbe_65536_tips_1000_replicates"]["from_newick_jobs_per_s"] = 4 / nwk_s
    ex["tbe_65536_tips_1000_replicates"]["from_newick_mean_support"] = float(np.nanmean(sup))
    # small bootstrap trees: 4,096 tips x 64 replicates, bitsets in (plain per
    # replicate vs one batched launch) and trees in
    ntip = 4096
    parent, leaves = synth.yule_parent(ntip, rng)
    ref = synth.clade_bitsets(parent, leaves, ntip)
    moves = int(np.ceil(0.05 * (ntip - 3)))
    allb, counts = replicate_block(synth, parent, leaves, ntip, moves, 64, rng)
    boots = np.split(allb, np.cumsum(counts)[:-1])
    cp.transfer_dis_batch(ref, allb, ntip, counts)
    cp.transfer_dis(ref, boots[0], ntip)
    plain_s, pl = timed(lambda: [cp.transfer_dis(ref, b, ntip) for b in boots], 1)
    batch_s, bt = timed(lambda: cp.transfer_dis_batch(ref, allb, ntip, counts), 1)
    for k in range(len(boots)):
        assert np.array_equal(bt[0][k], pl[k][0]) and np.array_equal(bt[2][k], pl[k][2])
    ex["tbe_4096_tips_64_replicates"] = {"plain_jobs_per_s": len(boots) / plain_s,
                                         "batched_jobs_per_s": len(boots) / batch_s,
                                         "speedup": plain_s / batch_s}
    return ex


TIPMERGE_RADII = np.arange(20, 340, 20, dtype=np.int32)
TIPMERGE_MEAN = 300   # autoSplit bound: mean neighbour count at most 300
TIPMERGE_P = 0.8      # p_shared default, R/CasPhy_functions.R:313,333


def tip_merge_step(cp, cells, q, with_cpu: bool) -> dict:
    """The split step of the autoSplit pipeline on resident cells: a device
    autoSplit round (radii 20..320, mean count bound 300) picks the radius and
    keeps its lists on the device, then tipMerge clusters them
    (R/CasPhy_functions.R:2024-2026 runs radiusNeighbor then tipMerge on files).
    The CPU leg runs the oracle's restatement of cu/tipMerge.cu on the same
    lists, single-threaded, and checks the labels."""
    r, total = cells.autosplit(TIPMERGE_RADII, q, float(np.log10(TIPMERGE_MEAN)))
    counts, idx = cells.fetch()
    cells.tip_merge(TIPMERGE_P)  # first call of the pair-tile shape class: autotuner
    dev_s, member = timed(lambda: cells.tip_merge(TIPMERGE_P), 5)
    file_s, m2 = timed(lambda: cp.tip_merge(counts, idx, TIPMERGE_P), 2)
    assert np.array_equal(member, m2)
    forms = {}
    for form in ("dense", "sparse"):
        os.environ["CASPHY_TIPMERGE"] = form
        cells.tip_merge(TIPMERGE_P)
        forms[form + "_ms"] = timed(lambda: cells.tip_merge(TIPMERGE_P), 3)[0] * 1e3
    os.environ.pop("CASPHY_TIPMERGE")
    n = counts.size
    out = {"radius": int(r), "neighbour_ids": int(total), "mean_list": float(counts.mean()),
           "clusters": int(np.unique(member).size), "largest_cluster": int(np.bincount(member).max()),
           "device_ms": dev_s * 1e3, "file_form_ms": file_s * 1e3, **forms,
           "tips_per_s": n / dev_s}
    if with_cpu:
        from oracle import oracle
        cpu_s, m3 = timed(lambda: oracle.tip_merge(counts, idx, TIPMERGE_P), 1)
        assert np.array_equal(m3, member), "CPU/GPU tipMerge mismatch"
        out["cpu_baseline"] = {"value": n / cpu_s, "unit": "tips/s", "cores": 1, "kind": "port",
                               "sample": f"the full {n}-tip workload, {cpu_s:.2f} s"}
        out["speedup_vs_cpu"] = cpu_s / dev_s
    return out, member


VOTE_TIPS = 10000      # voteRoot2: tree on 10,000 sequences of the DNA alignment
VOTE_SAMPLES = 16384   # outgroup samples, R's floor (R/DNA_functions.R:108)


def vote_root_step(cp, synth, dna, with_cpu: bool) -> dict:
    """voteRoot2 on the 100,000 x 1,000 DNA alignment: a Yule tree on its
    first 10,000 sequences with Fitch states, 16,384 further sequences as
    outgroup samples (R/DNA_functions.R:107-133); a host-API pass (upload,
    features, matrix-core counts with the argmax epilogue, votes back)."""
    st, edge, mat, idx = synth.vote_inputs(3, VOTE_TIPS, VOTE_SAMPLES, 1000, mat=dna)
    cp.vote_root(st, edge, mat, idx)  # first call of the shape class: autotuner
    dev_s, (root, nsh) = timed(lambda: cp.vote_root(st, edge, mat, idx), 3)
    nnode, nbin = st.shape
    pairs = nnode * VOTE_SAMPLES
    res = {"nodes": int(nnode), "samples": VOTE_SAMPLES, "positions": 4 * int(nbin), "ms": dev_s * 1e3,
           "node_sample_pairs_per_s": pairs / dev_s, "position_pairs_per_s": pairs * 4 * nbin / dev_s,
           "distinct_roots": int(np.unique(root).size)}
    if with_cpu:
        from oracle import oracle
        k = 1024
        cpu_s, (r2, n2) = timed(lambda: oracle.vote_root(st, edge, mat, idx[:k]), 1)
        assert np.array_equal(r2, root[:k]) and np.array_equal(n2, nsh[:k]), "CPU/GPU voteRoot2 mismatch"
        res["cpu_baseline"] = {"value": nnode * k / cpu_s, "unit": "node-sample pairs/s", "cores": 1,
                               "kind": "port", "sample": f"{k} of the {VOTE_SAMPLES} samples, {cpu_s:.2f} s"}
    return res


def fasta_step(cp, dna, with_cpu: bool) -> dict:
    """fasta2bin's column filter and encoding of the 100,000 x 1,000 DNA
    alignment as characters (host-API pass: characters up, codes back)."""
    b = dna.view(np.uint8).reshape(dna.shape[0], -1)[:, :1000]
    chars = np.frombuffer(b"?ATTCCCCGGGGGGGG-", np.uint8)
    seqs = np.ascontiguousarray(chars[b])
    cp.fasta_encode(seqs[:64])
    dev_s, (var, valid, out) = timed(lambda: cp.fasta_encode(seqs), 3)
    res = {"sequences": int(seqs.shape[0]), "columns": int(seqs.shape[1]), "kept": int(var.sum()),
           "ms": dev_s * 1e3, "characters_per_s": seqs.size / dev_s}
    if with_cpu:
        from oracle import oracle
        cpu_s, w = timed(lambda: oracle.fasta_encode(seqs), 1)
        assert all(np.array_equal(x, y) for x, y in zip((var, valid, out), w)), "CPU/GPU fasta2bin mismatch"
        res["cpu_baseline"] = {"value": seqs.size / cpu_s, "unit": "characters/s", "cores": 1, "kind": "port",
                               "sample": f"the full alignment, {cpu_s:.2f} s"}
        if os.path.exists(REF_FASTA2BIN):
            res["files_end_to_end"] = fasta_files(seqs)
    return res


REF_FASTA2BIN = os.path.join(ROOT, "oracle", "_ref", "fasta2bin")


def fasta_files(seqs: np.ndarray) -> dict:
    """FASTA file in, the three output files out: the drop-in fasta2bin
    executable against the reference's own fasta2bin (built from
    cu/fasta2bin.cu into oracle/_ref/) on the same file, outputs compared
    byte for byte."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="casphy_fasta_")
    try:
        src = os.path.join(d, "in.fa")
        with open(src, "wb") as f:
            for i in range(seqs.shape[0]):
                f.write(b">seq%d\n" % i)
                f.write(seqs[i].tobytes() + b"\n")
        times = {}
        for who, exe in (("drop_in", os.path.join(ROOT, "casphy_amd", "bin", "fasta2bin")), ("reference", REF_FASTA2BIN)):
            t0 = time.perf_counter()
            subprocess.run([exe, src, os.path.join(d, who)], check=True, stdout=subprocess.DEVNULL)
            times[who] = time.perf_counter() - t0
        same = all(open(os.path.join(d, "drop_in" + suf), "rb").read() == open(os.path.join(d, "reference" + suf), "rb").read()
                   for suf in (".bin", "_names.txt", "_positions.txt"))
        assert same, "drop-in / reference fasta2bin outputs differ"
        mb = os.path.getsize(src) / 1e6
        return {"fasta_mb": mb, "drop_in_s": times["drop_in"], "reference_s": times["reference"],
                "speedup": times["reference"] / times["drop_in"], "identical_files": same,
                "cpu_baseline": {"value": seqs.size / times["reference"], "unit": "characters/s", "cores": 1,
                                 "kind": "reference", "sample": "the full alignment as a FASTA file, wall clock "
                                 "of the reference executable built from cu/fasta2bin.cu"}}
    finally:
        shutil.rmtree(d, ignore_errors=True)


NODEKEEP_OUT = 1000    # outgroup rows (cells of the matrix)
NODEKEEP_CUTOFF = 1


def node_keep_step(cp, mat, with_cpu: bool) -> dict:
    """nodeKeepBit after the 10,000-cell join (R/CasPhy_functions.R:2204,
    2275-2286): the joined tree's .edge2 / .nodeState2 and 1,000 outgroup
    rows; the CPU leg runs the oracle's restatement on the same input."""
    eo, ns2 = cp.join_edge_bit_reverse(mat)
    out = mat[np.random.default_rng(5).choice(mat.shape[0], NODEKEEP_OUT, replace=False)]
    cp.node_keep(eo, ns2, out, NODEKEEP_CUTOFF)
    dev_s, kept = timed(lambda: cp.node_keep(eo, ns2, out, NODEKEEP_CUTOFF), 3)
    nnode = ns2.shape[0]
    res = {"internal_nodes": int(nnode), "outgroups": NODEKEEP_OUT, "cutoff": NODEKEEP_CUTOFF,
           "kept": int(kept.size), "ms": dev_s * 1e3,
           "node_outgroup_pairs_per_s": nnode * NODEKEEP_OUT / dev_s}
    if with_cpu:
        from oracle import oracle
        cpu_s, k2 = timed(lambda: oracle.node_keep(eo, ns2, out, NODEKEEP_CUTOFF), 1)
        assert np.array_equal(k2, kept), "CPU/GPU nodeKeep mismatch"
        res["cpu_baseline"] = {"value": 1 / cpu_s, "unit": "trees/s", "cores": 1, "kind": "port",
                               "sample": f"the full tree, {cpu_s * 1e3:.1f} ms"}
        res["speedup_vs_cpu"] = cpu_s / dev_s
    return res


def replicate_block(synth, parent, leaves, ntip, moves, nrep, rng):
    """Clade bitsets of nrep SPR-perturbed copies of a tree, back to back in
    one array (the layout a caller holding all replicates passes in)."""
    sets = [synth.clade_bitsets(synth.spr(parent, leaves, moves, rng), leaves, ntip) for _ in range(nrep)]
    counts = [b.shape[0] for b in sets]
    allb = np.empty((sum(counts), sets[0].shape[1]), np.uint32)
    o = 0
    for b in sets:
        allb[o:o + b.shape[0]] = b
        o += b.shape[0]
    return allb, counts


def scan_form(cp, synth, n: int = 100000, W: int = 2560, R: int = 4) -> dict:
    """Scan form of the query x cell scores: tryRadius for 4 query rows over
    100,000 cells x 81,920 sites (1.02 GB, four times the 256 MB MALL, so each
    pass streams from HBM), radii 1200..1350.  Bits are i.i.d. with density
    1/8 (the pass is bandwidth bound, the values do not matter); the counts
    are compared with the tile path on the same resident cells."""
    rng = np.random.default_rng(6)
    big = np.empty((n, W), np.uint32)
    for i in range(0, n, 10000):
        shape = (min(10000, n - i), W)
        blk = rng.integers(0, 2 ** 32, size=shape, dtype=np.uint32)
        blk &= rng.integers(0, 2 ** 32, size=shape, dtype=np.uint32)
        blk &= rng.integers(0, 2 ** 32, size=shape, dtype=np.uint32)
        big[i:i + shape[0]] = blk
    cells = cp.Cells(big, "bit", wordmajor=False)
    del big
    q = synth.query_rows(7, n, R)
    radii = (1200 + 10 * np.arange(16)).astype(np.int32)
    cells.try_radii(radii, q)
    ms = []
    for _ in range(10):
        counts = cells.try_radii(radii, q)
        ms.append(cells.last_ms()[0])
    os.environ["CASPHY_SCAN"] = "0"
    tiles = cells.try_radii(radii, q)
    os.environ.pop("CASPHY_SCAN")
    assert np.array_equal(counts, tiles)
    cells.close()
    pass_ms = float(np.median(ms))
    nbytes = n * W * 4
    return {"queries": R, "radii": 16, "bytes_per_pass": nbytes, "pass_ms": pass_ms,
            "GB_per_s": nbytes / (pass_ms / 1e3) / 1e9, "hbm_frac": nbytes / (pass_ms / 1e3) / 8e12,
            "kernel": "k_scan_bit", "bit_exact_vs_tiles": True}


def serving_mix(cp, synth, njobs: int = 60, seed: int = 5, max_cells: int = 30000) -> dict:
    """The serving configuration: tree builds (binary 4,096 sites and indel
    200 sites, alternating) over rows of resident matrices and transfer
    distances from trees (every fourth job: a Yule tree of the job's size
    against 3 replicates 5% SPR moves away), sizes log-uniform in
    [1k, max_cells], through the request loop in two forms:
      * plain: one worker, one job per launch, every job's input copied from
        the host (what one process per job does);
      * served: 4 workers whose launches overlap, small joins batched,
        matrices resident (rows gathered on the device), no head-of-line
        blocking.
    Each form runs once with every job submitted at once (capacity); then
    Poisson arrivals at 80% of each form's capacity give its p50 / p99
    latency.  Every result is compared with the plain capacity run."""
    import threading
    bmat = synth.binary_matrix(seed=seed, n=max_cells, nsite=4096, p=0.005, missing=0.05)
    depth = int(np.ceil(np.log2(max_cells)))
    imat = synth.indel_matrix(seed=seed, depth=depth, nsite=200)
    jobs = synth.serving_jobs(seed, njobs, 1000, max_cells, bmat.shape[0], imat.shape[0])
    print(f"# serving inputs ready ({njobs} jobs)", file=sys.stderr, flush=True)

    def run(served, rate):
        os.environ["CASPHY_SERVE_BATCH"] = "1" if served else "0"
        srv = cp.Server(hbm_budget_bytes=64 << 30, workers=4 if served else 1)
        mids = {"bit": srv.put_matrix(bmat, "bit"), "indel": srv.put_matrix(imat, "indel")} if served else None
        gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate, len(jobs)) if rate else np.zeros(len(jobs))
        ids = [None] * len(jobs)

        def feeder():
            t0 = time.perf_counter()
            due = 0.0
            for k, (kind, data) in enumerate(jobs):
                due += gaps[k]
                d = t0 + due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                if kind == "trees":
                    ids[k] = srv.submit_transfer_trees(*data)
                elif served:
                    ids[k] = srv.submit_join_rows(mids[kind], data)
                else:
                    ids[k] = srv.submit_join((bmat if kind == "bit" else imat)[data], kind)
        t0 = time.perf_counter()
        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        outs = [srv.wait(ids[k])[0] for k in range(len(jobs))]
        el = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        os.environ.pop("CASPHY_SERVE_BATCH", None)
        print(f"# serving run: served={served} rate={rate:.2f}/s {el:.1f} s", file=sys.stderr, flush=True)
        return outs, el, st

    def same(xs, ys):
        for a, b in zip(xs, ys):
            if isinstance(a, tuple):
                for x, y in zip(a, b):
                    if isinstance(x, list):
                        assert all(np.array_equal(u, v) for u, v in zip(x, y))
                    else:
                        assert np.array_equal(x, y)
            else:
                assert np.array_equal(a, b)

    outs_p, el_p, st_pb = run(False, 0)
    outs_s, el_s, st_sb = run(True, 0)
    same(outs_p, outs_s)
    rate_p, rate_s = 0.8 * len(jobs) / el_p, 0.8 * len(jobs) / el_s
    outs_p2, _, st_p = run(False, rate_p)
    outs_s2, _, st_s = run(True, rate_s)
    same(outs_p, outs_p2)
    same(outs_p, outs_s2)
    sizes = np.array([d[4] if k == "trees" else d.size for k, d in jobs])

    def form(el, burst, poisson, rate):
        return {"capacity_jobs_per_s": len(jobs) / el, "capacity_launches": burst["launches"],
                "poisson_rate_per_s": rate, "poisson_jobs_per_s": poisson["jobs_per_s"],
                "p50_ms": poisson["p50_ms"], "p99_ms": poisson["p99_ms"], "launches": poisson["launches"]}
    return {"jobs": len(jobs), "cells": f"log-uniform 1k..{max_cells // 1000}k",
            "mean_cells": float(sizes.mean()), "largest": int(sizes.max()),
            "tree_builds": sum(k != "trees" for k, _ in jobs), "tbe_jobs": sum(k == "trees" for k, _ in jobs),
            "plain": form(el_p, st_pb, st_p, rate_p), "served": form(el_s, st_sb, st_s, rate_s),
            "capacity_speedup": el_p / el_s, "bit_exact_between_forms": True}


def serving_poisson_segment(cp, synth, njobs: int, seed: int, max_cells: int, rate: float) -> dict:
    """One segment of the serving configuration at its stated size: `njobs`
    jobs of the mix above (sizes log-uniform in [1k, max_cells], its own seed)
    through the served form only, Poisson arrivals at `rate` jobs/s.  Returns
    every job's latency, so that segments of different seeds merge into one
    distribution; every 25th job is run again through a one-worker plain
    server and compared bit for bit."""
    import threading
    bmat = synth.binary_matrix(seed=seed, n=max_cells, nsite=4096, p=0.005, missing=0.05)
    depth = int(np.ceil(np.log2(max_cells)))
    imat = synth.indel_matrix(seed=seed, depth=depth, nsite=200)
    jobs = synth.serving_jobs(seed, njobs, 1000, max_cells, bmat.shape[0], imat.shape[0])
    print(f"# serving segment inputs ready ({njobs} jobs, seed {seed})", file=sys.stderr, flush=True)
    srv = cp.Server(hbm_budget_bytes=64 << 30, workers=4)
    mids = {"bit": srv.put_matrix(bmat, "bit"), "indel": srv.put_matrix(imat, "indel")}
    gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate, len(jobs))
    ids = [None] * len(jobs)

    def feeder():
        t0 = time.perf_counter()
        due = 0.0
        for k, (kind, data) in enumerate(jobs):
            due += gaps[k]
            d = t0 + due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            ids[k] = (srv.submit_transfer_trees(*data) if kind == "trees"
                      else srv.submit_join_rows(mids[kind], data))
    t0 = time.perf_counter()
    th = threading.Thread(target=feeder)
    th.start()
    done, lat = [], []
    for k in range(len(jobs)):
        while ids[k] is None:
            time.sleep(0.001)
        out, ms = srv.wait(ids[k])
        done.append(out)
        lat.append(ms)
        if k % 50 == 49:
            print(f"# segment {seed}: {k + 1} jobs done, {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    th.join()
    el = time.perf_counter() - t0
    srv.close()
    plain = cp.Server(hbm_budget_bytes=64 << 30, workers=1)
    checked = 0
    for k in range(0, len(jobs), 25):
        kind, data = jobs[k]
        j = (plain.submit_transfer_trees(*data) if kind == "trees"
             else plain.submit_join((bmat if kind == "bit" else imat)[data], kind))
        ref = plain.wait(j)[0]
        if isinstance(ref, tuple):
            for x, y in zip(ref, done[k]):
                if isinstance(x, list):
                    assert all(np.array_equal(u, v) for u, v in zip(x, y))
                else:
                    assert np.array_equal(x, y)
        else:
            assert np.array_equal(ref, done[k])
        checked += 1
    plain.close()
    sizes = np.array([d[4] if k == "trees" else d.size for k, d in jobs])
    lat = np.array(lat)
    return {"jobs": len(jobs), "seed": seed, "cells": f"log-uniform 1k..{max_cells // 1000}k",
            "mean_cells": float(sizes.mean()), "largest": int(sizes.max()),
            "tree_builds": sum(k != "trees" for k, _ in jobs), "tbe_jobs": sum(k == "trees" for k, _ in jobs),
            "poisson_rate_per_s": rate, "elapsed_s": el, "jobs_per_s": len(jobs) / el,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "checked_against_plain": checked, "latency_ms": [round(float(x), 1) for x in lat],
            "cells_per_job": [int(x) for x in sizes]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other configurations")
    ap.add_argument("--calibrate", action="store_true", help="also run a 256 MiB device copy (counter scaling)")
    ap.add_argument("--serving-jobs", type=int, default=60, help="jobs of the serving mix")
    ap.add_argument("--serving-max", type=int, default=30000, help="largest job of the serving mix (cells)")
    ap.add_argument("--serving-seed", type=int, default=11, help="seed of a serving segment")
    ap.add_argument("--serving-rate", type=float, default=1.4, help="Poisson arrivals of a serving segment (jobs/s)")
    ap.add_argument("--only", choices=["tipmerge", "nodekeep", "voteroot", "fasta", "tbe", "base", "indel", "scan",
                                       "dnavalu", "ceilings", "serving", "serving-segment"],
                    help="run only one of the other configurations (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import casphy_amd as cp
    from casphy_amd import synth

    if args.calibrate:
        calibration_copy()
    with_cpu = not args.no_cpu
    if args.only in ("tipmerge", "nodekeep"):
        mat = synth.binary_matrix(seed=2, n=N, nsite=NSITE, p=0.005, missing=0.05)
        if args.only == "nodekeep":
            print(json.dumps(node_keep_step(cp, mat, with_cpu)))
        else:
            cells = cp.Cells(np.ascontiguousarray(mat.T), "bit", wordmajor=True)
            out, member = tip_merge_step(cp, cells, synth.query_rows(2, N, R), with_cpu)
            out["split_layer"] = split_layer_step(cp, mat, cells, member, out["radius"])
            print(json.dumps(out))
        return
    if args.only in ("voteroot", "fasta"):
        dna = synth.dna_matrix(seed=3, n=N3, ncol=NCOL3)
        step = vote_root_step(cp, synth, dna, with_cpu) if args.only == "voteroot" else fasta_step(cp, dna, with_cpu)
        print(json.dumps(step))
        return
    if args.only == "tbe":
        print(json.dumps(tbe_step(cp, synth, with_cpu)))
        return
    if args.only == "base":
        print(json.dumps(base_editor_round(cp, synth, with_cpu, steps=args.steps, warmup=args.warmup)))
        return
    if args.only == "indel":
        print(json.dumps(indel_config(cp, synth, with_cpu)))
        return
    if args.only == "scan":
        print(json.dumps(scan_form(cp, synth)))
        return
    if args.only == "ceilings":
        print(json.dumps(ceilings()))
        return
    if args.only == "serving":
        print(json.dumps(serving_mix(cp, synth, njobs=args.serving_jobs, max_cells=args.serving_max)))
        return
    if args.only == "serving-segment":
        print(json.dumps(serving_poisson_segment(cp, synth, args.serving_jobs, args.serving_seed, args.serving_max,
                                                 args.serving_rate)))
        return
    if args.only == "dnavalu":
        # the VALU 2-bit + mask tiles of the DNA configuration: one tryRadius
        # and one radius pass at the headline's radius (counter profiles)
        print(json.dumps(ceilings()))
        dna = synth.dna_matrix(seed=3, n=N3, ncol=NCOL3)
        os.environ["CASPHY_DNA_FORM"] = "valu"
        cells = cp.Cells(dna, "dna")
        q = synth.query_rows(3, N3, R)
        counts = cells.try_radii(np.array([640], np.int32), q)
        total = cells.radius(640)
        print(json.dumps({"tryradius_mean": float(counts.mean()), "neighbour_ids": int(total),
                          "mask_tiles_ms": cells.last_ms()[0], "pairs": N3 * (N3 + 1) // 2}))
        cells.close()
        return

    t = time.perf_counter()
    dna = synth.dna_matrix(seed=3, n=N3, ncol=NCOL3)
    q = synth.query_rows(3, N3, R)
    print(f"# input generated in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    cells = cp.Cells(dna, "dna")  # upload, one-hot / 2-bit forms, form tuner
    form = [e for e in cp.mfma_tuning() if "dna_form" in e]
    form = form[-1] if form else {"dna_form": os.environ.get("CASPHY_DNA_FORM", "mfma")}
    bound = AUTOSPLIT_BOUND - np.log10(N3)

    # the device rounds equal the host-driven loop (tryRadius per radius, host
    # rule, radiusNeighbor); the first search also runs the tile autotuner
    r_host, t_host, tries = dna_search_host(cells, q, N3, bound)
    c_host, i_host = cells.fetch()
    r_dev, t_dev, rounds = dna_search(cells, q, N3, bound)
    c_dev, i_dev = cells.fetch()
    assert r_dev == r_host and t_dev == t_host, "device rounds / host loop mismatch"
    assert np.array_equal(c_dev, c_host) and np.array_equal(i_dev, i_host), "device rounds / host loop lists differ"
    del c_host, i_host
    print(f"# checked: radius {r_dev}, {t_dev} ids, {rounds} rounds", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        assert dna_search(cells, q, N3, bound)[:2] == (r_dev, t_dev)
    cp._lib.check(cp._lib.load().casphy_synchronize())
    tile_ms, round_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r, total, nr = dna_search(cells, q, N3, bound)
        tm, rm = cells.last_ms()  # the last round: the one that builds the lists
        tile_ms.append(tm)
        round_ms.append(rm)
    cp._lib.check(cp._lib.load().casphy_synchronize())
    el = time.perf_counter() - t0
    assert (r, total, nr) == (r_dev, t_dev, rounds)
    ms = 1000 * el / args.steps
    pairs_tile = N3 * (N3 + 1) // 2
    pairs = rounds * R * N3 + pairs_tile
    value = pairs * args.steps / el
    valu = form.get("dna_form") == "valu"
    # algorithmic work of the mask tiles: 4 one-hot sites (A, C, G, T) per
    # position, one multiply-add per site = 8,000 ops per pair (matrix cores);
    # 32 positions per 3 words x 6 lane-ops = 192 per pair (VALU 2-bit form)
    roof = mask_roofline(cp, float(np.mean(tile_ms)), pairs_tile, 2 * 4 * NCOL3, (17, 4096), TRAFFIC_TABLE_DNA,
                         valu=valu, valu_ops_per_pair=6 * ((NCOL3 + 31) // 32))
    headline = {"radius": r_dev, "neighbour_ids": int(t_dev), "ms_per_step": ms,
                "mask_tiles_ms": float(np.mean(tile_ms)), "form": "VALU 2-bit + mask tiles" if valu else
                f"one-hot sites on the {roof.get('variant')} matrix-core tiles"}
    out = {
        "metric": "cell-pair scores/s + tree-build & TBE jobs/s on 1 GPU, batched vs plain, bit-exact",
        "value": value, "unit": "cell-pair scores/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32" if valu else ("fp4" if roof["variant"].startswith("fp4") else "int8"),
        "data": "synthetic (Yule tree, Jukes-Cantor, 2% gaps, seed 3)",
        "config": {"workload": "DNA MSA 100,000 sequences x 1,000 columns: autoSplit radius search as R runs it "
                               "(tryRadiusDNA 1,024 queries at radii 100, 110, ..., 64 per device round) + "
                               "radiusNeighborDNA lists at the chosen radius",
                   "cells": N3, "columns": NCOL3, "queries": R, "radius": r_dev, "device_rounds": rounds,
                   "neighbour_ids": int(t_dev), "dna_form": form,
                   "last_round_ms": float(np.mean(round_ms))},
        "roofline": roof,
        "ceilings": ceilings(),
        "mfma_tuning": cp.mfma_tuning(),
    }
    if with_cpu:
        off = np.zeros(N3 + 1, np.int64)
        off[1:] = np.cumsum(c_dev, dtype=np.int64)
        out["cpu_baseline"] = cpu_radius_baseline("dna", dna, r_dev, c_dev, i_dev, off, budget_s=8.0, stride=97)
    del c_dev, i_dev
    cells.close()
    print("# headline done", file=sys.stderr, flush=True)
    if not args.no_extra:
        out["extra"] = extras(cp, synth, dna, q, headline, with_cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
