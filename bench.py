#!/usr/bin/env python3
"""Benchmark of the F4 linear-algebra step on one MI355X.

Workload (one step): the sparse pass -- reduce every to-reduce row of
one F4 matrix by its known pivots (gb's sparse_AB_CD_linear_algebra,
la_ff_32.c:1081-1179) -- on a seeded synthetic F4 matrix of 20000 rows x
30000 columns mod 65521: 12000 known pivots (staircase leads, lognormal
lengths 50-400, unit lead) and 8000 rows to reduce (~300 entries, 1 %
density).  Inputs are uploaded once; the timed region covers the whole pass
on the device (pivot layout, the reduction kernel, compaction of the
nonzero rows).

`value` is reduced rows per second of the blocked form (64-column groups
eliminated through precomputed inverses of the pivots' diagonal blocks);
the plain form (one pivot column at a time, ballot over the accumulator)
is timed beside it.  The JSON line also carries:
  * roofline: the reduction kernel's algorithmic bytes per launch (pivot
    entries streamed + row inputs + dense outputs) over its event-timed
    average duration, against 8 TB/s HBM;
  * cpu_baseline: the CPU oracle (the C restatement of the reference's
    routine, OpenMP) on the same matrix on this host;
  * extra: the full la_option 1 step (sparse pass + dense echelon form) on
    the same matrix in both forms, and the dense 4096 x 8192 block
    elimination in modular GOPS.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

P = 65521
NR, NC, NPIV = 20000, 30000, 12000
SEED = 2024
HBM_PEAK_GBS = 8000.0
# HBM bytes per k_sparse_blocked launch on this workload, from rocprofv3
# FETCH_SIZE (x2.00) and WRITE_SIZE (x1.00), each scaled by what the same
# counter reads for a plain 256 MiB device copy; the commands and raw counts
# are in profiles/k_sparse_blocked_f4_20000x30000_p65521.md
PMC_TRAFFIC_BYTES = 3.11e9 + 0.770e9


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def dense_ops(m: int, n: int, r: int) -> float:
    """Multiply-adds x 2 of a reduced echelon form of an m x n block of rank
    r: elimination 2mnr - r^2(m+n) + 2/3 r^3 plus back substitution
    r^2 (n - r)."""
    return 2.0 * m * n * r - r * r * (m + n) + (2.0 / 3.0) * r ** 3 + r * r * (n - r)


def cpu_baseline(m, nthreads: int):
    """The oracle's sparse pass (C + OpenMP) over the same matrix."""
    from oracle import oracle
    oracle.sparse_pass_work(m, nthreads=nthreads)  # warm-up (page-in)
    t0 = time.perf_counter()
    oracle.sparse_pass_work(m, nthreads=nthreads)
    t = time.perf_counter() - t0
    return t


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the extra workloads")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: the F4 step runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2

    import numpy as np
    from gb_amd import native, synth

    m = synth.f4_shaped(NR, NC, NPIV, P, seed=SEED)
    ctx = native.Context(0)
    d = ctx.upload(m)
    nrl = m.nr - len(np.unique(m.leads()))

    # algorithmic work of the reduction kernel (an instrumented, untimed run)
    nd, napp, nent = d.sparse_pass_work()
    lens = np.diff(m.rp)
    leads = m.leads()
    first = np.zeros(m.nr, bool)
    _, idx = np.unique(leads, return_index=True)
    first[idx] = True
    red_nnz = int(lens[~first].sum())
    bytes_per_launch = 8.0 * red_nnz + 8.0 * nent + 4.0 * m.ncr * nrl

    def timed(form, steps, warmup):
        for _ in range(warmup):
            d.sparse_pass(form)
        _sync()
        kernel_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            d.sparse_pass(form)
            kernel_ms.append(ctx.phase_ms(0))
        _sync()
        return (time.perf_counter() - t0) / steps, sum(kernel_ms) / len(kernel_ms)

    # headline: the blocked form (MFMA build of the library); the plain
    # ballot-chain form is timed beside it
    t_step, kavg = timed(native.FORM_MFMA, args.steps, args.warmup)
    t_plain, kavg_plain = timed(native.FORM_PLAIN, max(2, args.steps // 2), 1)
    ms_per_step = 1000.0 * t_step
    rows_per_s = nrl / t_step
    achieved = bytes_per_launch / (kavg / 1000.0) / 1e9

    out = {
        "metric": "reduced rows/s + dense modular GOPS, MFMA/batched vs plain HIP; bit-exact mod p",
        "value": rows_per_s,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint64 accumulators over uint32 residues mod p",
        "data": "synthetic (seeded F4-shaped matrix)",
        "config": {"workload": "sparse pass (known-pivot reduction), F4 matrix 20000x30000 mod 65521",
                   "form": "blocked (64-column group inverses)",
                   "rows": NR, "cols": NC, "known_pivots": NPIV, "to_reduce": nrl,
                   "p": P, "seed": SEED, "nonzero_remainders": nd},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": PMC_TRAFFIC_BYTES,
                     "kernel": "k_sparse_blocked", "kernel_ms": kavg,
                     "bytes_per_launch": bytes_per_launch,
                     "pivots_applied": napp, "pivot_entries": nent},
    }

    out["plain_form"] = {"rows_per_s": nrl / t_plain, "ms_per_step": 1000.0 * t_plain,
                         "kernel": "k_sparse_plain", "kernel_ms": kavg_plain,
                         "roofline_frac": bytes_per_launch / (kavg_plain / 1000.0) / 1e9 / HBM_PEAK_GBS}

    if not args.no_cpu:
        threads = int(os.environ.get("BENCH_CPU_THREADS", min(16, os.cpu_count() or 1)))
        tc = cpu_baseline(m, threads)
        out["cpu_baseline"] = {"value": nrl / tc, "unit": "rows/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the same matrix, all {nrl} to-reduce rows, one timed run after a warm-up"}

    if not args.no_extra:
        extra = {}
        for form, name in ((native.FORM_PLAIN, "plain"), (native.FORM_MFMA, "mfma")):
            d2 = ctx.upload(m)
            d2.la_sparse_dense(form)  # warm-up
            _sync()
            t0 = time.perf_counter()
            npiv = d2.la_sparse_dense(form)
            _sync()
            extra[f"la_option1_{name}_ms"] = 1000.0 * (time.perf_counter() - t0)
            extra[f"la_option1_{name}_dense_ms"] = ctx.phase_ms(1)
            extra["la_option1_new_pivots"] = npiv
            d2.free()
        # dense block elimination (4096 x 8192, full rank and rank 3000)
        for dp, rank in ((251, None), (P, None), (2147483629, None), (P, 3000)):
            a = synth.dense_block(4096, 8192, dp, seed=3, rank=rank)
            tag = f"dense_4096x8192_p{dp}" + ("" if rank is None else f"_rank{rank}")
            for form, name in ((native.FORM_PLAIN, "plain"), (native.FORM_MFMA, "mfma")):
                dd = ctx.dense(a, dp)
                dd.rref(form)
                dd.free()
                dd = ctx.dense(a, dp)
                _sync()
                t0 = time.perf_counter()
                r = dd.rref(form)
                _sync()
                td = time.perf_counter() - t0
                extra[f"{tag}_{name}_ms"] = 1000.0 * td
                extra[f"{tag}_{name}_gops"] = dense_ops(4096, 8192, r) / td / 1e9
                extra[f"{tag}_rank"] = r
                dd.free()
        # batched multimodular launch: 64 F4-shaped matrices of uneven sizes
        # (rows lognormal in [500, 20000], columns 1.5 x rows, density
        # 0.5-2 %), each modulo its own prime below 2^31; la_option 1 on all
        # of them as one graph launch, against the same matrices one by one
        mats = synth.batch_31bit(64, seed=7)
        nrl_b = sum(mm.nr - len(np.unique(mm.leads())) for mm in mats)
        bt = ctx.batch(mats)
        for form, name in ((native.FORM_MFMA, "mfma"), (native.FORM_PLAIN, "plain")):
            bt.la(form)  # device layout, tuning, graph capture
            ts, gs = [], []
            for _ in range(3):
                _sync()
                t0 = time.perf_counter()
                bt.la(form)
                _sync()
                ts.append(time.perf_counter() - t0)
                gs.append(ctx.phase_ms(2))
            tb = sorted(ts)[1]
            extra[f"batch64_{name}_ms"] = 1000.0 * tb
            extra[f"batch64_{name}_graph_ms"] = sorted(gs)[1]
            extra[f"batch64_{name}_rows_per_s"] = nrl_b / tb
            extra[f"batch64_{name}_tuned"] = bt.info()
        bt.free()
        singles = [ctx.upload(mm) for mm in mats]
        for dm in singles:
            dm.la_sparse_dense(native.FORM_MFMA)
        _sync()
        t0 = time.perf_counter()
        for dm in singles:
            dm.la_sparse_dense(native.FORM_MFMA)
        _sync()
        t1 = time.perf_counter() - t0
        extra["batch64_one_by_one_mfma_ms"] = 1000.0 * t1
        extra["batch64_to_reduce_rows"] = nrl_b
        for dm in singles:
            dm.free()
        out["extra"] = extra

    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
