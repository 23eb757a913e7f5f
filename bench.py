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
    routine, OpenMP) on the same matrix on this host: median of 5 timed runs
    on the host's CPU share, and a 1-thread median of 3;
  * extra: the end-to-end F4 runs on 14 dense quadratics in 14 variables
    (plain and MFMA forms, and the same run with the oracle's linear
    algebra on the host's cores), the full la_option 1 step on the same
    matrix in both forms, the dense 4096 x 8192 block elimination in
    modular GOPS, and the 64-matrix batch; each result checked against the
    oracle's digest of it (tests/golden/full_size_digests.json) outside the
    timed region, as `*_digest_ok`.
Every libgbgpu call synchronises its stream before it returns, so the host
clock around a call covers its device work.
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
# FETCH_SIZE (8,068,585 KiB, x2.00) and WRITE_SIZE (940,417 KiB, x1.00), each
# scaled by what the same counter reads for a plain 256 MiB device copy
# (131,088 and 262,144 KiB); the commands and raw counts are in
# profiles/k_sparse_blocked_f4_20000x30000_p65521.md
PMC_TRAFFIC_BYTES = 8068585 * 1024 * 2.0 + 940417 * 1024 * 1.0


# the CPU share of one GPU on the benchmark host: worker pools never exceed it
CPU_SHARE = 16


def full_size_digests():
    """The oracle's SHA-256 digests of the full-size results
    (tests/golden/full_size_digests.json, written on the CPU by
    tests/golden/make_digests.py): each extra's result is checked against
    them outside its timed region (`*_digest_ok`), or None without them."""
    path = os.path.join(ROOT, "tests", "golden", "full_size_digests.json")
    try:
        with open(path) as f:
            return json.load(f)
    except OSError:
        return None


def cpu_threads() -> int:
    n = os.cpu_count() or 1
    return int(os.environ.get("BENCH_CPU_THREADS", min(CPU_SHARE, n)))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class _RoundLinesToStderr:
    """Send what the C code prints on stdout (the F4 host's per-round lines)
    to stderr, so that stdout carries the JSON line only."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        import ctypes
        ctypes.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def dense_ops(m: int, n: int, r: int) -> float:
    """Multiply-adds x 2 of a reduced echelon form of an m x n block of rank
    r: elimination 2mnr - r^2(m+n) + 2/3 r^3 plus back substitution
    r^2 (n - r)."""
    return 2.0 * m * n * r - r * r * (m + n) + (2.0 / 3.0) * r ** 3 + r * r * (n - r)


def cpu_baseline(m, nthreads: int, runs: int) -> float:
    """Median wall time of the oracle's sparse pass (C + OpenMP) over the
    same matrix, after one warm-up run."""
    from oracle import oracle
    oracle.sparse_pass_work(m, nthreads=nthreads)  # warm-up (page-in)
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.sparse_pass_work(m, nthreads=nthreads)
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def f4_runs(nvars: int, ctx, with_cpu: bool, forms=("mfma", "plain"), cpu_vars: int = 14,
            info: int = 0) -> dict:
    """End to end F4 (the F4 host, every round's linear algebra on the GPU)
    on a seeded dense quadratic system: nvars quadratics in nvars
    variables, every monomial of degree <= 2, coefficients uniform in
    [1, p - 1] (splitmix64, seed 1), p = 65521, DRL; plain and MFMA forms;
    and, with with_cpu, the same run with the oracle's linear algebra on
    this host's cores (on the cpu_vars-variable system: the same one when
    cpu_vars >= nvars, otherwise a smaller system as the bounded sample,
    with the GPU run on it beside)."""
    import numpy as np
    from gb_amd import native, synth
    from tests import digests
    res = {"system": f"{nvars} dense quadratics in {nvars} variables mod 65521, DRL, seed 1"}
    s = synth.dense_quadratic_system(nvars, 65521, seed=1)
    fix = (full_size_digests() or {}).get("c5", {}).get(f"vars{nvars}")
    runs = {}
    for form, name in ((native.FORM_MFMA, "mfma"), (native.FORM_PLAIN, "plain")):
        if name not in forms:
            continue
        print(f"bench: F4 on {nvars} variables, {name} form", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        with _RoundLinesToStderr():
            b = ctx.f4(s, form=form, info_level=info)
        cold = time.perf_counter() - t0
        if name == "mfma":
            # again, with the device memory pool and the host's allocator
            # warm: the steady-state figure of a process that runs F4 more
            # than once (the first run's wall time is kept as total_s_first)
            print(f"bench: F4 on {nvars} variables, {name} form, second run", file=sys.stderr,
                  flush=True)
            t0 = time.perf_counter()
            b = ctx.f4(s, form=form)
        runs[name] = b
        st = b.stats
        res[name] = {"total_s": time.perf_counter() - t0, "total_s_first": cold,
                     "la_s": st["la_ms"] / 1e3,
                     "la_device_s": st["device_ms"] / 1e3,
                     "host_s": (st["total_ms"] - st["la_ms"]) / 1e3,
                     "host_phases_s": {k: st[k + "_ms"] / 1e3
                                       for k in ("select", "symbolic", "build", "update")},
                     "rounds": st["rounds"],
                     "basis_elements": len(b.lens), "max_matrix": [st["max_rows"], st["max_cols"]],
                     "max_nnz": st["max_nnz"], "rows_reduced": st["rows_total"],
                     "rows_per_s_la": st["rows_total"] / (st["la_ms"] / 1e3)}
        if fix:
            res[name]["basis_digest_ok"] = digests.basis(b) == fix["basis"]
    a = runs.get("mfma") or runs.get("plain")
    if len(runs) == 2:
        b2 = runs["plain"]
        res["forms_identical"] = bool(np.array_equal(a.lens, b2.lens) and np.array_equal(a.exps, b2.exps)
                                      and np.array_equal(a.cfs, b2.cfs))
        res["mfma_vs_plain_total"] = res["plain"]["total_s"] / res["mfma"]["total_s"]
        res["mfma_vs_plain_la"] = res["plain"]["la_s"] / res["mfma"]["la_s"]
    if with_cpu:
        from oracle import oracle
        n_cpu = min(nvars, cpu_vars)
        s_cpu = s if n_cpu == nvars else synth.dense_quadratic_system(n_cpu, 65521, seed=1)
        threads = cpu_threads()
        print(f"bench: F4 on {n_cpu} variables, oracle LA on {threads} threads",
              file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        with _RoundLinesToStderr():
            c = oracle.f4(s_cpu, nthreads=threads, info_level=info)
        tc = time.perf_counter() - t0
        cpu = {"vars": n_cpu, "threads": threads, "total_s": tc, "la_s": c.stats["la_ms"] / 1e3,
               "host_s": (c.stats["total_ms"] - c.stats["la_ms"]) / 1e3}
        if n_cpu == nvars:
            cpu["identical_to_gpu"] = bool(np.array_equal(c.lens, a.lens) and np.array_equal(c.exps, a.exps)
                                           and np.array_equal(c.cfs, a.cfs))
            for name in runs:
                res[f"{name}_vs_cpu_total"] = tc / res[name]["total_s"]
                res[f"{name}_vs_cpu_la"] = cpu["la_s"] / res[name]["la_s"]
        else:
            g = ctx.f4(s_cpu, form=native.FORM_MFMA)
            cpu["gpu_mfma_total_s_same_system"] = g.stats["total_ms"] / 1e3
            cpu["gpu_mfma_la_s_same_system"] = g.stats["la_ms"] / 1e3
            cpu["identical_to_gpu"] = bool(np.array_equal(c.lens, g.lens) and np.array_equal(c.exps, g.exps)
                                           and np.array_equal(c.cfs, g.cfs))
        res["cpu_oracle"] = cpu
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the extra workloads")
    ap.add_argument("--f4-vars", type=int, default=14,
                    help="variables of the dense quadratic system run end to end by the F4 host "
                         "(14 = the 14-variable configuration; plain and MFMA forms)")
    ap.add_argument("--only-f4", action="store_true", help="only the end-to-end F4 runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-matrix batch extra")
    ap.add_argument("--no-f4", action="store_true", help="skip the end-to-end F4 extra")
    ap.add_argument("--f4-forms", default="mfma,plain", help="forms of the F4 runs")
    ap.add_argument("--no-f4-info", action="store_true",
                    help="no per-round lines of the F4 runs (on stderr)")
    ap.add_argument("--f4-cpu-vars", type=int, default=14,
                    help="variables of the system the CPU (oracle LA) F4 run uses")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: the F4 step runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2

    import numpy as np
    from gb_amd import native, synth

    ctx = native.Context(0)
    if args.only_f4:
        print(json.dumps({"f4": f4_runs(args.f4_vars, ctx, not args.no_cpu,
                                        tuple(args.f4_forms.split(",")), args.f4_cpu_vars,
                                        0 if args.no_f4_info else 2)}))
        return 0
    m = synth.f4_shaped(NR, NC, NPIV, P, seed=SEED)
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
    # compulsory bytes of the reduction kernel: every byte it must touch
    # once -- the pivot layout (8 bytes per pivot entry, leads excluded),
    # the groups' inverses (64 x 64 words per 64 pivot columns), the rows to
    # reduce, and the D' rows it writes
    piv_entries = int(lens[first].sum()) - int(first.sum())
    compulsory = (8.0 * piv_entries + 4.0 * 4096 * ((m.ncl + 63) // 64) + 8.0 * red_nnz
                  + 4.0 * m.ncr * nrl)

    def timed(form, steps, warmup):
        for _ in range(warmup):
            d.sparse_pass(form)
        kernel_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            d.sparse_pass(form)
            kernel_ms.append(ctx.phase_ms(0))
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
                     "pivots_applied": napp, "pivot_entries": nent,
                     # algorithmic bytes count every pivot entry each time a
                     # row streams it (caches serve the repeats); HBM traffic
                     # and compulsory bytes beside them
                     "compulsory_bytes": compulsory,
                     "traffic_over_compulsory": PMC_TRAFFIC_BYTES / compulsory,
                     "hbm_gbs": PMC_TRAFFIC_BYTES / (kavg / 1000.0) / 1e9,
                     "hbm_frac": PMC_TRAFFIC_BYTES / (kavg / 1000.0) / 1e9 / HBM_PEAK_GBS},
    }

    out["plain_form"] = {"rows_per_s": nrl / t_plain, "ms_per_step": 1000.0 * t_plain,
                         "kernel": "k_sparse_plain", "kernel_ms": kavg_plain,
                         "roofline_frac": bytes_per_launch / (kavg_plain / 1000.0) / 1e9 / HBM_PEAK_GBS}

    if not args.no_cpu:
        threads = cpu_threads()
        print(f"bench: CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        tc = cpu_baseline(m, threads, 5)
        tc1 = cpu_baseline(m, 1, 3)
        out["cpu_baseline"] = {"value": nrl / tc, "unit": "rows/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the same matrix, all {nrl} to-reduce rows: median of 5 timed "
                                         f"runs after a warm-up",
                               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
                               "cpu_share": CPU_SHARE,
                               "one_thread": {"value": nrl / tc1, "unit": "rows/s",
                                              "sample": "median of 3 timed runs"},
                               "gpu_vs_cpu": rows_per_s / (nrl / tc)}

    if not args.no_extra:
        from tests import digests
        fix = full_size_digests() or {}
        extra = {}
        if not args.no_f4:
            extra["f4"] = f4_runs(args.f4_vars, ctx, not args.no_cpu, tuple(args.f4_forms.split(",")),
                                  args.f4_cpu_vars, 0 if args.no_f4_info else 2)
        for form, name in ((native.FORM_PLAIN, "plain"), (native.FORM_MFMA, "mfma")):
            d2 = ctx.upload(m)
            d2.la_sparse_dense(form)  # warm-up
            t0 = time.perf_counter()
            npiv = d2.la_sparse_dense(form)
            extra[f"la_option1_{name}_ms"] = 1000.0 * (time.perf_counter() - t0)
            extra[f"la_option1_{name}_dense_ms"] = ctx.phase_ms(1)
            extra["la_option1_new_pivots"] = npiv
            if "c2" in fix:
                extra[f"la_option1_{name}_digest_ok"] = digests.sparse_rows(d2.rows()) == fix["c2"]["la_option1"]
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
                t0 = time.perf_counter()
                r = dd.rref(form)
                td = time.perf_counter() - t0
                extra[f"{tag}_{name}_ms"] = 1000.0 * td
                extra[f"{tag}_{name}_gops"] = dense_ops(4096, 8192, r) / td / 1e9
                extra[f"{tag}_rank"] = r
                key = tag[len("dense_4096x8192_"):]
                if key in fix.get("c3", {}):
                    rows, lead = dd.result()
                    extra[f"{tag}_{name}_digest_ok"] = digests.dense(rows, lead) == fix["c3"][key]["rref"]
                dd.free()
        # batched multimodular launch: 64 F4-shaped matrices of uneven sizes
        # (rows lognormal in [500, 20000], columns 1.5 x rows, density
        # 0.5-2 %), each modulo its own prime below 2^31; la_option 1 on all
        # of them as one graph launch, against the same matrices one by one
        mats = [] if args.no_batch else synth.batch_31bit(64, seed=7)
        nrl_b = sum(mm.nr - len(np.unique(mm.leads())) for mm in mats)
        bt = ctx.batch(mats) if mats else None
        for form, name in ((native.FORM_MFMA, "mfma"), (native.FORM_PLAIN, "plain")):
            if bt is None:
                break
            bt.la(form)  # device layout, tuning, graph capture
            ts, gs = [], []
            for _ in range(3):
                t0 = time.perf_counter()
                bt.la(form)
                ts.append(time.perf_counter() - t0)
                gs.append(ctx.phase_ms(2))
            tb = sorted(ts)[1]
            extra[f"batch64_{name}_ms"] = 1000.0 * tb
            extra[f"batch64_{name}_graph_ms"] = sorted(gs)[1]
            extra[f"batch64_{name}_rows_per_s"] = nrl_b / tb
            extra[f"batch64_{name}_tuned"] = bt.info()
            if "c4" in fix:
                extra[f"batch64_{name}_digest_ok"] = all(
                    digests.sparse_rows(bt.rows(i)) == x["la_option1"]
                    for i, x in enumerate(fix["c4"]["matrices"]))
        if bt is not None:
            bt.free()
        singles = [ctx.upload(mm) for mm in mats]
        for dm in singles:
            dm.la_sparse_dense(native.FORM_MFMA)
        t0 = time.perf_counter()
        for dm in singles:
            dm.la_sparse_dense(native.FORM_MFMA)
        t1 = time.perf_counter() - t0
        if mats:
            extra["batch64_one_by_one_mfma_ms"] = 1000.0 * t1
            extra["batch64_to_reduce_rows"] = nrl_b
        for dm in singles:
            dm.free()
        out["extra"] = extra

    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
