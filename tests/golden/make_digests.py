"""Writes tests/golden/full_size_digests.json: SHA-256 digests of the
oracle's results on the benchmark configurations at their full sizes, for
the GPU tests that cannot run the oracle there in seconds.

  c2  la_option 1 on the seeded 20000 x 30000 F4 matrix mod 65521 (rows and
      count), and the sparse pass's work counts (pivot rows applied, pivot
      entries streamed: the roofline numerator)
  c3  the reduced echelon form of the 4096 x 8192 blocks mod 251, 65521 and
      2147483629, and of the rank-3000 block mod 65521
  c4  la_option 1 on each of the 64 matrices of synth.batch_31bit(64, 7)
      (rows 500-20000, each modulo its own prime below 2^31)
  c5  the F4 basis of 14 dense quadratics in 14 variables mod 65521, DRL
      (and of the 12-variable system), with the per-round trace: matrix
      shape, new rows and a hash of every round's echelon form

The oracle is pinned against gb's own golden vectors and brute force by
tests/test_oracle.py and tests/test_f4_host.py.  Run in a CPU container
(takes about half an hour on 8 cores):

    python tests/golden/make_digests.py [c2,c3,c4,c5] [--threads N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "full_size_digests.json")

# the inputs, as bench.py builds them
C2 = dict(nr=20000, nc=30000, npiv=12000, p=65521, seed=2024)
C3 = [dict(p=251, rank=None), dict(p=65521, rank=None), dict(p=2147483629, rank=None),
      dict(p=65521, rank=3000)]
C3_SHAPE = dict(m=4096, n=8192, seed=3)
C4 = dict(count=64, seed=7)
C5_VARS = (12, 14)


def c3_key(p, rank):
    return f"p{p}" + ("" if rank is None else f"_rank{rank}")


def main(argv=None) -> int:
    import numpy as np
    from gb_amd import synth
    from oracle import oracle
    from tests import digests

    ap = argparse.ArgumentParser()
    ap.add_argument("parts", nargs="?", default="c2,c3,c4,c5")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args(argv)
    parts = args.parts.split(",")
    th = args.threads
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}

    def save():
        with open(OUT, "w") as f:
            json.dump(res, f, indent=1, sort_keys=True)
            f.write("\n")

    if "c2" in parts:
        t = time.time()
        m = synth.f4_shaped(C2["nr"], C2["nc"], C2["npiv"], C2["p"], seed=C2["seed"])
        nd, napp, nent = oracle.sparse_pass_work(m, nthreads=th)
        r = oracle.la_sparse_dense(m, nthreads=th)
        res["c2"] = dict(C2, nonzero_remainders=nd, pivots_applied=napp, pivot_entries=nent,
                         new_pivots=r.np, la_option1=digests.sparse_rows(r))
        save()
        print(f"c2: {time.time() - t:.1f} s", flush=True)
    if "c3" in parts:
        res.setdefault("c3", {"shape": C3_SHAPE})
        for c in C3:
            t = time.time()
            a = synth.dense_block(C3_SHAPE["m"], C3_SHAPE["n"], c["p"], seed=C3_SHAPE["seed"],
                                  rank=c["rank"])
            rows, lead = oracle.dense_rref(a, c["p"], th)
            res["c3"][c3_key(c["p"], c["rank"])] = dict(p=c["p"], rank=len(lead),
                                                        rref=digests.dense(rows, lead))
            save()
            print(f"c3 {c}: rank {len(lead)}, {time.time() - t:.1f} s", flush=True)
    if "c4" in parts:
        t = time.time()
        mats = synth.batch_31bit(C4["count"], seed=C4["seed"])
        per = []
        for mm in mats:
            r = oracle.la_sparse_dense(mm, nthreads=th)
            per.append(dict(p=int(mm.p), nr=int(mm.nr), nc=int(mm.nc), new_pivots=r.np,
                            la_option1=digests.sparse_rows(r)))
        res["c4"] = dict(C4, matrices=per,
                         to_reduce_rows=int(sum(mm.nr - len(np.unique(mm.leads())) for mm in mats)))
        save()
        print(f"c4: {time.time() - t:.1f} s", flush=True)
    if "c5" in parts:
        res.setdefault("c5", {})
        for nv in C5_VARS:
            t = time.time()
            s = synth.dense_quadratic_system(nv, 65521, seed=1)
            b = oracle.f4(s, nthreads=th)
            res["c5"][f"vars{nv}"] = dict(vars=nv, p=65521, seed=1, basis=digests.basis(b),
                                          elements=len(b.lens), terms=int(np.sum(b.lens)),
                                          trace=[list(r.values()) for r in b.stats["trace"]])
            save()
            print(f"c5 {nv} vars: {time.time() - t:.1f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
