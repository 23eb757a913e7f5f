"""Extract the golden vectors held by gb's own tests into JSON fixtures.

Reads the reference test programs as text (they are C files whose data is
inline arrays) and writes:

* ``normalize_row.json`` -- the known-answer test of
  normalize_dense_matrix_row_ff_32 (test/normalize_row.c:15-44);
* ``f4_golden.json`` -- for every end-to-end F4 test over a prime field:
  the input system (lens/cfs/exps, field characteristic, monomial order,
  la_option, reduce_gb, thread count) and, where the test asserts one, the
  expected basis (tlen/tcf/texp plus the asserted size and term count);
* ``qq_golden.json`` -- the tests over the rationals (field_char 0): the
  input system with its coefficients as (numerator, denominator) pairs
  (``mpz_set_si(*(cfs[i]), v)``), and the asserted basis size, term count
  and integer coefficients (``mpz_set_si(tcfs[i], v)``).

Run once in a container where the reference tree is present:

    python tests/golden/extract_golden.py /root/reference/test

The fixtures are committed; nothing at test time reads the reference.
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _array(src: str, name: str):
    m = re.search(r"\b" + name + r"\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}", src)
    if not m:
        return None
    body = m.group(1).replace("\n", " ")
    return [int(t) for t in body.split(",") if t.strip()]


def _scalar(src: str, name: str):
    m = re.search(r"\b" + name + r"\s*=\s*(-?\d+)\s*;", src)
    return int(m.group(1)) if m else None


def _asserted_sizes(src: str):
    m = re.search(r"\(\*bld\)\s*!=\s*(\d+)\s*\|\|\s*nterms\s*!=\s*(\d+)", src)
    return (int(m.group(1)), int(m.group(2))) if m else (None, None)


def extract(test_dir: str):
    cases = {}
    for fn in sorted(os.listdir(test_dir)):
        if not fn.endswith(".c"):
            continue
        src = open(os.path.join(test_dir, fn)).read()
        if "f4_julia(" not in src:
            continue
        fc = _scalar(src, "field_char")
        if fc is None or fc == 0:
            continue  # rationals: out of scope
        lens = _array(src, "lens")
        cfs = _array(src, "cfs")
        exps = _array(src, "exps")
        if lens is None or cfs is None or exps is None:
            continue
        case = {
            "file": "test/" + fn,
            "field_char": fc,
            "mon_order": _scalar(src, "mon_order"),
            "nr_vars": _scalar(src, "nr_vars"),
            "nr_gens": _scalar(src, "nr_gens"),
            "nr_threads": _scalar(src, "nr_threads"),
            "la_option": _scalar(src, "la_option"),
            "reduce_gb": _scalar(src, "reduce_gb"),
            "lens": lens, "cfs": cfs, "exps": exps,
        }
        bld, nterms = _asserted_sizes(src)
        tlen, tcf, texp = _array(src, "tlen"), _array(src, "tcf"), _array(src, "texp")
        if tlen is not None and tcf is not None and texp is not None:
            case.update({"bld": bld, "nterms": nterms,
                         "tlen": tlen, "tcf": tcf, "texp": texp})
        cases[fn[:-2]] = case
    return cases


def _mpz_sets(src: str, name: str):
    """name[i] = v from mpz_set_si(*(name[i]), v) / mpz_set_si(name[i], v)."""
    pat = r"mpz_set_si\(\s*\*?\(?\s*" + name + r"\[(\d+)\]\s*\)?\s*,\s*(-?\d+)\s*\)"
    vals = {int(i): int(v) for i, v in re.findall(pat, src)}
    return [vals[i] for i in range(len(vals))] if vals else None


def extract_qq(test_dir: str):
    cases = {}
    for fn in sorted(os.listdir(test_dir)):
        if not fn.endswith(".c"):
            continue
        src = open(os.path.join(test_dir, fn)).read()
        if "f4_julia(" not in src or _scalar(src, "field_char") != 0:
            continue
        cfs = _mpz_sets(src, "cfs")
        lens, exps = _array(src, "lens"), _array(src, "exps")
        if cfs is None or lens is None or exps is None:
            continue
        case = {
            "file": "test/" + fn,
            "mon_order": _scalar(src, "mon_order"),
            "nr_vars": _scalar(src, "nr_vars"),
            "nr_gens": _scalar(src, "nr_gens"),
            "la_option": _scalar(src, "la_option"),
            "reduce_gb": _scalar(src, "reduce_gb"),
            "lens": lens, "exps": exps,
            "num": cfs[0::2], "den": cfs[1::2],
        }
        m = re.search(r"\(\*bld\)\s*!=\s*(\d+)\s*\|\|\s*ret\s*!=\s*(\d+)", src)
        tcfs = _mpz_sets(src, "tcfs")
        if m and tcfs is not None:
            case.update({"bld": int(m.group(1)), "nterms": int(m.group(2)), "tcf": tcfs})
        cases[fn[:-2]] = case
    return cases


def extract_normalize(test_dir: str):
    src = open(os.path.join(test_dir, "normalize_row.c")).read()
    fc = _scalar(src, "fc")
    checks = {int(i): int(v) for i, v in
              re.findall(r"row\[(\d+)\]\s*!=\s*(\d+)", src)}
    return {"file": "test/normalize_row.c", "fc": fc,
            "row0": 16, "row_i": "300 - i for i in 1..15", "len": 16,
            "expect": checks}


def main():
    test_dir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test"
    with open(os.path.join(HERE, "f4_golden.json"), "w") as f:
        json.dump(extract(test_dir), f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "qq_golden.json"), "w") as f:
        json.dump(extract_qq(test_dir), f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "normalize_row.json"), "w") as f:
        json.dump(extract_normalize(test_dir), f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
