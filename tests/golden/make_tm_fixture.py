#!/usr/bin/env python3
"""Extract the numeric values of the reference's generated transform tables
into a JSON fixture (``tm_literals.json``).

The reference's ``TransformMatrices.h`` is the SageMath generator's output: a
list of ``rslt(i,j)=<decimal>;`` assignments per table.  This script reads that
file as TEXT (it is never compiled or executed), keeps only the tables the
ADER-WENO step uses (GLL nodes/weights, gll<->coefs, coefs_to_deriv,
sten_to_coefs, coefs_to_gll_lower, weno_sten_to_coefs, and the coefficients of
the coefs_to_tv quadratic forms), and writes index -> decimal-string records.
The decimal strings are kept verbatim (38 significant digits) so the tests can
compare them with the derived tables at the precision they were printed with.

Run in the container that has the reference mounted:
    python tests/golden/make_tm_fixture.py /root/reference/TransformMatrices.h
"""
import json
import os
import re
import sys

WANT_MATRIX = {"get_gll_points", "get_gll_weights", "gll_to_coefs", "coefs_to_gll",
               "coefs_to_deriv", "sten_to_coefs", "coefs_to_gll_lower", "weno_sten_to_coefs"}
HEAD = re.compile(r"template <class FP> YAKL_INLINE (void|FP) (\w+)\((.*)\) \{")
ASSIGN = re.compile(r"rslt\(([0-9,]+)\)=(.+);")
TV_SQ = re.compile(r"([-0-9.e+]+)_fp\*\(a\((\d+)\)\*a\((\d+)\)\)")
TV_X = re.compile(r"([-0-9.e+]+)_fp\*a\((\d+)\)\*a\((\d+)\)")


def parse(path):
    out = {}
    cur = None
    for line in open(path):
        m = HEAD.search(line)
        if m:
            name, args = m.group(2), m.group(3)
            cur = None
            if "FP dx" in args:
                continue
            dims = [int(v) for v in re.search(r"SArray<FP,\d+,([0-9,]+)>", args).group(1).split(",")]
            if name in WANT_MATRIX:
                key = "%s_%s" % (name, "_".join(str(v) for v in dims))
                cur = key
                out[key] = {"dims": dims, "entries": []}
            elif name == "coefs_to_tv":
                cur = "coefs_to_tv_%d" % dims[0]
                out[cur] = {"dims": dims, "terms": []}
            continue
        if cur is None:
            continue
        if line.strip() == "}":
            cur = None
            continue
        if cur.startswith("coefs_to_tv"):
            if "rslt=" in line:
                expr = line.split("rslt=", 1)[1]
                for t in re.split(r"\+(?=[0-9-])", expr.strip().rstrip(";")):
                    m2 = TV_SQ.fullmatch(t) or TV_X.fullmatch(t)
                    if m2 is None:
                        raise ValueError("unparsed TV term %r" % t)
                    out[cur]["terms"].append([int(m2.group(2)), int(m2.group(3)), m2.group(1)])
            continue
        m = ASSIGN.search(line)
        if m:
            idx = [int(v) for v in m.group(1).split(",")]
            out[cur]["entries"].append(idx + [m.group(2)])
    return out


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/TransformMatrices.h"
    tabs = parse(src)
    keep = {}
    for k, v in tabs.items():
        n = v["dims"][0]
        if k.startswith(("sten_to_coefs", "coefs_to_gll_lower", "weno_sten_to_coefs")) and n not in (3, 5, 7, 9):
            continue
        keep[k] = v
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tm_literals.json")
    with open(dst, "w") as f:
        json.dump({"source": "TransformMatrices.h (reference, generated tables; numbers only)",
                   "tables": keep}, f, indent=0, sort_keys=True)
    print("wrote", dst, len(keep), "tables")


if __name__ == "__main__":
    main()
