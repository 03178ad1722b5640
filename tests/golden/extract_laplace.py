"""Extract the starting values of the minimax exponential sums 1/x ~ sum w e^{-a x}
on [1, R] that the reference's Laplace module reads
(src/math/laplace/data/init_para.txt and init_error.txt, tables from
http://www.mis.mpg.de/scicomp/EXP_SUM/1_x/), orders k = 1..12, into
chem_amd/data/laplace_init.json.

Usage: python tests/golden/extract_laplace.py /root/reference/src/math/laplace/data
"""
import json
import os
import sys


def parse_token_range(tok):
    # "1_xk08_4E1" -> 40.0 ; "1_xk01_1100" -> 1.1
    r = tok[7:]
    if r[1] == "E":
        return int(r[0]) * 10.0 ** int(r[2:])
    return int(r[0]) + int(r[1:4]) / 1000.0


def main(src):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "chem_amd", "data")
    errs = {}
    for line in open(os.path.join(src, "init_error.txt")):
        if line.startswith("1_xk"):
            t, v = line.split()[:2]
            errs[t] = float(v)
    out = {}
    cur = None
    vals = []
    for line in open(os.path.join(src, "init_para.txt")):
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        if s.startswith("1_xk"):
            if cur is not None:
                out[cur] = vals
            cur, vals = s, []
        else:
            vals.append(s.split()[0])
    if cur is not None:
        out[cur] = vals
    table = {}
    for tok, v in out.items():
        k = int(tok[4:6])
        if k > 12 or len(v) != 2 * k:
            continue
        table.setdefault(str(k), []).append({
            "token": tok, "R": parse_token_range(tok), "omega": v[:k], "alpha": v[k:],
            "err": errs.get(tok)})
    with open(os.path.join(here, "laplace_init.json"), "w") as f:
        json.dump({"source": "EXP_SUM 1/x tables (init_para.txt, init_error.txt)",
                   "orders": table}, f, indent=0)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/math/laplace/data")
