"""Write tests/golden/laplace_k08.json: the reference's tabulated 8-point
minimax solutions for 1/x on [1, R] at R = 10, 20, 30, 40, 100, 1000
(src/math/laplace/data/init_para.txt, init_error.txt), as a fixture.

Usage: python tests/golden/make_laplace_golden.py /root/reference/src/math/laplace/data
"""
import json
import os
import sys

here = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, here)
from extract_laplace import parse_token_range  # noqa: E402


def main(src):
    want = {"1_xk08_1E1", "1_xk08_2E1", "1_xk08_3E1", "1_xk08_4E1", "1_xk08_1E2", "1_xk08_1E3"}
    errs = {}
    for line in open(os.path.join(src, "init_error.txt")):
        if line.startswith("1_xk"):
            t, v = line.split()[:2]
            errs[t] = float(v)
    entries, cur, vals = [], None, []
    for line in list(open(os.path.join(src, "init_para.txt"))) + ["1_xkEND"]:
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        if s.startswith("1_xk"):
            if cur in want:
                entries.append({"token": cur, "R": parse_token_range(cur), "omega": vals[:8],
                                "alpha": vals[8:], "err": errs[cur]})
            cur, vals = s, []
        else:
            vals.append(s.split()[0])
    entries.sort(key=lambda e: e["R"])
    with open(os.path.join(here, "laplace_k08.json"), "w") as f:
        json.dump({"source": "init_para.txt / init_error.txt (k = 8)", "entries": entries}, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/math/laplace/data")
