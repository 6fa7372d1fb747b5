"""Writes large_cases.json: the oracle's results on full-size problems whose
single solve takes the oracle seconds to minutes, so the GPU tests compare
against stored values: status, iterations and objective of the problem as
generated, and the range of iteration counts the oracle itself takes under
1e-14 relative perturbations of q and b (8 runs), the envelope the GPU result
is held to.  Run from the repository root (a few minutes on one core):
python tests/golden/make_large_cases.py"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from cuclarabel_amd import problems  # noqa: E402
from oracle import pyoracle  # noqa: E402

CASES = {
    # exp/pow logistic regression at the full workload size: 5000 samples x 200 features, 2000 power cones
    "logistic_5000x200": ("logistic_exp_pow", [3, 5000, 200, 2000]),
}


def main():
    pyoracle.build()
    out = {}
    for name, (fn, args) in CASES.items():
        P, q, A, b, cones = getattr(problems, fn)(*args)
        r = pyoracle.solve(P, q, A, b, cones)
        rng = np.random.RandomState(0)
        its = []
        for _ in range(8):
            qq = q * (1 + 1e-14 * rng.standard_normal(q.size))
            bb = b * (1 + 1e-14 * rng.standard_normal(b.size))
            its.append(pyoracle.solve(P, qq, A, bb, cones)["iterations"])
        out[name] = {"generator": fn, "args": args, "status": r["status"], "iterations": r["iterations"],
                     "obj_val": r["obj_val"], "iterations_envelope": [min(its + [r["iterations"]]),
                                                                      max(its + [r["iterations"]])]}
        print(name, out[name], flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "large_cases.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
