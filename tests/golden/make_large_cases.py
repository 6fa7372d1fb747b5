"""Writes large_cases.json: the oracle's results on full-size problems whose
single solve takes the oracle seconds to minutes, so the GPU tests compare
against stored values: status, iterations and objective of the problem as
generated, and the range of iteration counts the oracle itself takes under
1e-14 relative perturbations of q and b (8 runs), the envelope the GPU result
is held to.  Run from the repository root (about an hour on four cores):
python tests/golden/make_large_cases.py [--single] [--engine-rule] [case ...];
cases not named keep their stored entries.  --single stores one oracle run
without the perturbed copies (the envelope is then that run's count alone):
used for the full-size PSD relaxation, whose single oracle solve takes 1.6
hours on one core.  The entry's top-level fields hold the oracle under the
reference's settings; --engine-rule writes the entry's "engine_rule" record
instead: the oracle with the engine's default dynamic-regularization rule
(a wrongly signed pivot keeps its magnitude above delta), what the GPU
tests compare against."""
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
    # PSD relaxation: 150 PSD cones of side 10, 300 nonnegative rows, n = 1000 (KKT 9550 rows)
    "psd_relaxation_1000x150": ("psd_relaxation", [4, 1000, 150, 10, 300, 0.2]),
    # the full PSD workload: 500 PSD cones of side 10, 1000 nonnegative rows, n = 3000 (KKT 31 500 rows)
    "psd_relaxation_3000x500": ("psd_relaxation", [4, 3000, 500, 10, 1000, 0.2]),
    # the random sparse SOCP at a quarter of the full size: n = 5000, m = 7500 (1250 zero, 3750 nonnegative,
    # 500 SOC(5)), 8 entries per row (KKT 12 500 rows, a near-dense root as at the full size)
    "socp_5000x7500": ("large_socp", [2, 5000, 1250, 3750, 500]),
}
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "large_cases.json")


def _run(job):
    fn, args, seed, keep = job
    P, q, A, b, cones = getattr(problems, fn)(*args)
    if seed >= 0:
        rng = np.random.RandomState(seed)
        q = q * (1 + 1e-14 * rng.standard_normal(q.size))
        b = b * (1 + 1e-14 * rng.standard_normal(b.size))
    pyoracle.build()
    return pyoracle.solve(P, q, A, b, cones, dynamic_regularization_keep_magnitude=keep)


def main(names):
    from multiprocessing import Pool
    single = "--single" in names
    keep = 1 if "--engine-rule" in names else 0
    names = [n for n in names if n not in ("--single", "--engine-rule")]
    pyoracle.build()
    for name in names or list(CASES):
        fn, args = CASES[name]
        with Pool(4) as pool:  # the problem as generated, then 8 perturbed copies
            rs = pool.map(_run, [(fn, args, -1, keep)] + ([] if single else [(fn, args, s, keep) for s in range(8)]))
        r, its = rs[0], [x["iterations"] for x in rs]
        out = json.load(open(OUT)) if os.path.exists(OUT) else {}  # read late: other runs may have written
        rec = {"status": r["status"], "iterations": r["iterations"], "obj_val": r["obj_val"],
               "iterations_envelope": [min(its), max(its)], "statuses": sorted({x["status"] for x in rs})}
        if keep:
            out.setdefault(name, {"generator": fn, "args": args})["engine_rule"] = rec
        else:
            prev = out.get(name, {}).get("engine_rule")
            out[name] = {"generator": fn, "args": args, **rec}
            if prev:
                out[name]["engine_rule"] = prev
        print(name, out[name], flush=True)
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
