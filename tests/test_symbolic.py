"""The host symbolic analysis: column counts by row subtrees and supernodal
row lists built from the children's lists (the default) against the
explicit per-column row structures (CLH_SYMBOLIC_EXPLICIT=1).  The two must
give the same factor shape and, the analysis being the same, bitwise the
same solution on the serial build of the engine.  (The library caches
analyses by shape for the process, so each rule runs in a process of its
own.)"""
import json
import os
import subprocess
import sys

import pytest

from helpers import EMU_LIB, ROOT, build_emu

_RUN = r'''
import json, sys, hashlib
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import cuclarabel_amd as ca
from cuclarabel_amd import _lib, problems
from helpers import LOGISTIC_CASES, SDP_CASES, SPARSE_SOC_CASES, SOCP_CASES, socp_case
emu = _lib.bind({lib!r})
cases = [("mpc", problems.mpc(2, T=20)), ("portfolio", problems.portfolio(1)),
         ("socp", socp_case(SOCP_CASES[0])), ("sparse_soc", socp_case(SPARSE_SOC_CASES[0])),
         ("sdp", problems.psd_relaxation(*SDP_CASES[1])), ("logistic", problems.logistic_exp_pow(*LOGISTIC_CASES[0]))]
out = {{}}
for engine in ("workgroup", "grid"):
    for method in ("hip", "hipmixed"):
        for name, prob in cases:
            sol = ca.Solver(*prob, ca.Settings(direct_solve_method=method, serve_engine=engine), _library=emu).solve()
            out[f"{{engine}}/{{method}}/{{name}}"] = [sol.info["nnz_L"], sol.info["factor_flops"], sol.info["n_supernodes"],
                                                   sol.info["n_levels"], sol.status, sol.iterations,
                                                   hashlib.sha1(sol.x.tobytes()).hexdigest()]
print(json.dumps(out))
'''


def _run(explicit):
    env = dict(os.environ)
    env.pop("CLH_SYMBOLIC_EXPLICIT", None)
    if explicit:
        env["CLH_SYMBOLIC_EXPLICIT"] = "1"
    code = _RUN.format(root=ROOT, tests=os.path.dirname(os.path.abspath(__file__)), lib=EMU_LIB)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_counts_and_row_lists_match_explicit_structures():
    build_emu()
    a, b = _run(True), _run(False)
    assert a.keys() == b.keys() and len(a) == 24
    for k in a:
        assert a[k] == b[k], (k, a[k], b[k])
