import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def reference_oracle():
    """the oracle with the reference's settings throughout"""
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


class EngineRuleOracle:
    """The oracle as parity tests compare against it: the reference's
    algorithm with the engine's default dynamic-regularization rule (a pivot
    on the wrong side of zero keeps its magnitude when that is above delta,
    Settings.dynamic_regularization_keep_magnitude).  The two rules differ
    only where cancellation drives a pivot across zero by more than delta,
    near the end of ill-conditioned solves; there the reference's rule can
    overflow the factorization (ALMOST_SOLVED where the other rule reaches
    SOLVED).  Pass dynamic_regularization_keep_magnitude=0 for the
    reference's rule."""

    def __init__(self, mod):
        self._mod = mod

    def __getattr__(self, name):
        return getattr(self._mod, name)

    def solve(self, *args, **kw):
        kw.setdefault("dynamic_regularization_keep_magnitude", 1)
        return self._mod.solve(*args, **kw)


@pytest.fixture(scope="session")
def oracle(reference_oracle):
    return EngineRuleOracle(reference_oracle)
