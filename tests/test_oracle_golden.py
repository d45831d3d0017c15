"""The CPU oracle against the reference test-suite's own expectations (test/test.cpp).

These pin the oracle before it is trusted as the parity checker: every case of
tests/golden/reference_expectations.json must hold within EXPECT_FLOAT_EQ's
4-ULP bound in the float build (the reference's default), and within 1e-5
relative in the double build (the -DDOUBLE_PRECISION build, whose own results
are closer to the exact values the reference quotes).
"""
import numpy as np
import pytest

from golden_runner import float_ulp_diff, load_cases, run_case

FIXTURES, CASES = load_cases()


class OracleBackend:
    def __init__(self, oracle_mod, dtype, *args):
        self.o = oracle_mod.Oracle(*args, dtype=dtype)
        self.grid = self.o.grid

    @property
    def densities(self):
        return self.o.densities

    @property
    def gradients(self):
        return self.o.gradients

    @property
    def internalized(self):
        return self.o.internalized

    def __getattr__(self, name):
        return getattr(self.o, name)

    def store(self):
        pass

    def load(self):
        pass


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_float_matches_reference_expectations(case, oracle_mod):
    make = lambda *a: OracleBackend(oracle_mod, np.float32, *a)
    n = 0
    for label, got, exp in run_case(case, FIXTURES, make):
        ulps = float_ulp_diff(got, exp)
        bad = np.nonzero(ulps > 4)[0]
        assert bad.size == 0, f"{case['name']} {label}: {bad.size} values beyond 4 ULP, e.g. idx {bad[:5]} " \
                              f"got {np.asarray(got)[bad[:5]]} expected {np.asarray(exp)[bad[:5]]}"
        n += 1
    assert n > 0


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_double_matches_reference_expectations(case, oracle_mod):
    make = lambda *a: OracleBackend(oracle_mod, np.float64, *a)
    for label, got, exp in run_case(case, FIXTURES, make):
        got = np.asarray(got, dtype=np.float64)
        exp = np.asarray(exp, dtype=np.float64)
        rel = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-30)
        assert np.all(rel <= 1e-5), f"{case['name']} {label}: max rel {rel.max():.3e}"
