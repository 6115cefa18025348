"""Every test case of the settings-file driver's table (cfdart_test_case) run for 10 steps through
cfdart_run on the GPU, bitwise equal to the oracle from the same initial state, with the same
fallback counters (exact arithmetic)."""
import math

import numpy as np
import pytest

import cfdart
import oracle
from test_test_cases import ALLAIRE_CASES, EULER_CASES
from util import bitwise_equal, interior

pytestmark = pytest.mark.gpu

CASES = [("allaire", n) for n in ALLAIRE_CASES] + [("euler", n) for n in EULER_CASES]


@pytest.mark.parametrize("problem,name", CASES, ids=[f"{p}-{n}" for p, n in CASES])
def test_case_10_steps_equals_oracle(gpu, problem, name):
    d, h, _ = cfdart.test_case(problem, name, 48, 40)
    splitting = cfdart.NAIVE if len(name) % 2 else cfdart.ALTERNATING
    cfl = 0.5
    print(f"case {problem} {name} 48x40 splitting={splitting}", flush=True)
    g = cfdart.Grid(d)
    g.upload(h)
    t, n = g.run(splitting, 10, T=math.inf, cfl=cfl)
    got = g.download()
    cnt = g.stats()
    g.close()
    s = oracle.Sim(d, h)
    k, _ = s.run(splitting, 10, cfl=cfl)
    assert n == k == 10 and t == s.t
    assert bitwise_equal(interior(got), interior(s.grid))
    assert cnt == s.counters()
