"""The OpenMP C restatement of the fused right-hand sides (the CPU baseline
bench.py times) against the numpy oracle, every equation at P1-P4, with
boundaries, and at several thread counts."""

import numpy as np
import pytest

from oracle import dg_cpu
from tests._helpers import discr_pair, field_errors, oracle_rhs, random_state

PARAM = {"wave": 1.3, "euler": 1.4, "maxwell": 0.5}


@pytest.mark.parametrize("equation", ["wave", "euler", "maxwell"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_c_restatement_matches_the_oracle(equation, order):
    d, od = discr_pair(order, 3)
    q = random_state(equation, od.x, seed=order + 7)
    got = dg_cpu.CpuOperator(d, equation, PARAM[equation])(q)
    ref = oracle_rhs(equation, od, q, PARAM[equation])
    assert max(field_errors(got, ref)) < 1e-12


def test_c_restatement_is_thread_count_independent():
    d, od = discr_pair(2, 4)
    q = random_state("euler", od.x, seed=1)
    a = dg_cpu.CpuOperator(d, "euler", 1.4, threads=1)(q)
    b = dg_cpu.CpuOperator(d, "euler", 1.4, threads=4)(q)
    assert np.array_equal(a, b)
