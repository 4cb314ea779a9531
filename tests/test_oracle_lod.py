"""The oracle's Flip3.lod and Flip3.lodTarget held to test/particle/Flip3LodTest.java's cases."""
import pytest

import lod_cases as cases


@pytest.mark.parametrize("case", cases.CASES, ids=[c.__name__ for c in cases.CASES])
def test_lod_case_on_the_oracle(case):
    case(cases.OracleLodRig)


def test_lod_target_on_the_oracle():
    cases.case_lod_target(lambda slots: cases.OracleLodRig(cases.N, slots))
