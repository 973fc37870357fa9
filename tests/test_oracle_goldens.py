"""The oracle against every golden vector of the reference's own test suite
(tests/test*.cpp of the reference, extracted to tests/golden/), at the
reference's tolerance."""
import pytest

from backends import OracleBackend
from golden_cases import case_id, check_case, kernel_cases, parser_cases, run_case, run_parser_case


@pytest.mark.parametrize("case", kernel_cases(), ids=case_id)
def test_oracle_reference_golden(case, oracle):
    final, env = run_case(case, OracleBackend(oracle))
    fails = check_case(case, final, env)
    assert not fails, "%s (%s): %s" % (case_id(case), case["source"], fails)


@pytest.mark.parametrize("case", parser_cases(), ids=case_id)
def test_layout_and_parser_golden(case):
    fails = run_parser_case(case)
    assert not fails, "%s (%s): %s" % (case_id(case), case["source"], fails)
