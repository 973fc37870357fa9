"""Every golden vector of the reference's own kernel tests, replayed on the
GPU through libhs_amd.so (the reference checks each at 1e-12)."""
import pytest

from backends import HipBackend
from golden_cases import case_id, check_case, kernel_cases, run_case


@pytest.mark.gpu
@pytest.mark.parametrize("case", kernel_cases(), ids=case_id)
def test_gpu_reference_golden(case, hs, dev):
    print("golden case", case_id(case), flush=True)
    final, env = run_case(case, HipBackend(hs, dev))
    fails = check_case(case, final, env)
    assert not fails, "%s (%s): %s" % (case_id(case), case["source"], fails)
