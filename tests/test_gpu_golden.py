"""The HIP backend, through the C ABI, against the golden fixtures from the
reference's own CPU backend: bit exact yuv16, device NV12 and device RGB on
every frame of every case."""
import pytest

import av1ongpu_amd as pkg
from golden_cases import case_id, compare, decode_hashes, load_cases

pytestmark = pytest.mark.gpu
CASES = load_cases()


@pytest.mark.parametrize("case", CASES, ids=[case_id(c) for c in CASES])
def test_hip_matches_reference(case):
    compare(case, decode_hashes(pkg.HipDecoder(), case))
