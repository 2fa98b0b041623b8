"""The CPU oracle against the golden fixtures from the reference's own CPU
backend: every frame of every case, bit exact (yuv16, NV12, RGB hashes) and
with the same slot assignment."""
import pytest

from golden_cases import case_id, compare, decode_hashes, load_cases
from oracle_lib import OracleDecoder

CASES = load_cases()


@pytest.mark.parametrize("case", CASES, ids=[case_id(c) for c in CASES])
def test_oracle_matches_reference(case):
    compare(case, decode_hashes(OracleDecoder(), case))
