"""The oracle against the committed golden vectors written by the reference
CPU backend (tests/golden/make_golden.py)."""
import glob
import os

import pytest

from oracle import harness as H

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_golden(oracle_built, path):
    p, expected = H.load_case(path)
    out = H.run_cpu(p, "oracle")
    assert not H.compare_to_fixture(out, expected)


def test_fixture_set_is_complete():
    names = {os.path.basename(p)[:-4] for p in FIXTURES}
    assert {"tm2d_o2_eonly", "3d_o4_s3_f32", "3d_o8_bf16_pec", "3d_o4_f16_faces", "2d_o8_coincident_s2",
            "3d_o4_physical_fwd"} <= names
