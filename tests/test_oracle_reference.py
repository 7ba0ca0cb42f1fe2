"""The oracle restatement against the reference CPU backend compiled from its
own sources (oracle/_ref): bitwise equality on a case matrix, and the
reference's thread-count invariance (tests/test_cpu_determinism.py:43-65)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import harness as H

pytestmark = pytest.mark.skipif(not os.path.isdir("/root/reference/src/DeepGPR/lib") and
                                not os.path.isfile(H.REF_LIB), reason="reference sources not present")


@pytest.fixture(scope="module")
def built():
    H.build_oracle(reference=True)
    if not H.reference_available():
        pytest.skip("reference build unavailable")
    return H


MATRIX = [
    dict(shape=(14, 13, 12), order=o, S=S, kind=k, pml=(3, 2, 2, 3, 2, 3), d=(0.02, 0.017, 0.013))
    for o in (2, 4, 8) for S, k in ((1, "eonly"), (3, "f32"), (1, "bf16"), (2, "f16"))
] + [
    dict(shape=(20, 18, 2), order=o, S=S, kind=k, pml=(4, 3, 2, 3, 0, 0), d=(0.02, 0.02, 0.02))
    for o in (2, 4, 8) for S, k in ((1, "eonly"), (3, "f32"), (1, "f16"))
] + [
    dict(shape=(9, 10, 8), order=8, S=1, kind="f32", pml=(1, 2, 2, 1, 1, 1), d=(0.02, 0.017, 0.013)),
    dict(shape=(12, 11, 10), order=4, S=1, kind="f32", pml=(0, 0, 0, 0, 0, 0), d=(0.02, 0.02, 0.02)),
]


@pytest.mark.parametrize("case", MATRIX, ids=lambda c: f"{c['shape']}-o{c['order']}-S{c['S']}-{c['kind']}")
def test_oracle_bitwise_equals_reference(built, case):
    p = H.make_problem(*case["shape"], order=case["order"], nt=9, nshot=2, nsrc=2, nrx=3, seed=5,
                       pml=case["pml"], d=case["d"], pec=case["order"] == 8)
    p.S = case["S"]
    p.storage_type = {"eonly": H.RHS_FROM_E, "f32": 0, "bf16": H.KIND_BF16, "f16": H.KIND_F16}[case["kind"]]
    ref = H.run_cpu(p, "ref")
    ora = H.run_cpu(p, "oracle")
    assert not H.bitwise_diff(ref, ora, H.KEYS_FWD + H.KEYS_BWD)
    assert np.array_equal(ref["E_saved"], ora["E_saved"])
    assert np.array_equal(ref["R_saved"], ora["R_saved"])
    assert np.linalg.norm(ora["rx"]) > 0 and np.linalg.norm(ora["g_eps"]) > 0


def test_sources_receivers_in_every_component(built):
    for sc in range(3):
        for rc in range(3):
            p = H.make_problem(12, 11, 10, order=4, nt=6, nshot=1, nsrc=2, nrx=2, seed=20 + 3 * sc + rc)
            p.src_comp, p.rx_comp = sc, rc
            assert not H.bitwise_diff(H.run_cpu(p, "ref"), H.run_cpu(p, "oracle"), H.KEYS_FWD + H.KEYS_BWD)


def test_reference_is_thread_count_invariant(built):
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); from oracle import harness as H;"
        "p = H.make_problem(14, 12, 11, order=4, nt=8, nshot=2, seed=9);"
        "o = H.run_cpu(p, sys.argv[1]); np.save(sys.argv[2], np.concatenate([o['rx'].ravel(), o['g_eps'].ravel()]))"
    ) % os.path.dirname(H.HERE)
    outs = []
    for which in ("ref", "oracle"):
        for threads in (1, 3):
            path = f"/tmp/dgx_threads_{which}_{threads}.npy"
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            subprocess.run([sys.executable, "-c", code, which, path], check=True, env=env)
            outs.append(np.load(path))
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)
