"""dgx_forward_noalloc (the forward that allocates no device memory) against
dgx_forward through the C ABI: every forward output bit for bit, on two calls
in a row (the second smaller, so it reuses the scratch array the first left
behind), and the refusals it documents."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(h, shape, shots, seed, storage):
    p = h.make_problem(*shape, order=4, nt=12, nshot=shots, nsrc=2, nrx=3, seed=seed, pml=(3, 2, 3, 2, 2, 3),
                       d=(0.02, 0.017, 0.013))
    p.storage_type = storage
    p.S = 2
    return p


def test_noalloc_forward_equals_dgx_forward(oracle_built, cuda_device):
    h = oracle_built
    from deepgpr_amd.native.loader import get_library
    assert get_library().dgx_forward_noalloc_scratch_bytes() == 4 << 20
    cases = [((22, 19, 70), 3, 21, 0), ((12, 11, 10), 1, 22, h.KIND_BF16)]  # second call smaller
    for shape, shots, seed, storage in cases:
        print(f"[no alloc] {shape} x {shots} shots", flush=True)
        p = _case(h, shape, shots, seed, storage)
        ref = h.run_gpu(p, backward=False, api="dgx")
        got = h.run_gpu(p, backward=False, api="dgx_noalloc")
        bad = h.bitwise_diff(got, ref, h.KEYS_FWD)
        assert not bad, bad
        assert np.array_equal(got["E_saved"], ref["E_saved"])
        assert np.array_equal(got["R_saved"], ref["R_saved"])
        # and the same as the oracle
        cpu = h.run_cpu(p, "oracle", backward=False)
        assert not h.bitwise_diff(got, cpu, h.KEYS_FWD)


def test_noalloc_refuses_int8_histories(oracle_built, cuda_device):
    h = oracle_built
    from deepgpr_amd.utils.exceptions import NativeLibraryError
    p = _case(h, (12, 11, 10), 1, 23, h.int8_storage_type((4, 4, 4)))
    with pytest.raises(NativeLibraryError, match="INT8"):
        h.run_gpu(p, backward=False, api="dgx_noalloc")
