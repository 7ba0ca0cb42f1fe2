"""The x-marching 3D sweeps (DEEPGPR_AMD_3D_SWEEP=xmarch) against the CPU
oracle, bit for bit: orders 2/4/8, asymmetric and zero CPML faces, anisotropic
spacing, every history format, sources and receivers in the CPML, tiles of 4
and 8 rows, and x-chunks short enough that every chunk boundary and the
deferred lambda psi_H update are crossed many times."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # (NX, NY, NZ, order, S, storage, pml, rows, planes)
    (14, 13, 12, 2, 1, "e_only", (3, 2, 2, 3, 2, 3), 8, 16),
    (14, 13, 12, 4, 3, "f32", (3, 2, 2, 3, 2, 3), 8, 3),
    (14, 13, 12, 4, 1, "f32", (3, 2, 2, 3, 2, 3), 4, 5),
    (14, 13, 12, 8, 1, "bf16", (3, 2, 2, 3, 2, 3), 8, 4),
    (15, 12, 13, 4, 1, "f16", (0, 3, 2, 0, 3, 2), 4, 2),
    (21, 70, 75, 4, 2, "f32", (4, 3, 5, 4, 6, 5), 8, 7),
    (19, 11, 130, 8, 1, "e_only", (4, 4, 3, 3, 5, 4), 8, 6),
    (17, 20, 66, 2, 1, "f32", (0, 0, 0, 0, 0, 0), 4, 9),
]


def _storage(h, name):
    return {"f32": 0, "e_only": h.RHS_FROM_E, "f16": h.KIND_F16, "bf16": h.KIND_BF16,
            "int8": h.int8_storage_type((4, 4, 4))}[name]


@pytest.fixture
def xmarch(monkeypatch):
    monkeypatch.setenv("DEEPGPR_AMD_3D_SWEEP", "xmarch")
    return monkeypatch


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}-o{c[3]}-S{c[4]}-{c[5]}-r{c[7]}-p{c[8]}")
def test_xmarch_matches_oracle_bitwise(oracle_built, cuda_device, xmarch, case):
    h = oracle_built
    NX, NY, NZ, order, S, st, pml, rows, planes = case
    xmarch.setenv("DEEPGPR_AMD_XM_TILE_ROWS", str(rows))
    xmarch.setenv("DEEPGPR_AMD_XM_PLANES", str(planes))
    print(f"[gpu xmarch] {case}", flush=True)
    p = h.make_problem(NX, NY, NZ, order=order, nt=11, nshot=2, nsrc=2, nrx=4, seed=7, pml=pml,
                       d=(0.02, 0.017, 0.013))
    p.S = S
    p.storage_type = _storage(h, st)
    # a source and a receiver inside the CPML of the low x and high z faces
    p.src_loc[0, 1] = [1, NY // 2, NZ - 2]
    p.rx_loc[1, 2] = [NX - 2, 1, NZ // 2]
    cpu = h.run_cpu(p, "oracle")
    gpu = h.run_gpu(p, api="dgx")
    bad = h.bitwise_diff(gpu, cpu, h.KEYS_FWD + h.KEYS_BWD)
    assert not bad, bad
    assert np.array_equal(gpu["E_saved"], cpu["E_saved"])
    assert np.array_equal(gpu["R_saved"], cpu["R_saved"])


def test_xmarch_int8_history_matches_oracle(oracle_built, cuda_device, xmarch):
    h = oracle_built
    xmarch.setenv("DEEPGPR_AMD_XM_PLANES", "4")
    print("[gpu xmarch] int8", flush=True)
    p = h.make_problem(16, 18, 20, order=4, nt=9, nshot=2, nsrc=1, nrx=3, seed=3, pml=(3, 3, 2, 4, 3, 2),
                       d=(0.02, 0.02, 0.02))
    p.storage_type = _storage(h, "int8")
    cpu = h.run_cpu(p, "oracle")
    gpu = h.run_gpu(p, api="dgx")
    bad = h.bitwise_diff(gpu, cpu, h.KEYS_FWD + h.KEYS_BWD)
    assert not bad, bad
