"""The block -> cell maps of the one-thread-per-cell 3D sweeps against the CPU
oracle, bit for bit, on grids large enough for an interior box: the default
map (interior box in 256-cell chunks, interleaved with the CPML shell),
without a box (DEEPGPR_AMD_3D_BOX=off) and without interleaving
(DEEPGPR_AMD_3D_INTERLEAVE=off); orders 2/4/8, every history format,
anisotropic spacing, no CPML, random initial states, and sources and
receivers in the z-CPML and z-edge cells beside the box, on plain box cells
and in the shell."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BASE = (60, 56, 70, 4, 1, "f32", (5, 4, 3, 6, 4, 5))
CASES = [
    # (NX, NY, NZ, order, S, storage, pml)
    (60, 56, 70, 2, 1, "e_only", (5, 4, 3, 6, 4, 5)),
    (60, 56, 70, 8, 2, "bf16", (5, 4, 3, 6, 4, 5)),
    (60, 56, 70, 4, 1, "int8", (5, 4, 3, 6, 4, 5)),
    (56, 52, 70, 4, 3, "f32", (0, 0, 0, 0, 0, 0)),  # the z-edge cells are the clamped-difference ones
    (60, 40, 270, 4, 1, "f32", (4, 5, 3, 4, 6, 5)),  # rows longer than a workgroup
]
MODES = {
    "default": {},
    "no_box": {"DEEPGPR_AMD_3D_BOX": "off"},
    "no_interleave": {"DEEPGPR_AMD_3D_INTERLEAVE": "off"},
}


def _storage(h, name):
    return {"f32": 0, "e_only": h.RHS_FROM_E, "bf16": h.KIND_BF16, "int8": h.int8_storage_type((4, 4, 4))}[name]


def _check(h, case):
    NX, NY, NZ, order, S, st, pml = case
    p = h.make_problem(NX, NY, NZ, order=order, nt=10, nshot=2, nsrc=3, nrx=4, seed=11, pml=pml,
                       d=(0.02, 0.017, 0.015))
    p.S = S
    p.storage_type = _storage(h, st)
    # sources and receivers in the z-CPML and z-edge cells of rows that cross
    # the box (low and high z), on a plain box cell and in the shell
    p.src_loc[0, 0] = [NX // 2, NY // 2, 2]
    p.src_loc[0, 1] = [NX // 2 + 1, NY // 2 - 1, NZ - 3]
    p.src_loc[1, 2] = [NX // 2, NY // 3, NZ // 2]
    p.rx_loc[0, 0] = [NX // 2 - 1, NY // 2, 1]
    p.rx_loc[1, 1] = [NX // 2, NY // 2 + 2, NZ - 2]
    p.rx_loc[1, 3] = [1, NY // 2, NZ // 2]
    cpu = h.run_cpu(p, "oracle")
    gpu = h.run_gpu(p, api="dgx")
    bad = h.bitwise_diff(gpu, cpu, h.KEYS_FWD + h.KEYS_BWD)
    assert not bad, bad
    assert np.array_equal(gpu["E_saved"], cpu["E_saved"])
    assert np.array_equal(gpu["R_saved"], cpu["R_saved"])


@pytest.mark.parametrize("mode", sorted(MODES))
def test_sweep_maps_match_oracle_bitwise(oracle_built, cuda_device, monkeypatch, mode):
    for key, value in MODES[mode].items():
        monkeypatch.setenv(key, value)
    print(f"[gpu sweep maps] {BASE} {mode}", flush=True)
    _check(oracle_built, BASE)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}-o{c[3]}-S{c[4]}-{c[5]}")
def test_box_grids_match_oracle_bitwise(oracle_built, cuda_device, case):
    print(f"[gpu sweep maps] {case}", flush=True)
    _check(oracle_built, case)
