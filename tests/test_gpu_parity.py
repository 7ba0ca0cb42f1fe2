"""HIP kernels vs the CPU oracle (float32 restatement of the reference CPU
backend, itself bitwise-checked against the reference build). With
-ffp-contract=off and the reference operation order the HIP path is expected
to match bit for bit; every case asserts that, and separately the stated
floating-point tolerances (receiver traces rel-L2 <= 1e-5, gradients <= 1e-4)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # (NX, NY, NZ, order, S, storage, pml)
    (14, 13, 12, 2, 1, "e_only", (3, 2, 2, 3, 2, 3)),
    (14, 13, 12, 4, 3, "f32", (3, 2, 2, 3, 2, 3)),
    (14, 13, 12, 8, 1, "bf16", (3, 2, 2, 3, 2, 3)),
    (15, 12, 13, 4, 1, "f16", (0, 3, 2, 0, 3, 2)),
    (20, 18, 2, 2, 1, "e_only", (3, 3, 3, 3, 0, 0)),
    (20, 18, 2, 4, 2, "f32", (4, 2, 3, 3, 0, 0)),
    (22, 19, 2, 8, 1, "bf16", (3, 3, 3, 3, 0, 0)),
]


def _storage(h, name):
    return {"f32": 0, "e_only": h.RHS_FROM_E, "f16": h.KIND_F16, "bf16": h.KIND_BF16}[name]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}-o{c[3]}-S{c[4]}-{c[5]}")
@pytest.mark.parametrize("api", ["abi6", "dgx"])
def test_gpu_matches_oracle_bitwise(oracle_built, cuda_device, case, api):
    h = oracle_built
    NX, NY, NZ, order, S, st, pml = case
    print(f"[gpu parity] {case} api={api}", flush=True)
    p = h.make_problem(NX, NY, NZ, order=order, nt=11, nshot=2, nsrc=2, nrx=4, seed=7, pml=pml,
                       d=(0.02, 0.017, 0.013) if NZ > 2 else (0.02, 0.02, 0.02))
    p.S = S
    p.storage_type = _storage(h, st)
    cpu = h.run_cpu(p, "oracle")
    gpu = h.run_gpu(p, api=api)
    for k in ("rx", "g_eps", "g_sig", "g_src"):
        err = h.rel_l2(gpu[k], cpu[k])
        assert err <= (1e-5 if k == "rx" else 1e-4), (k, err)
    bad = h.bitwise_diff(gpu, cpu, h.KEYS_FWD + h.KEYS_BWD)
    assert not bad, bad
    assert np.array_equal(gpu["E_saved"], cpu["E_saved"])
    assert np.array_equal(gpu["R_saved"], cpu["R_saved"])


@pytest.mark.parametrize("order,rx_comp", [(4, 2), (8, 0), (2, 1)])
def test_gpu_adjoint_receivers_match_oracle(oracle_built, cuda_device, order, rx_comp):
    """Receivers^T of the 3D adjoint: coincident receivers, receivers in x- and
    z-CPML, and receivers in both launches of a split cell sweep (z long
    enough), for each E component."""
    h = oracle_built
    print(f"[gpu parity] receivers order={order} comp={rx_comp}", flush=True)
    p = h.make_problem(12, 11, 74, order=order, nt=9, nshot=2, nsrc=1, nrx=5, seed=11, pml=(3, 2, 3, 2, 4, 3),
                       d=(0.02, 0.017, 0.013))
    p.rx_comp = rx_comp
    for s in range(2):
        p.rx_loc[s, 1] = p.rx_loc[s, 0]
        p.rx_loc[s, 2] = [5, 5, 2 + s]
        p.rx_loc[s, 3] = [6, 4, 40]
        p.rx_loc[s, 4] = [1, 9, 40]
    cpu = h.run_cpu(p, "oracle")
    gpu = h.run_gpu(p, api="dgx")
    assert np.abs(cpu["g_eps"]).max() > 0
    bad = h.bitwise_diff(gpu, cpu, h.KEYS_FWD + h.KEYS_BWD)
    assert not bad, bad


@pytest.mark.parametrize("order", [2, 4, 8])
def test_gpu_sources_in_both_parts_of_a_sweep(oracle_built, cuda_device, order):
    """z long enough that each cell sweep runs as a z-interior part and a
    z-CPML part in one launch: sources in the low and high z-CPML and in the
    z-interior, each found through the workgroup numbering of its part."""
    h = oracle_built
    print(f"[gpu parity] sweep parts order={order}", flush=True)
    p = h.make_problem(12, 11, 80, order=order, nt=9, nshot=2, nsrc=3, nrx=3, seed=21, pml=(3, 2, 3, 2, 4, 3),
                       d=(0.02, 0.017, 0.013))
    for s in range(2):
        p.src_loc[s, 0] = [5 + s, 5, 1 + s]
        p.src_loc[s, 1] = [6, 4 + s, 77 - s]
        p.src_loc[s, 2] = [4, 6, 40 + s]
        p.rx_loc[s, 0] = [6, 6, 3]
        p.rx_loc[s, 1] = [5, 5, 76]
        p.rx_loc[s, 2] = [7, 5, 41]
    cpu = h.run_cpu(p, "oracle")
    gpu = h.run_gpu(p, api="dgx")
    assert np.abs(cpu["rx"]).max() > 0
    bad = h.bitwise_diff(gpu, cpu, h.KEYS_FWD + h.KEYS_BWD)
    assert not bad, bad


def test_gpu_int8_history_matches_oracle(oracle_built, cuda_device):
    h = oracle_built
    # 64-cell tiles (one wavefront per tile) and a 16-cell tile (the general kernel)
    for NX, NY, NZ, block in ((14, 13, 12, (4, 4, 4)), (20, 18, 2, (8, 8, 1)), (14, 13, 12, (2, 4, 2))):
        p = h.make_problem(NX, NY, NZ, order=4, nt=9, nshot=2, nsrc=1, nrx=3, seed=3)
        p.storage_type = h.int8_storage_type(block)
        for e_only in (False, True):
            if e_only:
                p.storage_type |= h.RHS_FROM_E
            cpu = h.run_cpu(p, "oracle")
            gpu = h.run_gpu(p, api="dgx")
            assert np.array_equal(gpu["E_saved"], cpu["E_saved"])
            assert np.array_equal(gpu["R_saved"], cpu["R_saved"])
            assert not h.bitwise_diff(gpu, cpu, h.KEYS_FWD + h.KEYS_BWD)


def test_gpu_fp64_accumulators_close(oracle_built, cuda_device):
    h = oracle_built
    p = h.make_problem(16, 14, 13, order=4, nt=20, nshot=2, nsrc=1, nrx=4, seed=11)
    p.storage_type = h.RHS_FROM_E
    ref64 = h.run_cpu(p, "oracle", real=np.float64)
    g32 = h.run_gpu(p, api="dgx", fp64=False)
    g64 = h.run_gpu(p, api="dgx", fp64=True)
    for k in ("g_eps", "g_sig"):
        assert h.rel_l2(g64[k], ref64[k]) <= h.rel_l2(g32[k], ref64[k]) * 1.5 + 1e-7
        assert h.rel_l2(g64[k], ref64[k]) < 1e-5


def test_gpu_codec_backends_agree(cuda_device):
    import ctypes
    import torch
    from deepgpr_amd.native.loader import get_library
    lib = get_library()
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2**32, size=1 << 20, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    extra = np.array([0.0, -0.0, 65504.0, 65520.0, 65519.99, 6e-8, 3e-8, 5.96e-8, 1e-30, np.inf, -np.inf],
                     np.float32)
    vals = np.concatenate([vals, extra])
    x = torch.from_numpy(vals).to(cuda_device)
    for kind, npdt in ((1, np.float16), (2, None)):
        out = {}
        for backend in (0, 1):
            enc = torch.zeros(x.numel(), dtype=torch.int16, device=cuda_device)
            dec = torch.zeros_like(x)
            fp = ctypes.POINTER(ctypes.c_float)
            rc = lib.dgx_codec_roundtrip(ctypes.cast(x.data_ptr(), fp), ctypes.c_void_p(enc.data_ptr()),
                                         ctypes.cast(dec.data_ptr(), fp), x.numel(), kind, backend, None)
            assert rc == 0
            out[backend] = enc.cpu().numpy().view(np.uint16)
        finite = np.isfinite(vals)
        assert np.array_equal(out[0][finite], out[1][finite])
        if npdt is not None:
            assert np.array_equal(out[0][finite], vals[finite].astype(np.float16).view(np.uint16))
        else:
            ref = torch.from_numpy(vals[finite]).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
            assert np.array_equal(out[0][finite], ref)


TM2D = 1 << 6
INACTIVE_PHI = (0, 3, 4, 7, 8, 11, 12, 15)


def _tm2d_problem(h, NX, NY, order, seed, nsrc=2, nrx=4):
    p = h.make_problem(NX, NY, 2, order=order, nt=13, nshot=2, nsrc=nsrc, nrx=nrx, seed=seed,
                       pml=(4, 3, 3, 4, 0, 0))
    # the 2D TM contract: Ex, Ey, Hz and their CPML auxiliaries start at zero
    for arrs, idx in ((p.E0, (0, 1)), (p.H0, (2,)), (p.lam0, (0, 1, 5))):
        for k in idx:
            arrs[k] = np.zeros_like(arrs[k])
    for k in INACTIVE_PHI:
        p.phi0[k] = np.zeros_like(p.phi0[k])
        p.lam0[6 + k] = np.zeros_like(p.lam0[6 + k])
    return p


# tile rows of the fused step kernels: 4 and 8 (chosen for small solves) and 16
TILE_ROWS = ["4", "8", "16"]


@pytest.mark.parametrize("rows", TILE_ROWS)
@pytest.mark.parametrize("order", [2, 4, 8])
def test_tm2d_large_grid_matches_oracle(oracle_built, cuda_device, monkeypatch, order, rows):
    """A grid with tiles whose staged region (tile and halo) avoids the CPML
    and the grid faces, with sources and receivers inside such tiles and on
    tile edges."""
    h = oracle_built
    monkeypatch.setenv("DEEPGPR_AMD_TM2D_TILE_ROWS", rows)
    print(f"[tm2d large] order={order} rows={rows}", flush=True)
    p = _tm2d_problem(h, 90, 210, order, seed=60 + order, nsrc=3, nrx=6)
    for s in range(2):
        p.src_loc[s, 0] = [20 + 5 * s, 70, 0]
        p.src_loc[s, 1] = [47, 127 + s, 0]
        p.rx_loc[s, 0] = [33, 100, 0]
        p.rx_loc[s, 1] = [48, 128, 0]
        p.rx_loc[s, 2] = [64, 64 + s, 0]
    p.mode, p.S = 3, 1
    p.storage_type = _storage(h, "e_only")
    cpu = h.run_cpu(p, "oracle")
    p.storage_type |= TM2D
    gpu = h.run_gpu(p, api="dgx")
    bad = h.bitwise_diff(gpu, cpu, h.KEYS_FWD + ("g_src", "g_eps", "g_sig"))
    assert not bad, bad
    active = [2, 3, 4] + [6 + k for k in range(24) if k not in INACTIVE_PHI]
    for k in active:
        assert np.array_equal(gpu["L"][k], cpu["L"][k]), k
    assert np.array_equal(gpu["E_saved"], cpu["E_saved"])


@pytest.mark.parametrize("rows", TILE_ROWS)
@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("mode,S,st", [(2, 1, "e_only"), (3, 1, "e_only"), (3, 2, "f32"), (3, 1, "bf16")])
def test_tm2d_fast_path_matches_general_oracle(oracle_built, cuda_device, monkeypatch, order, mode, S, st, rows):
    h = oracle_built
    monkeypatch.setenv("DEEPGPR_AMD_TM2D_TILE_ROWS", rows)
    print(f"[tm2d] order={order} mode={mode} S={S} {st} rows={rows}", flush=True)
    p = _tm2d_problem(h, 26, 23, order, seed=40 + order)
    p.mode, p.S = mode, S
    p.storage_type = _storage(h, st)
    cpu = h.run_cpu(p, "oracle")
    p.storage_type |= TM2D
    gpu = h.run_gpu(p, api="dgx")
    bad = h.bitwise_diff(gpu, cpu, h.KEYS_FWD + ("g_src", "g_eps", "g_sig"))
    assert not bad, bad
    # adjoint states: the TM2D contract transposes only Ez, Hx, Hy and their
    # CPML auxiliaries and leaves the other cotangents untouched (deepgpr.h:77-83)
    active = [2, 3, 4] + [6 + k for k in range(24) if k not in INACTIVE_PHI]
    for k in active:
        assert np.array_equal(gpu["L"][k], cpu["L"][k]), k
    for k in set(range(30)) - set(active):
        assert np.array_equal(gpu["L"][k], p.lam0[k]), k
    assert np.array_equal(gpu["E_saved"], cpu["E_saved"])
    assert np.array_equal(gpu["R_saved"], cpu["R_saved"])
