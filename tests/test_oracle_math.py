"""Checks that pin the CPU oracle without the reference build: closed forms,
exact-transpose identities in FP64, Taylor remainders and codec definitions.
Each is chosen so that a plausible slip (a sign, a clamped radius, a missing
CPML term, a transposed index) breaks it."""
import ctypes
import math

import numpy as np
import pytest
import torch

from oracle import harness as H


def _lib(h):
    lib = h.cpu_lib("oracle")
    for pre, ct in (("oracle_f64_", ctypes.c_double), ("oracle_f32_", ctypes.c_float)):
        fn = getattr(lib, pre + "stencil_1d")
        fn.restype = ct
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        co = getattr(lib, pre + "stencil_coefficient")
        co.restype = ct
        co.argtypes = [ctypes.c_int, ctypes.c_int]
    return lib


def test_stencil_coefficients_are_the_staggered_taylor_weights(oracle_built):
    """c_r solve sum_r c_r (2r-1)^(2m+1) = delta_m0 (Fornberg staggered weights)."""
    lib = _lib(oracle_built)
    for R in (1, 2, 3, 4):
        c = np.array([lib.oracle_f64_stencil_coefficient(R, r) for r in range(1, R + 1)])
        A = np.array([[(2 * r - 1) ** (2 * m + 1) for r in range(1, R + 1)] for m in range(R)], float)
        b = np.zeros(R)
        b[0] = 1.0
        np.testing.assert_allclose(c, np.linalg.solve(A, b), rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("order", [2, 4, 8])
def test_clamped_differences_are_exact_on_polynomials(oracle_built, order):
    """A radius-R staggered difference is exact for polynomials of degree 2R-1;
    near a face the clamped radius must still differentiate a linear field."""
    lib = _lib(oracle_built)
    n = 17
    x = np.arange(n, dtype=np.float64)
    lin = np.ascontiguousarray(0.3 + 0.7 * x)
    for c in range(1, n):
        v = lib.oracle_f64_stencil_1d(lin.ctypes.data, n, c, order, 0)
        assert abs(v - 0.7) < 1e-12
    for c in range(0, n - 1):
        v = lib.oracle_f64_stencil_1d(lin.ctypes.data, n, c, order, 1)
        assert abs(v - 0.7) < 1e-12
    R = {2: 1, 4: 2, 8: 4}[order]
    cub = np.ascontiguousarray((x - 5.0) ** (2 * R - 1) / 100.0)
    c = 8  # interior for every order: derivative at x = c - 1/2 (backward) and c + 1/2 (forward)
    for kind, xm in ((0, c - 0.5), (1, c + 0.5)):
        v = lib.oracle_f64_stencil_1d(cub.ctypes.data, n, c, order, kind)
        exact = (2 * R - 1) * (xm - 5.0) ** (2 * R - 2) / 100.0
        assert abs(v - exact) < 1e-9 * max(1.0, abs(exact))


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("n", [2, 3, 6, 13])
def test_transposes_are_exact_adjoints_1d(oracle_built, order, n):
    """<D f, w> == <f, D^T w> over the coordinates where D is evaluated
    (restates src: tests/test_discrete_adjoint.py:337-351, 1e-13 in FP64)."""
    lib = _lib(oracle_built)
    rng = np.random.default_rng(n * 10 + order)
    for fwd, (dk, tk, coords) in enumerate(((0, 2, range(1, n)), (1, 3, range(0, n - 1)))):
        D = np.zeros((n, n))
        for col in range(n):
            e = np.zeros(n)
            e[col] = 1.0
            for c in coords:
                D[c, col] = lib.oracle_f64_stencil_1d(np.ascontiguousarray(e).ctypes.data, n, c, order, dk)
        w = np.zeros(n)
        w[list(coords)] = rng.standard_normal(len(list(coords)))
        w = np.ascontiguousarray(w)
        t = np.array([lib.oracle_f64_stencil_1d(w.ctypes.data, n, c, order, tk) for c in range(n)])
        np.testing.assert_allclose(t, D.T @ w, rtol=1e-13, atol=1e-13)


def _dot(a, b):
    return float(sum(np.vdot(np.asarray(x, np.float64).ravel(), np.asarray(y, np.float64).ravel())
                     for x, y in zip(a, b)))


@pytest.mark.parametrize("order,shape,pml", [
    (2, (11, 10, 9), (2, 3, 2, 2, 3, 2)),
    (4, (12, 11, 10), (3, 2, 2, 3, 2, 3)),
    (8, (13, 12, 11), (2, 2, 3, 2, 2, 2)),
    (4, (16, 14, 2), (3, 3, 2, 3, 0, 0)),
])
def test_adjoint_is_exact_transpose_of_the_state_map(oracle_built, order, shape, pml):
    """Dot-product test on the full linear map (E, H, 24 CPML, waveform) ->
    (E, H, 24 CPML, receivers) in FP64 (src: tests/test_discrete_adjoint.py:70-137)."""
    h = oracle_built
    p = h.make_problem(*shape, order=order, nt=7, nshot=2, nsrc=2, nrx=3, seed=order, pml=pml,
                       d=(0.02, 0.017, 0.013) if shape[2] > 2 else (0.02, 0.02, 0.02))
    p.storage_type, p.want_eps, p.want_sig = h.NO_HISTORY, 0, 0
    p.save_model_history = 0
    fwd = h.run_cpu(p, "oracle", real=np.float64, backward=False)
    # output duals: the lam0 states and data_grad
    out = fwd["E"] + fwd["H"] + fwd["phi"] + [fwd["rx"]]
    duals = list(p.lam0) + [p.data_grad]
    lhs = _dot(out, duals)
    bwd = h.run_cpu(p, "oracle", real=np.float64, backward=True)
    inp = list(p.E0) + list(p.H0) + list(p.phi0) + [p.wave]
    grads = bwd["L"] + [bwd["g_src"]]
    rhs = _dot(inp, grads)
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs), 1e-30), (lhs, rhs)


@pytest.mark.parametrize("storage", ["e_only", "stored"])
def test_material_gradient_matches_central_differences(oracle_built, storage):
    """Taylor check of d(sum rx*d)/d(eps, sigma) in FP64 along random directions."""
    h = oracle_built
    p = h.make_problem(13, 12, 11, order=4, nt=16, nshot=2, nsrc=1, nrx=5, seed=31, random_state=False,
                       pml=(2, 2, 2, 2, 2, 2))
    p.storage_type = h.RHS_FROM_E if storage == "e_only" else 0
    p.sig = p.sig * 500.0  # make the conductivity sensitivity visible
    core = (slice(0, p.NX - 1), slice(0, p.NY - 1), slice(0, p.NZ - 1))
    g = h.run_cpu(p, "oracle", real=np.float64)
    rng = np.random.default_rng(5)

    def J(eps, sig):
        q = h.with_storage(p, eps=eps, sig=sig)
        return float(np.sum(h.run_cpu(q, "oracle", real=np.float64, backward=False)["rx"] * p.data_grad))

    for name, key in (("eps", "g_eps"), ("sig", "g_sig")):
        base = getattr(p, name).astype(np.float64)
        delta = np.zeros_like(base)
        delta[core] = rng.standard_normal(delta[core].shape)
        step = 1e-4 * (np.abs(base[core]).max() + 1e-3)
        plus, minus = base + step * delta, base - step * delta
        if name == "eps":
            fd = (J(plus, p.sig) - J(minus, p.sig)) / (2 * step)
        else:
            fd = (J(p.eps, plus) - J(p.eps, minus)) / (2 * step)
        ad = float(np.sum(g[key] * delta[core]))
        assert abs(fd - ad) <= 1e-5 * abs(ad), (name, fd, ad)


def test_source_injection_known_answer(oracle_built):
    """One step from rest: E_c[src] = -cb w d_c / (dx dy dz) exactly; every
    receiver away from the source reads exactly 0 (tests/test_numerics.py:629-655)."""
    h = oracle_built
    for comp in range(3):
        p = h.make_problem(10, 11, 12, order=4, nt=1, nshot=1, nsrc=1, nrx=2, seed=2, random_state=False,
                           pml=(0, 0, 0, 0, 0, 0), d=(0.02, 0.017, 0.013))
        p.src_comp = p.rx_comp = comp
        p.src_loc[0, 0] = (4, 5, 6)
        p.rx_loc[0, 0] = (4, 5, 6)
        p.rx_loc[0, 1] = (7, 2, 3)
        p.wave[0, 0] = 0.75
        out = h.run_cpu(p, "oracle", real=np.float64, backward=False)
        cb = out["coeffs"][2][4, 5, 6]
        expect = -cb * (0.75 * p.d[comp] / (p.d[0] * p.d[1] * p.d[2]))
        assert out["rx"][0, 0, 0] == expect
        assert out["rx"][0, 0, 1] == 0.0


def test_update_coefficients_closed_form(oracle_built):
    h = oracle_built
    p = h.make_problem(9, 8, 7, order=2, nt=1, seed=4, pec=True)
    out = h.run_cpu(p, "oracle", real=np.float64, backward=False)
    ca, ccurl, cb, hh, hcurl, hrhs = out["coeffs"]
    core = (slice(0, p.NX - 1), slice(0, p.NY - 1), slice(0, p.NZ - 1))
    eps, sig, mu = (a.astype(np.float64)[core] for a in (p.eps, p.sig, p.mu))
    ea = 8.8541878128e-12 * eps / p.dt + 0.5 * sig
    eb = 8.8541878128e-12 * eps / p.dt - 0.5 * sig
    pec = sig > 100.0
    np.testing.assert_allclose(ca[core], np.where(pec, 0.0, eb / ea), rtol=1e-14)
    np.testing.assert_allclose(cb[core], np.where(pec, 0.0, 1.0 / ea), rtol=1e-14)
    np.testing.assert_allclose(ccurl[core], np.where(pec, 0.0, 1.0 / (p.d[0] * ea)), rtol=1e-14)
    ha = 1.25663706212e-06 * mu / p.dt
    np.testing.assert_allclose(hcurl[core], 1.0 / (p.d[0] * ha), rtol=1e-14)
    np.testing.assert_allclose(hrhs[core], 1.0 / ha, rtol=1e-14)
    assert np.all(hh[core] == 1.0)
    halo = np.ones(ca.shape, bool)
    halo[core] = False
    assert np.all(ca[halo] == 0) and np.all(hcurl[halo] == 0)


def test_half_and_bfloat16_encoders_round_to_nearest_even(oracle_built):
    """Against numpy's binary16 and torch's bfloat16 conversions (both IEEE RNE)."""
    lib = oracle_built.cpu_lib("oracle")
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2**32, size=1 << 18, dtype=np.uint64).astype(np.uint32)
    # dense coverage of the binary16 range, its subnormals and the overflow edge
    dense = np.arange(0x33000000, 0x47800400, 97, dtype=np.uint32)
    vals = np.concatenate([bits.view(np.float32), dense.view(np.float32), -dense.view(np.float32),
                           np.array([65504, 65519.996, 65520, 2.0 ** -25, 2.0 ** -24 * 1.5], np.float32)])
    vals = np.ascontiguousarray(vals[np.isfinite(vals)])
    enc = np.zeros(vals.size, np.uint16)
    dec = np.zeros(vals.size, np.float32)
    lib.oracle_f32_codec_roundtrip(vals.ctypes.data_as(ctypes.c_void_p), enc.ctypes.data_as(ctypes.c_void_p),
                                   dec.ctypes.data_as(ctypes.c_void_p), ctypes.c_longlong(vals.size), 1)
    with np.errstate(over="ignore"):
        ref16 = vals.astype(np.float16)
    assert np.array_equal(enc, ref16.view(np.uint16))
    assert np.array_equal(dec, ref16.astype(np.float32))
    lib.oracle_f32_codec_roundtrip(vals.ctypes.data_as(ctypes.c_void_p), enc.ctypes.data_as(ctypes.c_void_p),
                                   dec.ctypes.data_as(ctypes.c_void_p), ctypes.c_longlong(vals.size), 2)
    refb = torch.from_numpy(vals).to(torch.bfloat16)
    assert np.array_equal(enc, refb.view(torch.int16).numpy().view(np.uint16))
    assert np.array_equal(dec, refb.to(torch.float32).numpy())


@pytest.mark.parametrize("shape,block", [((14, 13, 12), (4, 4, 4)), ((20, 18, 2), (8, 8, 1)),
                                         ((15, 12, 11), (2, 4, 2))])
def test_int8_history_bounds_and_layout(oracle_built, shape, block):
    """INT8 E/R frames decode (with the KAT-pinned decoder) to within half a
    quantisation step of the exact FP32 frames; scales are max|v|/127."""
    from deepgpr_amd.solver.storage import decompress_wavefield_history, saved_history_shape
    h = oracle_built
    p = h.make_problem(*shape, order=4, nt=7, nshot=2, nsrc=1, nrx=2, seed=8)
    p.S = 2
    ref = h.run_cpu(p, "oracle", backward=False)
    p.storage_type = h.int8_storage_type(block)
    q = h.run_cpu(p, "oracle", backward=False)
    assert not h.bitwise_diff(ref, q, h.KEYS_FWD)  # the codec never touches the fields
    (x0, nx), (y0, ny), (z0, nz) = p.box()
    hist = saved_history_shape(3, p.nt_saved, p.nshot, nx, ny, nz)
    for key in ("E_saved", "R_saved"):
        exact = ref[key].view(np.float32).reshape(hist)
        dec = decompress_wavefield_history(torch.from_numpy(q[key].view(np.int8)), hist, block).numpy()
        bx, by, bz = block
        pad = [(0, (-s) % b) for s, b in zip(exact.shape[-3:], block)]
        ex = np.pad(np.abs(exact), [(0, 0)] * 3 + pad)
        tiles = ex.reshape(*ex.shape[:3], ex.shape[3] // bx, bx, ex.shape[4] // by, by, ex.shape[5] // bz, bz)
        peak = tiles.max(axis=(4, 6, 8))
        scale = np.where(peak >= 1e-30, peak / 127.0, 1.0).astype(np.float32)
        full = np.repeat(np.repeat(np.repeat(scale, bx, 3), by, 4), bz, 5)[..., :nx, :ny, :nz]
        assert np.all(np.abs(dec - exact) <= 0.5 * full * (1 + 1e-6) + 1e-38)
