"""Harness around the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Builds fully specified FDTD problems (numpy) and runs them through
  * the reference CPU backend compiled from source (oracle/_ref, optional),
  * the oracle restatement (oracle/liboracle.so, float32 and float64),
  * the HIP library (deepgpr_amd/libdeepgpr_amd.so, through its ABI-6 exports
    or the dgx_* API) when a GPU is present,
all with the same positional ABI-6 argument list
(src/DeepGPR/lib/deepgpr.h:89-152). Only tests, the smoke check and the
benchmark's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field, replace
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libdeepgpr_ref.so")

EPS0 = 8.8541878128e-12
C0 = 299792458.0
FACES = ("x0", "xm", "y0", "ym", "z0", "zm")
KIND_F32, KIND_F16, KIND_BF16, KIND_INT8 = 0, 1, 2, 3
RHS_FROM_E, PHYSICAL, NO_HISTORY = 1 << 4, 1 << 5, 1 << 26


def build_oracle(reference: bool = True) -> None:
    """Compile the oracle (and the reference when its sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if reference and os.path.isdir("/root/reference/src/DeepGPR/lib"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def int8_storage_type(block) -> int:
    bx, by, bz = block
    return KIND_INT8 | (bx << 8) | (by << 14) | (bz << 20)


@dataclass
class Problem:
    """Every input of one forward + adjoint pair, on the extended grid."""

    NX: int
    NY: int
    NZ: int
    nshot: int
    nt: int
    order: int
    pml: List[int]
    d: tuple
    dt: float
    eps: np.ndarray
    sig: np.ndarray
    mu: np.ndarray
    coef_e: List[np.ndarray]
    coef_h: List[np.ndarray]
    E0: List[np.ndarray]
    H0: List[np.ndarray]
    phi0: List[np.ndarray]
    src_loc: np.ndarray
    wave: np.ndarray
    rx_loc: np.ndarray
    src_comp: int = 2
    rx_comp: int = 2
    S: int = 1
    mode: int = 3
    storage_type: int = 0
    save_model_history: int = 1
    data_grad: Optional[np.ndarray] = None
    lam0: Optional[List[np.ndarray]] = None
    want_src: int = 1
    want_eps: int = 1
    want_sig: int = 1
    meta: Dict = field(default_factory=dict)

    # ---------------------------------------------------------------- shapes
    @property
    def kind(self) -> int:
        return self.storage_type & 0x0F

    @property
    def comps(self) -> int:
        return 3 if self.mode == 3 else 1

    @property
    def nt_saved(self) -> int:
        return (self.nt + self.S - 1) // self.S

    def tile(self):
        if self.kind != KIND_INT8:
            return (1, 1, 1)
        st = self.storage_type
        return ((st >> 8) & 0x3F, (st >> 14) & 0x3F, (st >> 20) & 0x3F)

    def box(self):
        physical = bool(self.storage_type & PHYSICAL)
        out = []
        for a, n in enumerate((self.NX - 1, self.NY - 1, self.NZ - 1)):
            t = self.tile()[a]
            if not physical:
                out.append((0, n))
                continue
            lo, hi = self.pml[2 * a], self.pml[2 * a + 1]
            begin = (lo // t) * t
            end = min(n, ((n - hi + t - 1) // t) * t)
            out.append((begin, end - begin))
        return out

    def box_cells(self) -> int:
        b = self.box()
        return b[0][1] * b[1][1] * b[2][1]

    def e_only(self) -> bool:
        return bool(self.storage_type & RHS_FROM_E)

    def history_bytes(self, frames: int, elem: int) -> int:
        cells = self.comps * frames * self.nshot * self.box_cells()
        if self.kind == KIND_INT8:
            b = self.box()
            t = self.tile()
            tiles = math.prod((b[a][1] + t[a] - 1) // t[a] for a in range(3))
            return ((cells + 3) & ~3) + 4 * self.comps * frames * self.nshot * tiles
        return cells * (elem if self.kind == KIND_F32 else 2)

    def psi_shape(self, face: int, which: int) -> tuple:
        """Shape of phi 'which' (0 e1, 1 e2, 2 h1, 3 h2) of a face (pml.py:310-348)."""
        p = self.pml[face]
        if p <= 0:
            return (0,)
        nx, ny, nz = self.NX - 1, self.NY - 1, self.NZ - 1
        a, b, c = [(p, ny, nz), (p, ny, nz), (nx, p, nz), (nx, p, nz), (nx, ny, p), (nx, ny, p)][face]
        axis = face // 2
        s = self.nshot
        table = {
            0: [(s, a + 1, b, c + 1), (s, a + 1, b + 1, c), (s, a, b + 1, c), (s, a, b, c + 1)],
            1: [(s, a, b + 1, c + 1), (s, a + 1, b + 1, c), (s, a + 1, b, c), (s, a, b, c + 1)],
            2: [(s, a, b + 1, c + 1), (s, a + 1, b, c + 1), (s, a + 1, b, c), (s, a, b + 1, c)],
        }
        return table[axis][which]


def cpml_profile(p: int, d: float, dt: float, electric: bool, v_ref: float = C0) -> np.ndarray:
    """RA, RB, RE, RF of a quartic CFS-CPML (alpha 0, kappa 1), float32 (4, 1, p)."""
    k = np.arange(p, dtype=np.float64) + (0.0 if electric else 0.5)
    sigma = (k / p) ** 4 * (0.8 * 5 * EPS0 * v_ref / d)
    den = 2 * EPS0 + dt * sigma
    out = np.stack([2 * EPS0 / den, 2 * EPS0 / den, (2 * EPS0 - dt * sigma) / den, 2 * sigma * dt / den])
    return out.astype(np.float32).reshape(4, 1, p)


def make_problem(NX: int, NY: int, NZ: int, *, nshot: int = 1, nt: int = 12, order: int = 2,
                 pml=(3, 3, 3, 3, 3, 3), d=(0.02, 0.02, 0.02), seed: int = 0, nsrc: int = 1, nrx: int = 3,
                 random_state: bool = True, pec: bool = False, cfl: float = 0.9, **kw) -> Problem:
    """Random but fully specified inputs on an extended field grid (NX, NY, NZ)."""
    rng = np.random.default_rng(seed)
    pml = list(pml)
    if NZ == 2:
        pml[4] = pml[5] = 0
    shape = (NX, NY, NZ)
    eps = np.zeros(shape, np.float32)
    sig = np.zeros(shape, np.float32)
    mu = np.zeros(shape, np.float32)
    core = (slice(0, NX - 1), slice(0, NY - 1), slice(0, NZ - 1))
    eps[core] = rng.uniform(1.0, 9.0, size=(NX - 1, NY - 1, NZ - 1))
    sig[core] = rng.uniform(0.0, 0.01, size=(NX - 1, NY - 1, NZ - 1))
    mu[core] = rng.uniform(1.0, 1.5, size=(NX - 1, NY - 1, NZ - 1))
    if pec:
        m = rng.random((NX - 1, NY - 1, NZ - 1)) < 0.05
        sig[core][m] = 200.0
        sig[core] = np.where(m, 200.0, sig[core])
    inv2 = sum(1.0 / (dd * dd) for n, dd in zip((NX - 1, NY - 1, NZ - 1), d) if n > 1)
    ssum = {2: 1.0, 4: 9 / 8 + 1 / 24, 8: 1225 / 1024 + 245 / 3072 + 49 / 5120 + 5 / 7168}[order]
    dt = cfl / (C0 * ssum * math.sqrt(inv2))
    coef_e, coef_h = [], []
    for f in range(6):
        p = pml[f]
        if p > 0:
            coef_e.append(cpml_profile(p, d[f // 2], dt, True))
            coef_h.append(cpml_profile(p, d[f // 2], dt, False))
        else:
            coef_e.append(np.zeros(0, np.float32))
            coef_h.append(np.zeros(0, np.float32))
    fshape = (nshot, NX, NY, NZ)

    def rand(shape_):
        if not random_state or int(np.prod(shape_)) == 0:
            return np.zeros(shape_, np.float32)
        return (rng.standard_normal(shape_) * 1e-3).astype(np.float32)

    prob = Problem(NX, NY, NZ, nshot, nt, order, pml, tuple(d), dt, eps, sig, mu, coef_e, coef_h,
                   [rand(fshape) for _ in range(3)], [rand(fshape) for _ in range(3)], [],
                   np.zeros((nshot, nsrc, 3), np.int32), np.zeros((nsrc, nt), np.float32),
                   np.zeros((nshot, nrx, 3), np.int32))
    for k, v in kw.items():
        setattr(prob, k, v)
    prob.phi0 = [rand(prob.psi_shape(f, w)) if pml[f] > 0 else np.zeros(0, np.float32)
                 for f in range(6) for w in range(4)]
    lo = [1, 1, 0 if NZ == 2 else 1]
    hi = [NX - 2, NY - 2, 0 if NZ == 2 else NZ - 2]
    for s in range(nshot):
        for k in range(nsrc):
            prob.src_loc[s, k] = [rng.integers(lo[a], hi[a] + 1) for a in range(3)]
        for r in range(nrx):
            prob.rx_loc[s, r] = [rng.integers(lo[a], hi[a] + 1) for a in range(3)]
    t = np.arange(nt) * dt
    f0 = 1.0 / (20 * dt)
    for k in range(nsrc):
        a2 = (math.pi * f0 * (t - 1.2 / f0)) ** 2
        prob.wave[k] = ((1 - 2 * a2) * np.exp(-a2) * (1.0 + 0.3 * k)).astype(np.float32)
    prob.data_grad = rng.standard_normal((nshot, nt, nrx)).astype(np.float32)
    prob.lam0 = [rand(fshape) for _ in range(6)] + [rand(x.shape) if x.size else np.zeros(0, np.float32)
                                                    for x in prob.phi0]
    return prob


def workload_problem(w, shots: int, nt: int, storage_type: int = 0) -> Problem:
    """Native-level inputs of a synthetic workload (deepgpr_amd.synthetic) for
    its first `shots` shots and `nt` time steps: the extended model, the CPML
    coefficients of build_pml_coeffs, the PML-shifted locations and the
    waveform, as compute() hands them to the library."""
    import torch

    from deepgpr_amd.preprocessing.model_setup import initialization, shift_locations
    from deepgpr_amd.solver.pml import build_pml_coeffs

    src_loc, rx_loc = w.source_location[:shots], w.receiver_location[:shots]
    prep = initialization("cpu", w.eps_r, w.sigma, None, w.source_amplitudes, src_loc, rx_loc, w.dx, w.dt, w.pml,
                          w.fdtd_order)
    coeffs = build_pml_coeffs(prep.er, prep.mr, w.dt, w.dx, prep.nx, prep.ny, prep.nz, prep.pmlthick, "cpu",
                              torch.float32)[6:]
    src = shift_locations(src_loc, prep.pmlthick, "cpu").numpy().astype(np.int32)
    rx = shift_locations(rx_loc, prep.pmlthick, "cpu").numpy().astype(np.int32)
    p = make_problem(prep.nx + 1, prep.ny + 1, prep.nz + 1, order=w.fdtd_order, nt=nt, nshot=shots,
                     nsrc=src.shape[1], nrx=rx.shape[1], seed=1, random_state=False,
                     pml=[int(v) for v in prep.pmlthick], d=(w.dx, w.dx, w.dx))
    p.dt = w.dt
    p.eps, p.sig, p.mu = (t.detach().numpy().astype(np.float32) for t in (prep.ere, prep.see, prep.mr))
    p.coef_e = [coeffs[2 * f].numpy() for f in range(6)]
    p.coef_h = [coeffs[2 * f + 1].numpy() for f in range(6)]
    p.src_loc, p.rx_loc = src, rx
    wave = w.source_amplitudes[:, :nt, 0].numpy().astype(np.float32)
    p.wave = np.ascontiguousarray(np.broadcast_to(wave, (src.shape[1], nt)))
    p.mode, p.S = w.mode, w.sampling_interval
    p.storage_type = storage_type
    p.data_grad = np.random.default_rng(0).standard_normal((shots, nt, rx.shape[1])).astype(np.float32)
    return p


# --------------------------------------------------------------------- CPU runs


def _cptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


def _load(path: str):
    lib = ctypes.CDLL(path)
    return lib


_LIBS: Dict[str, ctypes.CDLL] = {}


def cpu_lib(which: str) -> ctypes.CDLL:
    """'ref' (reference compiled from source) or 'oracle'."""
    path = REF_LIB if which == "ref" else ORACLE_LIB
    if which not in _LIBS:
        if not os.path.isfile(path):
            raise FileNotFoundError(path)
        _LIBS[which] = _load(path)
    return _LIBS[which]


def reference_available() -> bool:
    return os.path.isfile(REF_LIB)


def _forward_args(p: Problem, st: Dict[str, np.ndarray], real, E_saved, R_saved, coeffs, rdata):
    F = lambda x: ctypes.c_float(x) if real == np.float32 else ctypes.c_double(x)  # noqa: E731
    args = [_cptr(st["eps"]), _cptr(st["sig"]), _cptr(st["mu"]), _cptr(E_saved), _cptr(R_saved)]
    args += [_cptr(x) for x in st["E"]] + [_cptr(x) for x in st["H"]]
    args += [_cptr(c) for c in coeffs]
    args += [_cptr(x) for x in st["phi"]]
    args += [ctypes.c_int(v) for v in p.pml]
    args += [_cptr(x) for x in st["coef_e"]] + [_cptr(x) for x in st["coef_h"]]
    args += [F(p.dt), ctypes.c_int(p.nt), ctypes.c_int(p.nshot), ctypes.c_int(p.rx_loc.shape[1]),
             F(p.d[0]), F(p.d[1]), F(p.d[2])]
    args += [_cptr(p.rx_loc), _cptr(rdata), ctypes.c_int(p.rx_comp)]
    args += [ctypes.c_int(p.NX), ctypes.c_int(p.NY), ctypes.c_int(p.NZ), ctypes.c_int(p.src_loc.shape[1]),
             _cptr(p.src_loc), _cptr(st["wave"]), ctypes.c_int(p.src_comp), ctypes.c_int(p.S),
             ctypes.c_int(p.mode), ctypes.c_int(p.storage_type), ctypes.c_int(p.save_model_history), ctypes.c_int(0)]
    return args


def _backward_args(p: Problem, st, real, E_saved, R_saved, coeffs, dgrad, gsrc, geps, gsig):
    F = lambda x: ctypes.c_float(x) if real == np.float32 else ctypes.c_double(x)  # noqa: E731
    args = [_cptr(st["eps"]), _cptr(st["sig"]), _cptr(st["mu"]), _cptr(E_saved), _cptr(R_saved)]
    args += [_cptr(x) for x in st["L"][:6]]
    args += [_cptr(c) for c in coeffs]
    args += [_cptr(x) for x in st["L"][6:]]
    args += [ctypes.c_int(v) for v in p.pml]
    args += [_cptr(x) for x in st["coef_e"]] + [_cptr(x) for x in st["coef_h"]]
    args += [F(p.dt), ctypes.c_int(p.nt), ctypes.c_int(p.nshot), ctypes.c_int(p.rx_loc.shape[1]),
             F(p.d[0]), F(p.d[1]), F(p.d[2])]
    args += [ctypes.c_int(p.NX), ctypes.c_int(p.NY), ctypes.c_int(p.NZ), ctypes.c_int(p.rx_loc.shape[1]),
             _cptr(p.rx_loc), _cptr(dgrad), ctypes.c_int(p.rx_comp), ctypes.c_int(p.src_loc.shape[1]),
             _cptr(p.src_loc), ctypes.c_int(p.src_comp), _cptr(gsrc), ctypes.c_int(p.want_src),
             _cptr(geps), _cptr(gsig), ctypes.c_int(p.want_eps), ctypes.c_int(p.want_sig), ctypes.c_int(p.S),
             ctypes.c_int(p.mode), ctypes.c_int(p.storage_type), ctypes.c_int(0)]
    return args


def _state(p: Problem, real) -> Dict:
    c = lambda a: np.ascontiguousarray(a.astype(real))  # noqa: E731
    return {"eps": c(p.eps), "sig": c(p.sig), "mu": c(p.mu), "E": [c(x) for x in p.E0], "H": [c(x) for x in p.H0],
            "phi": [c(x) for x in p.phi0], "coef_e": [c(x) for x in p.coef_e], "coef_h": [c(x) for x in p.coef_h],
            "wave": c(p.wave)}


def _history_buffers(p: Problem, real):
    elem = np.dtype(real).itemsize
    if p.storage_type & NO_HISTORY:
        return np.zeros(1, np.uint8), np.zeros(1, np.uint8)
    e = np.zeros(p.history_bytes(p.nt_saved, elem), np.uint8)
    if not p.save_model_history:
        return e, np.zeros(1, np.uint8)
    r_frames = 1 if p.e_only() else p.nt_saved
    return e, np.zeros(p.history_bytes(r_frames, elem), np.uint8)


def run_cpu(p: Problem, which: str = "oracle", real=np.float32, backward: bool = True,
            threads: Optional[int] = None, timing: Optional[Dict] = None) -> Dict:
    """Forward (+ adjoint) on the CPU: which = 'ref' | 'oracle'. With a
    `timing` dict, the wall seconds of the forward and backward calls alone
    (buffers allocated beforehand) are stored under "seconds"."""
    import time
    if threads is not None:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    lib = cpu_lib(which)
    st = _state(p, real)
    coeffs = [np.zeros((p.NX, p.NY, p.NZ), real) for _ in range(6)]
    rdata = np.zeros((p.nshot, p.nt, p.rx_loc.shape[1]), real)
    E_saved, R_saved = _history_buffers(p, real)
    if which == "ref":
        assert real == np.float32
        lib.set_fdtd_order(ctypes.c_int(p.order))
        fwd, bwd = lib.forward, lib.backward
        extra = []
        err = lib.deepgpr_last_error
    else:
        pre = "oracle_f32_" if real == np.float32 else "oracle_f64_"
        fwd, bwd = getattr(lib, pre + "forward"), getattr(lib, pre + "backward")
        extra = [ctypes.c_int(p.order)]
        err = getattr(lib, pre + "last_error")
    err.restype = ctypes.c_char_p
    fwd.restype = None
    fargs = _forward_args(p, st, real, E_saved, R_saved, coeffs, rdata)
    t0 = time.perf_counter()
    fwd(*fargs, *extra)
    spent = time.perf_counter() - t0
    msg = err()
    if msg:
        raise RuntimeError(msg.decode())
    out = {"E": st["E"], "H": st["H"], "phi": st["phi"], "rx": rdata, "E_saved": E_saved, "R_saved": R_saved,
           "coeffs": coeffs}
    if not backward:
        if timing is not None:
            timing["seconds"] = spent
        return out
    st["L"] = [np.ascontiguousarray(x.astype(real)) for x in p.lam0]
    dgrad = np.ascontiguousarray(p.data_grad.astype(real))
    gsrc = np.zeros(p.wave.shape, real)
    gshape = (p.NX - 1, p.NY - 1, p.NZ - 1)
    geps = np.zeros(gshape, real)
    gsig = np.zeros(gshape, real)
    coeffs = [np.zeros((p.NX, p.NY, p.NZ), real) for _ in range(6)]
    bwd.restype = None
    bargs = _backward_args(p, st, real, E_saved, R_saved, coeffs, dgrad, gsrc, geps, gsig)
    t0 = time.perf_counter()
    bwd(*bargs, *extra)
    spent += time.perf_counter() - t0
    if timing is not None:
        timing["seconds"] = spent
    msg = err()
    if msg:
        raise RuntimeError(msg.decode())
    out.update({"L": st["L"], "g_src": gsrc, "g_eps": geps, "g_sig": gsig})
    return out


# --------------------------------------------------------------------- GPU runs


def run_gpu(p: Problem, backward: bool = True, api: str = "abi6", fp64: bool = False, device: str = "cuda:0") -> Dict:
    """The same problem through the HIP library: api 'abi6' (the ABI-6
    exports), 'dgx' (dgx_forward/dgx_backward) or 'dgx_noalloc'
    (dgx_forward_noalloc, then dgx_backward)."""
    import torch

    from deepgpr_amd.native.bridge import invoke
    from deepgpr_amd.native.loader import get_library

    lib = get_library()
    dev = torch.device(device)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    st = {"eps": T(p.eps), "sig": T(p.sig), "mu": T(p.mu), "E": [T(x) for x in p.E0], "H": [T(x) for x in p.H0],
          "phi": [T(x) for x in p.phi0], "coef_e": [T(x) for x in p.coef_e], "coef_h": [T(x) for x in p.coef_h],
          "wave": T(p.wave)}
    src, rx = T(p.src_loc), T(p.rx_loc)
    coeffs = [torch.zeros((p.NX, p.NY, p.NZ), device=dev) for _ in range(6)]
    rdata = torch.zeros((p.nshot, p.nt, p.rx_loc.shape[1]), device=dev)
    e_np, r_np = _history_buffers(p, np.float32)
    E_saved, R_saved = T(e_np), T(r_np)
    common = {"fdtd_order": p.order, "pml": p.pml, "dt": p.dt, "nt": p.nt, "nshot": p.nshot,
              "nreceiver": p.rx_loc.shape[1], "dx": p.d[0], "dy": p.d[1], "dz": p.d[2], "nx_fields": p.NX,
              "ny_fields": p.NY, "nz_fields": p.NZ, "receiver_component": p.rx_comp,
              "nsource": p.src_loc.shape[1], "source_component": p.src_comp, "sampling_interval": p.S,
              "fwi_mode": p.mode, "storage_type": p.storage_type, "accumulate_fp64": int(fp64),
              "eps_r_pad": st["eps"], "sigma_pad": st["sig"], "mu_r_pad": st["mu"], "E_saved": E_saved,
              "R_saved": R_saved, "pml_coeff": st["coef_e"] + st["coef_h"], "receiver_location": rx,
              "source_location": src}
    if api == "abi6":
        ptr = lambda t: ctypes.c_void_p(t.data_ptr() if t.numel() else 0)  # noqa: E731
        lib.set_fdtd_order(p.order)
        lib.forward.argtypes = None
        fargs = [ptr(st["eps"]), ptr(st["sig"]), ptr(st["mu"]), ptr(E_saved), ptr(R_saved)]
        fargs += [ptr(x) for x in st["E"] + st["H"] + coeffs + st["phi"]]
        fargs += [ctypes.c_int(v) for v in p.pml] + [ptr(x) for x in st["coef_e"] + st["coef_h"]]
        fargs += [ctypes.c_float(p.dt), ctypes.c_int(p.nt), ctypes.c_int(p.nshot), ctypes.c_int(p.rx_loc.shape[1]),
                  ctypes.c_float(p.d[0]), ctypes.c_float(p.d[1]), ctypes.c_float(p.d[2]), ptr(rx), ptr(rdata),
                  ctypes.c_int(p.rx_comp), ctypes.c_int(p.NX), ctypes.c_int(p.NY), ctypes.c_int(p.NZ),
                  ctypes.c_int(p.src_loc.shape[1]), ptr(src), ptr(st["wave"]), ctypes.c_int(p.src_comp),
                  ctypes.c_int(p.S), ctypes.c_int(p.mode), ctypes.c_int(p.storage_type),
                  ctypes.c_int(p.save_model_history), ctypes.c_int(0)]
        torch.cuda.synchronize(dev)
        lib.forward(*fargs)
        msg = lib.deepgpr_last_error()
        if msg:
            raise RuntimeError(msg.decode())
    else:
        args = dict(common, E=st["E"], H=st["H"], update_coeff=coeffs, phi=st["phi"], receiver_data=rdata,
                    source_waveform=st["wave"], save_model_history=p.save_model_history)
        invoke("forward_noalloc" if api == "dgx_noalloc" else "forward", args, dev)
    torch.cuda.synchronize(dev)
    N = lambda t: t.cpu().numpy()  # noqa: E731
    out = {"E": [N(x) for x in st["E"]], "H": [N(x) for x in st["H"]], "phi": [N(x) for x in st["phi"]],
           "rx": N(rdata), "E_saved": N(E_saved), "R_saved": N(R_saved), "coeffs": [N(c) for c in coeffs]}
    if not backward:
        return out
    L = [T(x) for x in p.lam0]
    dgrad = T(p.data_grad)
    gsrc = torch.zeros(p.wave.shape, device=dev)
    gshape = (p.NX - 1, p.NY - 1, p.NZ - 1)
    geps = torch.zeros(gshape, device=dev)
    gsig = torch.zeros(gshape, device=dev)
    coeffs = [torch.zeros((p.NX, p.NY, p.NZ), device=dev) for _ in range(6)]
    if api == "abi6":
        ptr = lambda t: ctypes.c_void_p(t.data_ptr() if t.numel() else 0)  # noqa: E731
        lib.backward.argtypes = None
        bargs = [ptr(st["eps"]), ptr(st["sig"]), ptr(st["mu"]), ptr(E_saved), ptr(R_saved)]
        bargs += [ptr(x) for x in L[:6] + coeffs + L[6:]]
        bargs += [ctypes.c_int(v) for v in p.pml] + [ptr(x) for x in st["coef_e"] + st["coef_h"]]
        bargs += [ctypes.c_float(p.dt), ctypes.c_int(p.nt), ctypes.c_int(p.nshot), ctypes.c_int(p.rx_loc.shape[1]),
                  ctypes.c_float(p.d[0]), ctypes.c_float(p.d[1]), ctypes.c_float(p.d[2]), ctypes.c_int(p.NX),
                  ctypes.c_int(p.NY), ctypes.c_int(p.NZ), ctypes.c_int(p.rx_loc.shape[1]), ptr(rx), ptr(dgrad),
                  ctypes.c_int(p.rx_comp), ctypes.c_int(p.src_loc.shape[1]), ptr(src), ctypes.c_int(p.src_comp),
                  ptr(gsrc), ctypes.c_int(p.want_src), ptr(geps), ptr(gsig), ctypes.c_int(p.want_eps),
                  ctypes.c_int(p.want_sig), ctypes.c_int(p.S), ctypes.c_int(p.mode), ctypes.c_int(p.storage_type),
                  ctypes.c_int(0)]
        lib.backward(*bargs)
        msg = lib.deepgpr_last_error()
        if msg:
            raise RuntimeError(msg.decode())
    else:
        args = dict(common, E=L[:3], H=L[3:6], update_coeff=coeffs, phi=L[6:], data_grad=dgrad, grad_source=gsrc,
                    grad_eps_r=geps, grad_sigma=gsig, source_requires_grad=p.want_src,
                    eps_r_requires_grad=p.want_eps, sigma_requires_grad=p.want_sig)
        invoke("backward", args, dev)
    torch.cuda.synchronize(dev)
    out.update({"L": [N(x) for x in L], "g_src": N(gsrc), "g_eps": N(geps), "g_sig": N(gsig)})
    return out


# --------------------------------------------------------------------- comparisons

KEYS_FWD = ("E", "H", "phi", "rx")
KEYS_BWD = ("L", "g_src", "g_eps", "g_sig")


def flatten(out: Dict, keys: Sequence[str]) -> Dict[str, np.ndarray]:
    flat = {}
    for k in keys:
        v = out[k]
        if isinstance(v, list):
            for i, a in enumerate(v):
                flat[f"{k}[{i}]"] = np.asarray(a)
        else:
            flat[k] = np.asarray(v)
    return flat


def bitwise_diff(a: Dict, b: Dict, keys: Sequence[str]) -> List[str]:
    """Names whose arrays differ in any bit (NaN == NaN; -0 == +0)."""
    fa, fb = flatten(a, keys), flatten(b, keys)
    bad = []
    for k in fa:
        x, y = fa[k], fb[k]
        if x.shape != y.shape:
            bad.append(f"{k}: shape {x.shape} vs {y.shape}")
            continue
        if x.size and not np.array_equal(x, y, equal_nan=True):
            n = int(np.sum(~((x == y) | (np.isnan(x) & np.isnan(y)))))
            bad.append(f"{k}: {n}/{x.size} differ")
    return bad


def rel_l2(x: np.ndarray, y: np.ndarray) -> float:
    x = np.asarray(x, np.float64).ravel()
    y = np.asarray(y, np.float64).ravel()
    den = np.linalg.norm(y)
    return float(np.linalg.norm(x - y) / den) if den > 0 else float(np.linalg.norm(x - y))


def with_storage(p: Problem, **kw) -> Problem:
    return replace(p, **kw)


# --------------------------------------------------------------------- fixtures

_SCALARS = ("NX", "NY", "NZ", "nshot", "nt", "order", "dt", "src_comp", "rx_comp", "S", "mode", "storage_type",
            "save_model_history", "want_src", "want_eps", "want_sig")


def save_case(path: str, p: Problem, out: Dict) -> None:
    """Inputs and expected outputs of one case as a compressed .npz fixture."""
    arrays = {f"in_{k}": np.asarray(getattr(p, k)) for k in _SCALARS}
    arrays["in_pml"] = np.asarray(p.pml, np.int32)
    arrays["in_d"] = np.asarray(p.d, np.float64)
    for k in ("eps", "sig", "mu", "src_loc", "wave", "rx_loc", "data_grad"):
        arrays[f"in_{k}"] = getattr(p, k)
    for name in ("coef_e", "coef_h", "E0", "H0", "phi0", "lam0"):
        for i, a in enumerate(getattr(p, name)):
            arrays[f"in_{name}_{i}"] = a
    for k, v in flatten(out, KEYS_FWD + KEYS_BWD).items():
        arrays[f"out_{k}"] = v
    arrays["out_E_saved"] = out["E_saved"]
    arrays["out_R_saved"] = out["R_saved"]
    np.savez_compressed(path, **arrays)


def load_case(path: str):
    z = np.load(path, allow_pickle=False)
    sc = {k: z[f"in_{k}"].item() for k in _SCALARS}
    lists = {}
    for name, count in (("coef_e", 6), ("coef_h", 6), ("E0", 3), ("H0", 3), ("phi0", 24), ("lam0", 30)):
        lists[name] = [z[f"in_{name}_{i}"] for i in range(count)]
    p = Problem(int(sc["NX"]), int(sc["NY"]), int(sc["NZ"]), int(sc["nshot"]), int(sc["nt"]), int(sc["order"]),
                [int(v) for v in z["in_pml"]], tuple(float(v) for v in z["in_d"]), float(sc["dt"]), z["in_eps"],
                z["in_sig"], z["in_mu"], lists["coef_e"], lists["coef_h"], lists["E0"], lists["H0"],
                lists["phi0"], z["in_src_loc"], z["in_wave"], z["in_rx_loc"])
    for k in ("src_comp", "rx_comp", "S", "mode", "storage_type", "save_model_history", "want_src", "want_eps",
              "want_sig"):
        setattr(p, k, int(sc[k]))
    p.data_grad = z["in_data_grad"]
    p.lam0 = lists["lam0"]
    expected = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return p, expected


def compare_to_fixture(out: Dict, expected: Dict) -> List[str]:
    flat = flatten(out, KEYS_FWD + KEYS_BWD)
    flat["E_saved"] = out["E_saved"]
    flat["R_saved"] = out["R_saved"]
    bad = []
    for k, v in expected.items():
        got = np.asarray(flat[k])
        if got.shape != v.shape or not np.array_equal(got, v, equal_nan=True):
            bad.append(k)
    return bad
