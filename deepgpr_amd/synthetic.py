"""Seeded synthetic models and acquisitions of the benchmark workloads.

All randomness comes from torch CPU generators in float64, then cast to
float32, so a workload is identical on every machine. The depth axis is x.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Tuple

import torch

from .preprocessing.wavelets import ricker


@dataclass
class Workload:
    name: str
    eps_r: torch.Tensor
    sigma: torch.Tensor
    source_location: torch.Tensor  # (nshot, nsrc, 3) physical-frame
    receiver_location: torch.Tensor  # (nshot, nrx, 3)
    source_amplitudes: torch.Tensor  # (1, nt, 1)
    dx: float
    dt: float
    nt: int
    fdtd_order: int
    pml: int
    mode: int
    sampling_interval: int
    kwargs: Dict

    @property
    def physical_shape(self) -> Tuple[int, ...]:
        return tuple(self.eps_r.shape)

    def extended_cells(self) -> int:
        shape = list(self.physical_shape) + [1] * (3 - self.eps_r.ndim)
        active = [True, True, len(self.physical_shape) == 3]
        return math.prod(n + (2 * self.pml if a else 0) for n, a in zip(shape, active))

    def cell_updates(self) -> int:
        """Forward + adjoint cell-updates of one solve (CPML included, halo not)."""
        return 2 * self.extended_cells() * self.nt * self.source_location.shape[0]


def _gaussian_field(shape, corr: float, gen: torch.Generator) -> torch.Tensor:
    """Unit-variance Gaussian random field: white noise convolved (periodically)
    with a Gaussian kernel of standard deviation ``corr`` cells."""
    noise = torch.randn(shape, generator=gen, dtype=torch.float64)
    freqs = [torch.fft.fftfreq(n, dtype=torch.float64) for n in shape]
    grids = torch.meshgrid(*freqs, indexing="ij")
    k2 = sum(g * g for g in grids)
    filt = torch.exp(-2.0 * (math.pi ** 2) * (corr ** 2) * k2)
    field = torch.fft.ifftn(torch.fft.fftn(noise) * filt).real
    field = field - field.mean()
    return field / field.std()


def c1_small_2d() -> Workload:
    """100x100 2D, order 2, eps 2 above row 50 and 5 below, one Ricker source."""
    nx = ny = 100
    eps = torch.full((nx, ny), 2.0)
    eps[50:] = 5.0
    sig = torch.zeros(nx, ny)
    nt, dt = 2000, 3e-11
    w = ricker(2e8, nt, dt, 1.0 / 2e8, dtype=torch.float64).float().reshape(1, nt, 1)
    return Workload("c1_small_2d_100x100_o2_nt2000", eps, sig, torch.tensor([[[10, 10, 0]]]),
                    torch.tensor([[[10, 90, 0]]]), w, 0.02, dt, nt, 2, 10, 2, 1, {})


def c2_fwi_2d(nshot: int = 20, nt: int = 3000) -> Workload:
    """400x200 2D multi-shot FWI gradient: log-normal Gaussian random field
    (correlation 10 cells) plus 3-6 ellipses, order 4, CPML 20, mode 3."""
    nx, ny = 400, 200
    g = _gaussian_field((nx, ny), 10.0, torch.Generator().manual_seed(2026))
    eps = torch.clamp(torch.exp(math.log(6.0) + 0.5 * g), 1.0, 25.0)
    gen = torch.Generator().manual_seed(2027)
    count = int(torch.randint(3, 7, (1,), generator=gen))
    xs = torch.arange(nx, dtype=torch.float64)[:, None]
    ys = torch.arange(ny, dtype=torch.float64)[None, :]
    for _ in range(count):
        cx = float(torch.rand(1, generator=gen, dtype=torch.float64)) * nx
        cy = float(torch.rand(1, generator=gen, dtype=torch.float64)) * ny
        a, b = (5.0 + 25.0 * torch.rand(2, generator=gen, dtype=torch.float64)).tolist()
        th = float(torch.rand(1, generator=gen, dtype=torch.float64)) * math.pi
        val = 1.0 + 24.0 * float(torch.rand(1, generator=gen, dtype=torch.float64))
        u = (xs - cx) * math.cos(th) + (ys - cy) * math.sin(th)
        v = -(xs - cx) * math.sin(th) + (ys - cy) * math.cos(th)
        eps = torch.where((u / a) ** 2 + (v / b) ** 2 <= 1.0, torch.tensor(val, dtype=torch.float64), eps)
    sig = 0.01 * torch.rand((nx, ny), generator=torch.Generator().manual_seed(2028), dtype=torch.float64)
    sy = torch.round(torch.linspace(10, 189, nshot)).long()
    src = torch.stack([torch.tensor([[5, int(y), 0]]) for y in sy])
    ry = torch.round(torch.linspace(0, 199, 100)).long()
    rx_one = torch.stack([torch.full_like(ry, 5), ry, torch.zeros_like(ry)], dim=1)
    rx = rx_one.unsqueeze(0).repeat(nshot, 1, 1)
    dt = 3e-11
    w = ricker(2.5e8, nt, dt, 4e-9, dtype=torch.float64).float().reshape(1, nt, 1)
    return Workload(f"c2_fwi_2d_400x200_o4_{nshot}shots_nt{nt}", eps.float(), sig.float(), src, rx, w, 0.02, dt,
                    nt, 4, 20, 3, 1, {})


def c3_layered_3d(nshot: int = 4, nt: int = 2000, shape=(256, 256, 192), S: int = 4) -> Workload:
    """3D 4-layer model with 10 spheres, order 4, CPML 16, 16x16 receiver grid."""
    nx, ny, nz = shape
    gen = torch.Generator().manual_seed(3031)
    eps = torch.empty(shape, dtype=torch.float64)
    sig = torch.empty(shape, dtype=torch.float64)
    bounds = [0, nx // 4, nx // 2, 3 * nx // 4, nx]
    for layer, val in enumerate((3.0, 6.0, 9.0, 16.0)):
        eps[bounds[layer]:bounds[layer + 1]] = val
        sig[bounds[layer]:bounds[layer + 1]] = 0.005 * float(torch.rand(1, generator=gen, dtype=torch.float64))
    gen = torch.Generator().manual_seed(3032)
    xs = torch.arange(nx, dtype=torch.float64)[:, None, None]
    ys = torch.arange(ny, dtype=torch.float64)[None, :, None]
    zs = torch.arange(nz, dtype=torch.float64)[None, None, :]
    for _ in range(10):
        c = torch.rand(3, generator=gen, dtype=torch.float64) * torch.tensor(shape, dtype=torch.float64)
        r = 5.0 + 10.0 * float(torch.rand(1, generator=gen, dtype=torch.float64))
        e_val = 1.0 + 24.0 * float(torch.rand(1, generator=gen, dtype=torch.float64))
        s_val = 0.005 * float(torch.rand(1, generator=gen, dtype=torch.float64))
        inside = (xs - c[0]) ** 2 + (ys - c[1]) ** 2 + (zs - c[2]) ** 2 <= r * r
        eps[inside] = e_val
        sig[inside] = s_val
    quad = [(ny // 4, nz // 4), (ny // 4, 3 * nz // 4), (3 * ny // 4, nz // 4), (3 * ny // 4, 3 * nz // 4)]
    src = torch.tensor([[[5, y, z]] for y, z in quad[:nshot]])
    gy = torch.round(torch.linspace(0, ny - 1, 16)).long()
    gz = torch.round(torch.linspace(0, nz - 1, 16)).long()
    yy, zz = torch.meshgrid(gy, gz, indexing="ij")
    rx_one = torch.stack([torch.full_like(yy.reshape(-1), 5), yy.reshape(-1), zz.reshape(-1)], dim=1)
    rx = rx_one.unsqueeze(0).repeat(nshot, 1, 1)
    dt = 2.5e-11
    w = ricker(3e8, nt, dt, 1.0 / 3e8, dtype=torch.float64).float().reshape(1, nt, 1)
    return Workload(f"c3_layered_3d_{nx}x{ny}x{nz}_o4_{nshot}shots_nt{nt}_S{S}", eps.float(), sig.float(), src, rx,
                    w, 0.02, dt, nt, 4, 16, 3, S, {"wavefield_history_region": "extended"})


WORKLOADS = {"c1": c1_small_2d, "c2": c2_fwi_2d, "c3": c3_layered_3d}


def _layered_model(shape, gen: torch.Generator, eps_hi: float = 25.0, sig_hi: float = 0.02):
    """Layers along x plus 1-4 spherical (circular) inclusions, eps_r in
    [1, eps_hi], sigma in [0, sig_hi]."""
    nx = shape[0]
    nlay = int(torch.randint(2, 7, (1,), generator=gen))
    bounds = sorted(torch.randint(1, nx, (nlay - 1,), generator=gen).tolist())
    eps = torch.empty(shape, dtype=torch.float64)
    sig = torch.empty(shape, dtype=torch.float64)
    edges = [0] + bounds + [nx]
    for a, b in zip(edges[:-1], edges[1:]):
        eps[a:b] = 1.0 + (eps_hi - 1.0) * float(torch.rand(1, generator=gen, dtype=torch.float64))
        sig[a:b] = sig_hi * float(torch.rand(1, generator=gen, dtype=torch.float64))
    axes = [torch.arange(n, dtype=torch.float64).reshape([-1 if k == d else 1 for k in range(len(shape))])
            for d, n in enumerate(shape)]
    for _ in range(int(torch.randint(1, 5, (1,), generator=gen))):
        c = [float(torch.rand(1, generator=gen, dtype=torch.float64)) * n for n in shape]
        r = 3.0 + 0.1 * min(shape) * float(torch.rand(1, generator=gen, dtype=torch.float64))
        inside = sum((ax - cc) ** 2 for ax, cc in zip(axes, c)) <= r * r
        eps = torch.where(inside, 1.0 + (eps_hi - 1.0) * float(torch.rand(1, generator=gen, dtype=torch.float64)), eps)
        sig = torch.where(inside, sig_hi * float(torch.rand(1, generator=gen, dtype=torch.float64)), sig)
    return eps.float(), sig.float()


def serving_stream(n: int, rate: float = 50.0, seed: int = 5051, nt_scale: float = 1.0):
    """Mixed Poisson request stream: 80 % 2D forward-only (grids 128^2-512^2,
    nt 1000-4000, order 2/4/8), 15 % 2D forward+gradient, 5 % 3D forward
    (96^3-192^3, nt 1000-2000). Layered-plus-inclusion models, Ricker sources
    at 100-500 MHz, CPML 10, dx 0.02 m; dt 3e-11 s in 2D and 2.5e-11 s in 3D.
    ``nt_scale`` shortens every run (for quick checks)."""
    from .serving import SolveRequest

    gen = torch.Generator().manual_seed(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += float(-math.log(1.0 - float(torch.rand(1, generator=gen, dtype=torch.float64)))) / rate
        u = float(torch.rand(1, generator=gen, dtype=torch.float64))
        three_d = u >= 0.95
        gradient = 0.80 <= u < 0.95
        if three_d:
            side = lambda: int(torch.randint(96, 193, (1,), generator=gen)) // 8 * 8  # noqa: E731
            shape = (side(), side(), side())
            nt = int(torch.randint(1000, 2001, (1,), generator=gen))
            order = int([2, 4][int(torch.randint(0, 2, (1,), generator=gen))])
            dt = 2.5e-11
        else:
            side = lambda: int(torch.randint(128, 513, (1,), generator=gen)) // 16 * 16  # noqa: E731
            shape = (side(), side())
            nt = int(torch.randint(1000, 4001, (1,), generator=gen))
            order = int([2, 4, 8][int(torch.randint(0, 3, (1,), generator=gen))])
            dt = 3e-11
        nt = max(16, int(nt * nt_scale))
        eps, sig = _layered_model(shape, gen)
        freq = 1e8 + 4e8 * float(torch.rand(1, generator=gen, dtype=torch.float64))
        wave = ricker(freq, nt, dt, 1.0 / freq, dtype=torch.float64).float().reshape(1, nt, 1)
        nz = shape[2] // 2 if three_d else 0
        src = torch.tensor([[[5, shape[1] // 2, nz]]])
        nrx = 32
        ry = torch.round(torch.linspace(0, shape[1] - 1, nrx)).long()
        rx = torch.stack([torch.full_like(ry, 5), ry, torch.full_like(ry, nz)], dim=1).unsqueeze(0)
        out.append(SolveRequest(eps, sig, wave, src, rx, 0.02, dt, fdtd_order=order, pmlthick=10,
                                gradient=gradient, options={"mode": 2}, rid=rid, arrival=t))
    return out


def batch_instances(n: int = 64, seed: int = 4041, nt_scale: float = 1.0):
    """Independent 2D forward+adjoint solves of uneven size: nx, ny i.i.d. from
    {128, 192, 256, 384}, order 2 or 4, nt uniform in [1000, 3000] (times
    ``nt_scale``), CPML 10, dx 0.02 m, dt 3e-11 s; 2-6 layers along x with
    eps_r in [1, 20] and sigma in [0, 0.01] S/m; 1-4 sources and 32-128
    receivers on row 5; Ricker 100-500 MHz."""
    from .serving import SolveRequest

    gen = torch.Generator().manual_seed(seed)
    sizes = (128, 192, 256, 384)
    out = []
    for rid in range(n):
        nx = sizes[int(torch.randint(0, 4, (1,), generator=gen))]
        ny = sizes[int(torch.randint(0, 4, (1,), generator=gen))]
        order = (2, 4)[int(torch.randint(0, 2, (1,), generator=gen))]
        nt = max(16, int(int(torch.randint(1000, 3001, (1,), generator=gen)) * nt_scale))
        nlay = int(torch.randint(2, 7, (1,), generator=gen))
        bounds = sorted(torch.randint(1, nx, (nlay - 1,), generator=gen).tolist())
        eps = torch.empty((nx, ny), dtype=torch.float64)
        sig = torch.empty((nx, ny), dtype=torch.float64)
        for a, b in zip([0] + bounds, bounds + [nx]):
            eps[a:b] = 1.0 + 19.0 * float(torch.rand(1, generator=gen, dtype=torch.float64))
            sig[a:b] = 0.01 * float(torch.rand(1, generator=gen, dtype=torch.float64))
        nsrc = int(torch.randint(1, 5, (1,), generator=gen))
        nrx = int(torch.randint(32, 129, (1,), generator=gen))
        sy = torch.round(torch.linspace(ny / (nsrc + 1), ny * nsrc / (nsrc + 1), nsrc)).long()
        src = torch.stack([torch.full_like(sy, 5), sy, torch.zeros_like(sy)], dim=1).unsqueeze(0)
        ry = torch.round(torch.linspace(0, ny - 1, nrx)).long()
        rx = torch.stack([torch.full_like(ry, 5), ry, torch.zeros_like(ry)], dim=1).unsqueeze(0)
        freq = 1e8 + 4e8 * float(torch.rand(1, generator=gen, dtype=torch.float64))
        wave = ricker(freq, nt, 3e-11, 1.0 / freq, dtype=torch.float64).float().reshape(1, nt, 1)
        out.append(SolveRequest(eps.float(), sig.float(), wave, src, rx, 0.02, 3e-11, fdtd_order=order,
                                pmlthick=10, gradient=True, options={"mode": 2}, rid=rid))
    return out
