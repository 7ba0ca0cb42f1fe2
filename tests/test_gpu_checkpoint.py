"""Segment checkpointing on the GPU (restates the reference's
tests/test_external_pml.py:145-191 for this build): a solve split into
segments under torch.utils.checkpoint, each segment starting from the states
the previous one returned (checkpoint_initial_field for the first), gives the
traces and final states of one unsegmented solve bit for bit, and the eps_r,
sigma and source-waveform gradients of the unsegmented adjoint within float32
summation round-off, through two FWI updates, with re-entrant and
non-re-entrant checkpoints, in 2D and 3D.

The gradients cannot be the same bits: the unsegmented adjoint sums every
time step's terms into one accumulator in reverse time order, while autograd
adds the per-segment sums (each summed from zero), which associates the same
terms differently."""
import pytest
import torch
from torch.utils.checkpoint import checkpoint

pytestmark = pytest.mark.gpu


def _problem(shape, pml, device):
    n = 1
    for v in shape:
        n *= v
    model = torch.linspace(3.0, 5.0, steps=n, device=device).reshape(shape)
    sigma = torch.full_like(model, 2.0e-4)
    mu = torch.linspace(1.0, 1.2, steps=n, device=device).reshape(shape)
    nt = 35
    source = torch.zeros((1, nt, 1), device=device)
    source[0, 1:4, 0] = torch.tensor([1.0, -0.5, 0.25], device=device)
    z = 0 if len(shape) == 2 else 2
    args = dict(device=device, dx=(0.02, 0.017, 0.015), dt=1.0e-11, source_amplitudes=source,
                source_location=torch.tensor([[[0, 3, z]], [[3, 4, z]]], device=device),
                receiver_location=torch.tensor([[[0, 3, z], [3, 4, z]], [[3, 4, z], [2, 3, z]]], device=device),
                pmlthick=pml, mu_r=mu, mode=2 if z == 0 else 3, gradient_accumulator="float32")
    return model, sigma, args


@pytest.mark.parametrize("shape,pml,reentrant", [
    ((8, 9), (2, 3, 1, 2), True), ((6, 7, 5), (1, 2, 2, 1, 1, 2), True),
    ((8, 9), (2, 3, 1, 2), False), ((6, 7, 5), (1, 2, 2, 1, 1, 2), False),
    ((30, 26, 34), (4, 5, 3, 4, 5, 3), False),
])
def test_checkpoint_segments_match_the_full_solve(cuda_device, shape, pml, reentrant):
    import deepgpr_amd as dg
    print(f"[gpu checkpoint] {shape} reentrant={reentrant}", flush=True)
    base, sigma_base, args = _problem(shape, pml, cuda_device)
    model = base.clone().requires_grad_()
    sigma = sigma_base.clone().requires_grad_()
    opt = torch.optim.SGD([model, sigma], lr=1.0e-7)
    seg_len = 12
    for _ in range(2):
        ref_model = model.detach().clone().requires_grad_()
        ref_sigma = sigma.detach().clone().requires_grad_()
        ref_source = args["source_amplitudes"].clone().requires_grad_()
        full = dg.compute(eps_r=ref_model, sigma=ref_sigma, **{**args, "source_amplitudes": ref_source})
        full[-1].square().sum().backward()

        source = args["source_amplitudes"].clone().requires_grad_()
        init = {k: v for k, v in args.items() if k not in ("mode", "mu_r", "gradient_accumulator")}
        E, H, P = dg.checkpoint_initial_field(er=model, se=sigma, mr=args["mu_r"], **init)
        state = (*E, *H, *P)
        traces = []
        for start in range(0, source.shape[1], seg_len):
            def segment(er, se, wave, *boundary):
                work = tuple(t.clone() for t in boundary)
                r = dg.compute(eps_r=er, sigma=se, **{**args, "source_amplitudes": wave}, E=work[:3], H=work[3:6],
                               PML=work[6:])
                return (*r[1], *r[2], *r[3], r[-1])
            out = checkpoint(segment, model, sigma, source[:, start:start + seg_len], *state, use_reentrant=reentrant)
            state, data = out[:-1], out[-1]
            traces.append(data)
        data = torch.cat(traces, dim=1)
        assert torch.equal(data, full[-1])
        for a, b in zip(state, (*full[1], *full[2], *full[3])):
            assert torch.equal(a, b)
        opt.zero_grad()
        data.square().sum().backward()
        for a, b in ((model.grad, ref_model.grad), (sigma.grad, ref_sigma.grad), (source.grad, ref_source.grad)):
            torch.testing.assert_close(a, b, rtol=2.0e-5, atol=max(1.0e-5, 1.0e-7 * float(b.abs().max())))
        opt.step()
        with torch.no_grad():
            model.clamp_(min=1.0)
            sigma.clamp_(min=0.0)
    assert not torch.equal(model, base)


@pytest.mark.parametrize("shape,pml,segment_steps,history", [
    ((8, 9), (2, 3, 1, 2), 12, None), ((30, 26, 34), (4, 5, 3, 4, 5, 3), 12, None),
    ((30, 26, 34), (4, 5, 3, 4, 5, 3), None, None),
    ((30, 26, 34), (4, 5, 3, 4, 5, 3), 12, "bfloat16"), ((30, 26, 34), (4, 5, 3, 4, 5, 3), 12, "int8"),
    ((30, 26, 34), (4, 5, 3, 4, 5, 3), 12, "sampled"), ((30, 26, 34), (4, 5, 3, 4, 5, 3), 12, "bare_cuda"),
])
def test_segmented_compute_matches_the_full_solve(cuda_device, shape, pml, segment_steps, history, monkeypatch):
    """The library's segment driver: traces and final states bit for bit, and
    gradients within float32 summation round-off, against one compute();
    segment_steps None plans the segments from a memory budget a third of the
    whole solve's need. The lossy histories (BF16 and INT8 E+R, which the
    re-runs take from the driver's history pool as two buffers) store the same
    frames as the unsegmented solve, and so do histories sampled every second
    step, so their gradients meet the same bound."""
    import deepgpr_amd as dg
    print(f"[gpu checkpoint] segmented_compute {shape} segment_steps={segment_steps} history={history}", flush=True)
    base, sigma_base, args = _problem(shape, pml, cuda_device)
    args.update({None: {}, "bfloat16": {"wavefield_storage_dtype": torch.bfloat16},
                 "int8": {"wavefield_compression": "int8"},
                 "sampled": {"model_gradient_sampling_interval": 2}, "bare_cuda": {}}[history])
    device = "cuda" if history == "bare_cuda" else cuda_device  # the pool matches devices by index
    ref_model = base.clone().requires_grad_()
    ref_sigma = sigma_base.clone().requires_grad_()
    full = dg.compute(eps_r=ref_model, sigma=ref_sigma, **args)
    full[-1].square().sum().backward()

    model = base.clone().requires_grad_()
    sigma = sigma_base.clone().requires_grad_()
    kw = {k: v for k, v in args.items() if k != "device"}
    budget = None
    if segment_steps is None:
        from deepgpr_amd.solver.memory import solve_footprint
        # the driver's own estimate (E-only FP32 history at S 1, FP32 sums) of a third of the steps
        n = [s + p0 + p1 for s, p0, p1 in zip(shape, pml[0::2], pml[1::2])]
        nt = args["source_amplitudes"].shape[1]
        opts = dict(mode=args["mode"], reconstruct_rhs=True, fp64_accumulators=False, nrx=2)
        budget = solve_footprint(*n, nt // 3, 2, list(pml), **opts)["total"]
        planned = dg.plan_segments(nt, budget, n, 2, list(pml), **opts)
        assert planned < nt, planned  # several segments
    from deepgpr_amd.solver import segments
    pools = []

    class Recorded(segments.HistoryPool):
        def __init__(self, device):
            super().__init__(device)
            self.takes = 0
            pools.append(self)

        def take(self, specs):
            self.takes += 1
            return super().take(specs)

    monkeypatch.setattr(segments, "HistoryPool", Recorded)
    out = dg.segmented_compute(device, eps_r=model, sigma=sigma, segment_steps=segment_steps,
                               memory_budget=budget, **kw)
    assert torch.equal(out[-1], full[-1])
    for a, b in zip((*out[1], *out[2], *out[3]), (*full[1], *full[2], *full[3])):
        assert torch.equal(a, b)
    out[-1].square().sum().backward()
    for a, b in ((model.grad, ref_model.grad), (sigma.grad, ref_sigma.grad)):
        torch.testing.assert_close(a, b, rtol=2.0e-5, atol=max(1.0e-5, 1.0e-7 * float(b.abs().max())))
    # the last segment's run and every re-run took their histories from the
    # pool, one take per segment, and the pool let go of them after the backward
    assert len(pools) == 1 and pools[0].buffer is None and not pools[0].keep
    if segment_steps is None:
        assert pools[0].takes >= 2
    else:
        S = kw.get("model_gradient_sampling_interval", 1)
        assert pools[0].takes == len(segments.segment_bounds(args["source_amplitudes"].shape[1], segment_steps, S))


def test_segmented_compute_on_the_c3_grid(cuda_device):
    """The segment driver at C3's real grid (256x256x192 + CPML 16, order 4,
    all four shots in one call, sampling interval 4, FP32 E+R history, FP32
    accumulators), shortened to nt 200 so that the unsegmented solve it is
    checked against (about 89 GB of history) still fits: a 30 GB budget
    plans several segments, whose re-runs take their histories from the
    driver's pool. Traces and final states bit for bit, gradients within
    float32 summation round-off."""
    import deepgpr_amd as dg
    from deepgpr_amd import synthetic
    print("[gpu checkpoint] segmented_compute, C3 grid, 4 shots, nt 200", flush=True)
    w = synthetic.c3_layered_3d(nshot=4, nt=200)
    kw = dict(pmlthick=w.pml, fdtd_order=w.fdtd_order, mode=w.mode,
              model_gradient_sampling_interval=w.sampling_interval, gradient_accumulator="float32", **w.kwargs)
    src, rxl, amp = (w.source_location.to(cuda_device), w.receiver_location.to(cuda_device),
                     w.source_amplitudes.to(cuda_device))
    budget = 30 * 10**9
    eps = w.eps_r.to(cuda_device).requires_grad_(True)
    sig = w.sigma.to(cuda_device).requires_grad_(True)
    out = dg.segmented_compute(cuda_device, w.dx, w.dt, amp, src, rxl, eps_r=eps, sigma=sig, memory_budget=budget,
                               **kw)
    out[-1].square().sum().backward()
    seg = (out[-1].detach().clone(), [t.detach().clone() for t in (*out[1], *out[2])], eps.grad.clone(),
           sig.grad.clone())
    del out
    torch.cuda.empty_cache()
    eps2 = w.eps_r.to(cuda_device).requires_grad_(True)
    sig2 = w.sigma.to(cuda_device).requires_grad_(True)
    full = dg.compute(cuda_device, w.dx, w.dt, amp, src, rxl, eps_r=eps2, sigma=sig2, **kw)
    full[-1].square().sum().backward()
    assert torch.equal(seg[0], full[-1])
    for a, b in zip(seg[1], (*full[1], *full[2])):
        assert torch.equal(a, b)
    for a, b in ((seg[2], eps2.grad), (seg[3], sig2.grad)):
        torch.testing.assert_close(a, b, rtol=2.0e-5, atol=max(1.0e-5, 1.0e-7 * float(b.abs().max())))


def test_segmented_compute_forward_only(cuda_device, monkeypatch):
    """Without a gradient to take, every segment runs without a history (the
    pool is never used), and traces and final states equal one compute()."""
    import deepgpr_amd as dg
    from deepgpr_amd.solver import segments
    print("[gpu checkpoint] segmented_compute forward only", flush=True)
    base, sigma_base, args = _problem((30, 26, 34), (4, 5, 3, 4, 5, 3), cuda_device)
    full = dg.compute(eps_r=base, sigma=sigma_base, **args)
    takes = []

    class Recorded(segments.HistoryPool):
        def take(self, specs):
            takes.append(specs)
            return super().take(specs)

    monkeypatch.setattr(segments, "HistoryPool", Recorded)
    kw = {k: v for k, v in args.items() if k != "device"}
    out = dg.segmented_compute(cuda_device, eps_r=base, sigma=sigma_base, segment_steps=12, **kw)
    assert not takes
    assert out[0] is None and torch.equal(out[-1], full[-1])
    for a, b in zip((*out[1], *out[2], *out[3]), (*full[1], *full[2], *full[3])):
        assert torch.equal(a, b)
