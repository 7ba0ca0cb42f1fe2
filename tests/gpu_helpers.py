"""Capture the native calls compute() makes (inputs before, outputs after) and
rebuild them as oracle problems, so an end-to-end run can be replayed through
the CPU oracle (test infrastructure; imported by the GPU tests)."""
import torch


def capture(monkeypatch):
    """Record every native call of the autograd layer as (function, inputs, outputs)."""
    import deepgpr_amd.solver.autograd as ag
    calls = []
    real = ag.invoke

    def spy(function, args, device):
        snap = {k: ([t.detach().clone() if isinstance(t, torch.Tensor) else t for t in v]
                    if isinstance(v, (list, tuple)) else (v.detach().clone() if isinstance(v, torch.Tensor) else v))
                for k, v in args.items()}
        real(function, args, device)
        calls.append((function, snap, args))

    monkeypatch.setattr(ag, "invoke", spy)
    return calls


def np_of(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else t


def problem(h, fwd_args, bwd_args=None):
    """The oracle problem of a captured forward (and backward) call."""
    a = fwd_args
    p = h.Problem(a["nx_fields"], a["ny_fields"], a["nz_fields"], a["nshot"], a["nt"], a["fdtd_order"],
                  list(a["pml"]), (a["dx"], a["dy"], a["dz"]), a["dt"], np_of(a["eps_r_pad"]),
                  np_of(a["sigma_pad"]), np_of(a["mu_r_pad"]), [np_of(t) for t in a["pml_coeff"][:6]],
                  [np_of(t) for t in a["pml_coeff"][6:]], [np_of(t) for t in a["E"]], [np_of(t) for t in a["H"]],
                  [np_of(t) for t in a["phi"]], np_of(a["source_location"]),
                  np_of(a["source_waveform"]).reshape(a["nsource"], a["nt"]), np_of(a["receiver_location"]))
    p.src_comp, p.rx_comp = a["source_component"], a["receiver_component"]
    p.S, p.mode = a["sampling_interval"], a["fwi_mode"]
    p.storage_type = a["storage_type"] & ~(1 << 6)
    p.save_model_history = a["save_model_history"]
    if bwd_args is not None:
        b = bwd_args
        p.lam0 = [np_of(t) for t in b["E"]] + [np_of(t) for t in b["H"]] + [np_of(t) for t in b["phi"]]
        p.data_grad = np_of(b["data_grad"])
        p.want_src, p.want_eps, p.want_sig = (b["source_requires_grad"], b["eps_r_requires_grad"],
                                              b["sigma_requires_grad"])
    return p
