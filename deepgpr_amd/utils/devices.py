"""Device arguments of the public entry points."""
from __future__ import annotations

import torch


def resolve_device(device) -> torch.device:
    """`device` as a torch.device with its index filled in: a HIP device given
    without one ("cuda") is the current device, as the reference's native
    bridge takes it (src/DeepGPR/native/bridge.py:211-212). Other devices are
    returned unchanged."""
    device = torch.device(device)
    if device.type == "cuda" and device.index is None:
        return torch.device("cuda", torch.cuda.current_device())
    return device
