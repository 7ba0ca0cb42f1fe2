"""Device arguments: the public entry points take a HIP device without an
index ("cuda") as the current device, as the reference's native bridge does
(src/DeepGPR/native/bridge.py:211-212). The bare form needs a GPU and is
covered by the GPU tests; here, devices that carry an index or are not HIP
devices pass through unchanged."""
import torch

from deepgpr_amd.utils.devices import resolve_device


def test_indexed_and_cpu_devices_pass_through():
    assert resolve_device("cpu") == torch.device("cpu")
    assert resolve_device(torch.device("cuda", 1)) == torch.device("cuda", 1)
    assert resolve_device("cuda:3") == torch.device("cuda", 3)
    assert resolve_device(torch.device("meta")) == torch.device("meta")
