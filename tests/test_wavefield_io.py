"""On-disk history format of ``save_forward_wavefield_path``
(src/DeepGPR/postprocessing/io.py:17-92, tests/test_modeling.py of the reference
covers the same contract through compute)."""

import os
import re
from datetime import datetime

import pytest
import torch

from deepgpr_amd.solver.modeling import save_forward_wavefield, wavefield_directory

WHEN = datetime(2024, 5, 17, 9, 4, 30)


def test_bare_tensor_and_name(tmp_path):
    e = torch.arange(24, dtype=torch.float32).reshape(1, 2, 3, 4)
    out = tmp_path / "new" / "dir"
    path = save_forward_wavefield(e, str(out), WHEN)
    assert os.path.isabs(path)
    assert os.path.basename(path) == "forward_wavefield_09-04.pt"
    back = torch.load(path, weights_only=True)
    assert isinstance(back, torch.Tensor) and back.device.type == "cpu"
    assert torch.equal(back, e)


def test_collisions_get_numbered_names(tmp_path):
    e = torch.zeros(3)
    names = [os.path.basename(save_forward_wavefield(e + k, str(tmp_path), WHEN)) for k in range(4)]
    assert names == ["forward_wavefield_09-04.pt", "forward_wavefield_09-04_01.pt",
                     "forward_wavefield_09-04_02.pt", "forward_wavefield_09-04_03.pt"]
    for k, n in enumerate(names):
        assert torch.equal(torch.load(tmp_path / n, weights_only=True), e + k)


def test_metadata_dict_format(tmp_path):
    packed = torch.randint(-127, 128, (64,), dtype=torch.int8)
    meta = {"compression": "int8", "block_size": (4, 4, 4), "uncompressed_shape": (1, 2, 1, 4, 4, 4)}
    path = save_forward_wavefield(packed, str(tmp_path), WHEN, metadata=meta)
    back = torch.load(path, weights_only=True)
    assert set(back) == {"wavefield", "compression", "block_size", "uncompressed_shape"}
    assert torch.equal(back["wavefield"], packed)
    assert back["compression"] == "int8" and tuple(back["block_size"]) == (4, 4, 4)


def test_failed_write_leaves_no_file(tmp_path):
    with pytest.raises(Exception):
        save_forward_wavefield(torch.zeros(2), str(tmp_path), WHEN, metadata={"bad": lambda: 0})
    assert not any(re.match(r"forward_wavefield_", p.name) for p in tmp_path.iterdir())


def test_directory_validation(tmp_path):
    assert wavefield_directory(None) is None
    assert wavefield_directory(tmp_path) == str(tmp_path)
    assert wavefield_directory("~/x").startswith(os.path.expanduser("~"))
    with pytest.raises(TypeError):
        wavefield_directory(3.5)
    f = tmp_path / "a_file"
    f.write_text("x")
    with pytest.raises(NotADirectoryError):
        wavefield_directory(str(f))


@pytest.mark.gpu
def test_compute_saves_history_file(tmp_path):
    """compute writes the returned history, bare, under the HH-MM name
    (tests/test_numerics.py:304-338 of the reference)."""
    import deepgpr_amd as DeepGPR

    dev = torch.device("cuda:0")
    nx, ny, nt = 10, 12, 40
    er = torch.full((nx, ny), 4.0, device=dev, requires_grad=True)
    se = torch.zeros((nx, ny), device=dev, requires_grad=True)
    source = torch.zeros((1, nt, 1), device=dev)
    source[0, 5, 0] = 1.0
    result = DeepGPR.compute(device=dev, dx=0.02, dt=3.0e-11, source_amplitudes=source,
                             source_location=torch.tensor([[[3, 3, 0]]], dtype=torch.int32, device=dev),
                             receiver_location=torch.tensor([[[3, 6, 0]]], dtype=torch.int32, device=dev),
                             er=er, se=se, pmlthick=2, save_forward_wavefield_path=str(tmp_path))
    files = list(tmp_path.glob("*.pt"))
    assert len(files) == 1 and re.fullmatch(r"forward_wavefield_\d{2}-\d{2}\.pt", files[0].name)
    assert torch.equal(torch.load(files[0], weights_only=True), result[0].detach().cpu())
    result[-1].square().mean().backward()
    assert torch.isfinite(er.grad).all() and torch.isfinite(se.grad).all()


@pytest.mark.gpu
def test_compute_saves_physical_region_payload(tmp_path):
    """A physical-region history is saved as a dict that records its box
    (tests/test_history_region.py:155-170 of the reference)."""
    import deepgpr_amd as DeepGPR

    dev = torch.device("cuda:0")
    nx, ny, nt = 17, 19, 40
    x = torch.linspace(0.0, 1.0, nx)[:, None]
    y = torch.linspace(0.0, 1.0, ny)[None, :]
    eps_r = (4.0 + 0.3 * x + 0.2 * y).to(dev)
    sigma = (2.0e-4 + 1.0e-4 * x + 0.0 * y).to(dev)
    dt = 2.5e-11
    result = DeepGPR.compute(
        device=dev, dx=(0.020, 0.016, 0.012), dt=dt,
        source_amplitudes=DeepGPR.ricker(4.0e8, nt, dt, 1.5e-9).reshape(1, nt, 1).to(dev),
        source_location=torch.tensor([[[7, 6, 0]], [[9, 8, 0]]], dtype=torch.int32, device=dev),
        receiver_location=torch.tensor([[[6, 11, 0], [8, 13, 0]], [[5, 12, 0], [10, 14, 0]]],
                                       dtype=torch.int32, device=dev),
        pmlthick=[3, 5, 4, 2, 0, 0], fdtd_order=4, mode=2, er=eps_r, se=sigma, wavefield_history_region="physical",
        save_forward_wavefield_path=str(tmp_path))
    files = list(tmp_path.glob("forward_wavefield_*.pt"))
    assert len(files) == 1
    payload = torch.load(files[0], weights_only=True)
    assert payload["history_region"] == "physical"
    assert tuple(payload["history_origin"]) == (3, 4, 0)
    assert list(payload["pmlthick"]) == [3, 5, 4, 2, 0, 0]
    assert tuple(payload["uncompressed_shape"]) == tuple(result[0].shape)
    assert torch.equal(payload["wavefield"], result[0].detach().cpu())
