"""The pybind11 module and the autograd wrapper import on a host without a GPU (no compute)."""
import numpy as np


def test_core_module_and_autograd_import():
    from msosm_amd import autograd as ag

    assert hasattr(ag._core, "ProfileEngine")
    assert ag.ProfileEngine is ag._core.ProfileEngine
    import torch

    prof = torch.tensor([[1.0, 2.0, 3.0], [2.0, 2.0, 2.0]], dtype=torch.float64)
    loss = ag.snr_loss(prof, 2.0)
    assert np.allclose(loss.numpy(), [-1.0, 0.0])
