"""Argument checks of the Python layer that run before any device call: host
tensors and short rows are refused with ValueError (no device pointer to host
memory ever reaches a kernel)."""
import numpy as np
import pytest


def test_device_entry_points_refuse_host_tensors(nb):
    import torch

    n = 8
    host = torch.zeros(3, n)
    params = nb.default_params(num_particles=n)
    with pytest.raises(ValueError):
        nb.adjoint_step(params, host, host, host, host, host)
    with pytest.raises(ValueError):
        nb.fwd_adjoint_step(params, host, host, host, host, host, host, host)
    with pytest.raises(ValueError):
        nb.forward_step(params, list(host), list(host), list(host))
    with pytest.raises(ValueError):
        nb.trajectory(host, host, params, 4, 2)
    with pytest.raises(ValueError):
        nb.pack_vec4(host)


def test_trajectory_shapes(nb):
    import torch

    params = nb.default_params(num_particles=4)
    with pytest.raises(ValueError):
        nb.trajectory(torch.zeros(4, 3), torch.zeros(4, 3), params, 2, 1)


def test_simulator_checks_arrays_and_copies_params(nb):
    params = nb.default_params(num_particles=16)
    bad = [np.zeros(15, np.float32)] * 3
    with pytest.raises(ValueError):
        nb.Simulator(params, bad, bad)
    assert params.sim_iterations_per_frame == 4  # untouched by the failed create
