"""fp64 diagnostics on the GPU against the CPU oracle's fp64 sums."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_forces_f64_match_cpu(nb, orc, gpu):
    import torch

    n = 6000
    pos, _ = nb.disk_ic(n)
    idx = np.array([0, 1, 17, 2999, 5999, 4242, 123], np.int32)
    P = torch.from_numpy(np.stack(pos)).to(gpu)
    params = nb.default_params(num_particles=n)
    got = nb.forces_f64(params, P, torch.from_numpy(idx).to(gpu)).cpu().numpy()
    ref = np.stack(orc.forces_f64(pos, idx.astype(np.int64), eps=float(np.float32(1e-7))))
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_forces_f64_invalid_index_is_nan(nb, gpu):
    import torch

    n = 100
    pos, _ = nb.disk_ic(n)
    P = torch.from_numpy(np.stack(pos)).to(gpu)
    idx = torch.tensor([3, 100, -1], dtype=torch.int32, device=gpu)
    got = nb.forces_f64(nb.default_params(num_particles=n), P, idx).cpu().numpy()
    assert np.isfinite(got[:, 0]).all() and np.isnan(got[:, 1:]).all()


@pytest.mark.parametrize("n", [1, 2, 300, 5000])
def test_energy_momentum_match_cpu(nb, orc, gpu, n):
    import torch

    pos, vel = nb.plummer_ic(n, seed=4, g_mass=1.0)
    params = nb.default_params(num_particles=n, G=1.0 / n, dist_eps=1e-4)
    e = nb.energy_momentum(params, torch.from_numpy(np.stack(pos)).to(gpu),
                           torch.from_numpy(np.stack(vel)).to(gpu))
    ref = orc.energy_momentum_f64(orc.params(G=1.0 / n, dist_eps=1e-4), pos, vel)
    assert np.isclose(e["kinetic"], ref[0], rtol=1e-12)
    assert np.isclose(e["potential"], ref[1], rtol=1e-10, atol=1e-300)
    assert np.allclose(e["momentum"], ref[2:], rtol=1e-10, atol=1e-9)
