"""nb_forward_step_noalloc_dev: the forward step with its partial sums in
the library's static device buffer, called through the C ABI, against
nb_forward_step_dev (same shape, same kernels: bit-equal) on two calls in a
row, the second with fewer particles; sizes past the buffer are refused; no
per-stream scratch is allocated by it."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def step(nb, torch, gpu, entry, params, pos, vel, cfg):
    """One step of `entry` (a C-ABI function) on fresh device copies."""
    P = [torch.from_numpy(a.copy()).to(gpu) for a in pos]
    V = [torch.from_numpy(a.copy()).to(gpu) for a in vel]
    N = [torch.empty_like(t) for t in P]
    torch.cuda.synchronize()
    ptrs = [ctypes.c_void_p(t.data_ptr()) for t in P + V + N]
    st = entry(ctypes.byref(params), *ptrs, len(pos[0]), cfg, nb.NB_ACC_PLAIN,
               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, [t.cpu().numpy() for t in N + V]


def targets_split_config(nb):
    for c in range(nb.forward_config_count()):
        i = nb.forward_config_info(c)
        if i["scheme"] == "targets" and (i["block"], i["unroll"], i["split"]) == (256, 1, 16):
            return c
    raise AssertionError("no 256 x 1 split-16 shape")


@pytest.mark.parametrize("shape", ["default", "targets_split"])
def test_noalloc_matches_allocating_step(nb, gpu, shape):
    import torch

    lib = nb.load()
    cfg = -1 if shape == "default" else targets_split_config(nb)
    torch.cuda.synchronize()
    nb.scratch_release()
    for n in (25600, 12800):  # the second call with fewer particles
        pos, vel = nb.disk_ic(n)
        params = nb.default_params(num_particles=n)
        st, got = step(nb, torch, gpu, lib.nb_forward_step_noalloc_dev, params, pos, vel, cfg)
        assert st == nb.NB_OK, lib.nb_last_error()
        assert nb.scratch_held() == 0  # the allocation-free call took no scratch
        st, ref = step(nb, torch, gpu, lib.nb_forward_step_dev, params, pos, vel, cfg)
        assert st == nb.NB_OK, lib.nb_last_error()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), (shape, n)
        nb.scratch_release()


def test_noalloc_refuses_past_its_buffer(nb, gpu):
    import torch

    lib = nb.load()
    n = 26624 + 128  # one tile past the 64 MiB of tile-pair partials
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n)
    st, _ = step(nb, torch, gpu, lib.nb_forward_step_noalloc_dev, params, pos, vel, -1)
    assert st == 5, st  # NB_ERR_UNSUPPORTED
    assert b"64 MiB" in lib.nb_last_error()
    n = 26624  # the largest default-shape size that fits
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n)
    st, got = step(nb, torch, gpu, lib.nb_forward_step_noalloc_dev, params, pos, vel, -1)
    assert st == nb.NB_OK, lib.nb_last_error()
    st, ref = step(nb, torch, gpu, lib.nb_forward_step_dev, params, pos, vel, -1)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
