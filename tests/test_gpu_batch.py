"""Batched persistent kernel: results per system equal the single-system
kernel's (same launch shape => same summation order => bit-equal), hipGraph
replays equal direct launches on every replay, the autotuner restores the
state it trials on, and one system is checked against the CPU oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_batch(nb, sizes, seed0=100):
    params, pos, vel = [], [[], [], []], [[], [], []]
    for k, n in enumerate(sizes):
        p, v = nb.plummer_ic(n, seed=seed0 + k, g_mass=1.0)
        params.append(nb.default_params(num_particles=n, G=1.0 / max(n, 1),
                                        dt=0.005 * (1 + k % 3)))
        for c in range(3):
            pos[c].append(p[c])
            vel[c].append(v[c])
    cat = lambda a: tuple(np.concatenate(x) if x else np.zeros(0, np.float32) for x in a)
    return params, cat(pos), cat(vel)


SIZES = [1, 100, 1000, 1024, 3000, 257, 5000]


def single_system_step(nb, torch, gpu, params, pos, vel, off, n):
    dp = [torch.from_numpy(a[off:off + n].copy()).to(gpu) for a in pos]
    dv = [torch.from_numpy(a[off:off + n].copy()).to(gpu) for a in vel]
    dn = [torch.empty_like(t) for t in dp]
    nb.forward_step(params, dp, dv, dn, cfg=0)  # 256 threads x 1 target, no split
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in dn], [t.cpu().numpy() for t in dv]


def test_batch_equals_single_system_kernel(nb, gpu):
    import torch

    params, pos, vel = make_batch(nb, SIZES)
    b = nb.Batch(params, pos, vel)
    b.set_config(0)  # 256 x 1, the shape of forward config 0
    b.step(1)
    bp, bv = b.positions(), b.velocities()
    off = 0
    for p, n in zip(params, SIZES):
        sp, sv = single_system_step(nb, torch, gpu, p, pos, vel, off, n)
        for c in range(3):
            assert np.array_equal(bp[c][off:off + n], sp[c]), n
            assert np.array_equal(bv[c][off:off + n], sv[c]), n
        off += n
    b.close()


def test_graph_replays_match_direct(nb, gpu):
    params, pos, vel = make_batch(nb, SIZES)
    g = nb.Batch(params, pos, vel)
    d = nb.Batch(params, pos, vel)
    for rep in range(4):  # eager reads between replays
        g.step(2)
        d.step_direct(2)
        for a, b in zip(g.positions() + g.velocities(), d.positions() + d.velocities()):
            assert np.array_equal(a, b), rep
    g.step(3)  # odd: one graph replay + one direct launch
    d.step_direct(3)
    for a, b in zip(g.positions(), d.positions()):
        assert np.array_equal(a, b)
    g.close()
    d.close()


def test_every_batch_shape_agrees(nb, gpu):
    params, pos, vel = make_batch(nb, SIZES)
    ref = None
    for cfg in range(nb.batch_config_count()):
        b = nb.Batch(params, pos, vel)
        b.set_config(cfg)
        b.step(2)
        got = np.stack(b.positions()).astype(np.float64)
        b.close()
        if ref is None:
            ref = got
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-5, (cfg, nb.batch_config_info(cfg), err)


def test_autotune_restores_state(nb, gpu):
    params, pos, vel = make_batch(nb, SIZES)
    b = nb.Batch(params, pos, vel)
    before = b.positions() + b.velocities()
    best = b.autotune(2)
    assert 0 <= best < nb.batch_config_count() and b.config == best
    after = b.positions() + b.velocities()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    b.close()


def test_separate_launch_baseline_agrees(nb, gpu):
    params, pos, vel = make_batch(nb, SIZES)
    a = nb.Batch(params, pos, vel)
    s = nb.Batch(params, pos, vel)
    a.step(3)
    s.step_separate(3)
    x = np.stack(a.positions()).astype(np.float64)
    y = np.stack(s.positions()).astype(np.float64)
    assert np.abs(x - y).max() / np.abs(y).max() < 1e-5
    a.close()
    s.close()


def test_batch_system_vs_oracle(nb, orc, gpu):
    params, pos, vel = make_batch(nb, [700, 2000])
    b = nb.Batch(params, pos, vel)
    b.step(1)
    bp = b.positions()
    b.close()
    off, n = 700, 2000
    p = orc.params(G=params[1].G, dt=params[1].dt, num_particles=n)
    rp, _ = orc.step_f32(p, [a[off:] for a in pos], [a[off:] for a in vel])
    got = np.stack([a[off:] for a in bp]).astype(np.float64)
    ref = np.stack(rp).astype(np.float64)
    e = np.abs(got - ref).max(0) / np.maximum(np.abs(ref).max(0), 1.0)
    assert e.max() < 1e-5


def test_empty_and_tiny_batches(nb, gpu):
    b = nb.Batch([], (np.zeros(0, np.float32),) * 3, (np.zeros(0, np.float32),) * 3)
    b.step(3)
    b.close()
    params, pos, vel = make_batch(nb, [1])
    b = nb.Batch(params, pos, vel)
    b.step(5)
    x = b.positions()
    assert np.isfinite(np.stack(x)).all()
    b.close()


def sym_batch_configs(nb):
    return [c for c in range(nb.batch_config_count())
            if nb.batch_config_info(c)["scheme"] == "symmetric"]


def forward_twin(nb, bcfg):
    """Single-system symmetric shape with the batch shape's super-tile and one
    super-tile pair per work item."""
    b = nb.batch_config_info(bcfg)
    for c in range(nb.forward_config_count()):
        f = nb.forward_config_info(c)
        if (f["scheme"], f["unroll"], f["tile"], f["split"]) == ("symmetric", b["unroll"],
                                                                  b["tile"], 1):
            return c
    raise AssertionError(f"no forward twin of {b}")


SYM_SIZES = [1, 100, 1023, 1024, 1025, 3000, 0, 2048, 5000, 9000]


def test_symmetric_batch_equals_single_system(nb, gpu):
    """Every system of a batched symmetric step equals the single-system
    symmetric step (same super-tile pairs, same partial order), bit for bit;
    sizes around the super-tile length and an empty system."""
    import torch

    params, pos, vel = make_batch(nb, SYM_SIZES)
    for bcfg in sym_batch_configs(nb):
        fcfg = forward_twin(nb, bcfg)
        b = nb.Batch(params, pos, vel)
        b.set_config(bcfg)
        b.step(1)
        bp, bv = b.positions(), b.velocities()
        b.close()
        off = 0
        for p, n in zip(params, SYM_SIZES):
            if n:
                dp = [torch.from_numpy(a[off:off + n].copy()).to(gpu) for a in pos]
                dv = [torch.from_numpy(a[off:off + n].copy()).to(gpu) for a in vel]
                dn = [torch.empty_like(t) for t in dp]
                nb.forward_step(p, dp, dv, dn, cfg=fcfg)
                torch.cuda.synchronize()
                for c in range(3):
                    assert np.array_equal(bp[c][off:off + n], dn[c].cpu().numpy()), (bcfg, n)
                    assert np.array_equal(bv[c][off:off + n], dv[c].cpu().numpy()), (bcfg, n)
            off += n


def test_symmetric_batch_graph_replays(nb, gpu):
    """Captured two-step graphs (pair kernel + finalize, twice) equal direct
    launches on every replay, with eager reads between replays."""
    params, pos, vel = make_batch(nb, SYM_SIZES)
    cfg = sym_batch_configs(nb)[0]
    g = nb.Batch(params, pos, vel)
    d = nb.Batch(params, pos, vel)
    g.set_config(cfg)
    d.set_config(cfg)
    for rep in range(4):
        g.step(2)
        d.step_direct(2)
        for a, b in zip(g.positions() + g.velocities(), d.positions() + d.velocities()):
            assert np.array_equal(a, b), rep
    g.close()
    d.close()


def test_symmetric_batch_vs_oracle(nb, orc, gpu):
    params, pos, vel = make_batch(nb, [700, 5000])
    b = nb.Batch(params, pos, vel)
    b.set_config(sym_batch_configs(nb)[0])
    b.step(1)
    bp = b.positions()
    b.close()
    off, n = 700, 5000
    p = orc.params(G=params[1].G, dt=params[1].dt, num_particles=n)
    rp, _ = orc.step_f32(p, [a[off:] for a in pos], [a[off:] for a in vel])
    got = np.stack([a[off:] for a in bp]).astype(np.float64)
    ref = np.stack(rp).astype(np.float64)
    e = np.abs(got - ref).max(0) / np.maximum(np.abs(ref).max(0), 1.0)
    print(f"symmetric batch vs oracle: max {e.max():.2e} p99 {np.percentile(e, 99):.2e}")
    assert np.percentile(e, 99) <= 1e-5 and e.max() <= 1e-4


def test_symmetric_batch_vanishing_eps(nb, gpu):
    """A system with dist_eps = 0 in a symmetric batch: self pairs masked,
    results equal to the single-system symmetric step bit for bit."""
    import torch

    sizes = [300, 1500]
    params, pos, vel = make_batch(nb, sizes)
    params[1].dist_eps = 0.0
    for bcfg in sym_batch_configs(nb):
        fcfg = forward_twin(nb, bcfg)
        b = nb.Batch(params, pos, vel)
        b.set_config(bcfg)
        b.step(1)
        bp, bv = b.positions(), b.velocities()
        b.close()
        off = 0
        for p, n in zip(params, sizes):
            dp = [torch.from_numpy(a[off:off + n].copy()).to(gpu) for a in pos]
            dv = [torch.from_numpy(a[off:off + n].copy()).to(gpu) for a in vel]
            dn = [torch.empty_like(t) for t in dp]
            nb.forward_step(p, dp, dv, dn, cfg=fcfg)
            torch.cuda.synchronize()
            for c in range(3):
                assert np.isfinite(bp[c][off:off + n]).all(), (bcfg, n)
                assert np.array_equal(bp[c][off:off + n], dn[c].cpu().numpy()), (bcfg, n)
                assert np.array_equal(bv[c][off:off + n], dv[c].cpu().numpy()), (bcfg, n)
            off += n
