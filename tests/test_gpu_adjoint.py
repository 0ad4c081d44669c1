"""GPU adjoint kernels and the checkpointed trajectory gradient, against the
fp64 CPU adjoint of the oracle (itself pinned by central differences in
test_oracle.py) and against central finite differences of an fp64 loss.

Gradient error metric: ||g_gpu - g_fp64|| / ||g_fp64|| (2-norm over all
particles and components); target <= 1e-3.

Two kinds of reference:
  * the fp64 adjoint evaluated along the GPU's own fp32 states (the same
    trajectory, the same residual): this measures the adjoint's arithmetic,
    and is the comparison for the target-matching loss of the differentiable
    trajectory (a residual x_T - target of ~1e-3 on positions of ~50 is
    changed by fp32 rounding of the trajectory itself, so an independent
    fp64 trajectory gives a different, equally valid gradient);
  * an independent fp64 trajectory and central differences, for a loss whose
    residual is far above fp32 rounding (target 10 units away).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def to_dev(torch, arrs, gpu):
    return torch.from_numpy(np.ascontiguousarray(np.stack(arrs), np.float32)).to(gpu)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def sym_adjoint_configs(nb):
    return [c for c in range(nb.adjoint_config_count())
            if nb.adjoint_config_info(c)["scheme"] in ("symmetric", "symmetric_tiles")]


def matching_forward_config(nb, acfg):
    """The forward shape with the adjoint shape's decomposition (same scheme,
    unroll, tile and split): the fused step's forward part equals its bits."""
    a = nb.adjoint_config_info(acfg)
    for c in range(nb.forward_config_count()):
        if nb.forward_config_info(c) == a:
            return c
    raise AssertionError(f"no forward shape like {a}")


@pytest.mark.parametrize("which", ["default", "symmetric"])
def test_adjoint_pair_sum_vs_fp64(nb, orc, gpu, which):
    """With dt = G = 1 and lam_x' = 0 the kernel's lam_x is the pair sum
    J^T g itself; compare it per particle with the fp64 sum (3000 particles:
    a partial last tile for every shape), for the library's shape and for
    every symmetric shape."""
    import torch

    cfgs = [-1] if which == "default" else sym_adjoint_configs(nb)
    n = 3000
    pos, _ = nb.plummer_ic(n, seed=21, g_mass=1.0)
    rng = np.random.default_rng(2)
    g = rng.normal(size=(3, n)).astype(np.float32)
    params = nb.default_params(num_particles=n, dt=1.0, G=1.0, damping=0.5)
    p = orc.params(dt=1.0, G=1.0, damping=0.5)
    ref_lx, ref_lv = orc.adjoint_step_f64(p, pos, np.zeros((3, n)), g.astype(np.float64))
    ref = np.stack(ref_lx)
    P = to_dev(torch, pos, gpu)
    lxn = torch.zeros(3, n, device=gpu)
    lvn = torch.from_numpy(g).to(gpu)
    for cfg in cfgs:
        lx = torch.empty(3, n, device=gpu)
        lv = torch.empty(3, n, device=gpu)
        nb.adjoint_step(params, P, lxn, lvn, lx, lv, cfg=cfg)
        torch.cuda.synchronize()
        got = lx.cpu().numpy().astype(np.float64)
        e = np.abs(got - ref).max(0) / np.abs(ref).max()
        info = nb.adjoint_config_info(cfg) if cfg >= 0 else "default"
        print(f"J^T g vs fp64 ({info}): max {e.max():.2e} normwise {rel(got, ref):.2e}")
        assert rel(got, ref) < 1e-5 and e.max() < 1e-4, info
        assert np.array_equal(lv.cpu().numpy(), np.float32(0.5) * g), info


def test_adjoint_vanishing_eps(nb, orc, gpu):
    """dist_eps = 0: every symmetric adjoint shape (and the fused step)
    masks the self pairs; the pair sum matches the fp64 adjoint, which skips
    j == k."""
    import torch

    n = 3000
    pos, vel = nb.plummer_ic(n, seed=23, g_mass=1.0)
    rng = np.random.default_rng(5)
    g = rng.normal(size=(3, n)).astype(np.float32)
    params = nb.default_params(num_particles=n, dt=1.0, G=1.0, damping=0.5, dist_eps=0.0)
    p = orc.params(dt=1.0, G=1.0, damping=0.5, dist_eps=0.0)
    ref_lx, _ = orc.adjoint_step_f64(p, pos, np.zeros((3, n)), g.astype(np.float64))
    ref = np.stack(ref_lx)
    P = to_dev(torch, pos, gpu)
    V = to_dev(torch, vel, gpu)
    lxn = torch.zeros(3, n, device=gpu)
    lvn = torch.from_numpy(g).to(gpu)
    for cfg in [-1] + sym_adjoint_configs(nb):
        for fused in (False, True):
            lx = torch.empty(3, n, device=gpu)
            lv = torch.empty(3, n, device=gpu)
            if fused:
                nb.fwd_adjoint_step(params, P, V.clone(), torch.empty_like(P), lxn, lvn,
                                    lx, lv, cfg=cfg)
            else:
                nb.adjoint_step(params, P, lxn, lvn, lx, lv, cfg=cfg)
            torch.cuda.synchronize()
            got = lx.cpu().numpy().astype(np.float64)
            what = (cfg, fused)
            assert np.isfinite(got).all(), what
            assert rel(got, ref) < 1e-5, (what, rel(got, ref))


@pytest.mark.parametrize("eps", ["default", 0.0])
def test_adjoint_coincident_particles(nb, orc, gpu, eps):
    """Distinct particles at one position (3001 particles; pairs within a
    tile, across tiles, in the last partial tile, and a triple), every
    adjoint shape, plain and fused.  With SimParam's eps a coincident pair
    adds K (g_j - g_k), K = eps^-3/2: the pair sum matches the fp64 adjoint.
    With eps = 0 the coincident pairs give NaN (d = 0), for exactly those
    particles, in the fp64 adjoint and in every shape; the others match."""
    import torch

    n = 3001
    pos, vel = nb.disk_ic(n)
    pos = [a.copy() for a in pos]
    for dst, src in ((1, 0), (1500, 5), (3000, 2999), (2990, 100), (701, 700), (2000, 700)):
        for a in pos:
            a[dst] = a[src]
    kw = {} if eps == "default" else {"dist_eps": eps}
    rng = np.random.default_rng(11)
    g = rng.normal(size=(3, n)).astype(np.float32)
    params = nb.default_params(num_particles=n, dt=1.0, G=1.0, damping=0.5, **kw)
    p = orc.params(dt=1.0, G=1.0, damping=0.5, **kw)
    with np.errstate(invalid="ignore", divide="ignore"):
        ref_lx, _ = orc.adjoint_step_f64(p, pos, np.zeros((3, n)), g.astype(np.float64))
    ref = np.stack(ref_lx)
    ref_nan = ~np.isfinite(ref).all(0)
    dup = np.zeros(n, bool)
    dup[[1, 0, 1500, 5, 3000, 2999, 2990, 100, 701, 700, 2000]] = True
    assert np.array_equal(ref_nan, dup if eps == 0.0 else np.zeros(n, bool))
    # the coincident pairs' K = eps^-3/2 dominates their particles' sums:
    # those and the others are compared separately
    groups = [~dup] if eps == 0.0 else [~dup, dup]
    P = to_dev(torch, pos, gpu)
    V = to_dev(torch, vel, gpu)
    lxn = torch.zeros(3, n, device=gpu)
    lvn = torch.from_numpy(g).to(gpu)
    for cfg in range(nb.adjoint_config_count()):
        for fused in (False, True):
            lx = torch.empty(3, n, device=gpu)
            lv = torch.empty(3, n, device=gpu)
            if fused:
                nb.fwd_adjoint_step(params, P, V.clone(), torch.empty_like(P), lxn, lvn,
                                    lx, lv, cfg=cfg)
            else:
                nb.adjoint_step(params, P, lxn, lvn, lx, lv, cfg=cfg)
            torch.cuda.synchronize()
            got = lx.cpu().numpy().astype(np.float64)
            what = (eps, cfg, fused, nb.adjoint_config_info(cfg))
            assert np.array_equal(~np.isfinite(got).all(0), ref_nan), what
            for sel in groups:
                e = rel(got[:, sel], ref[:, sel])
                assert e < 1e-5, (what, e)


@pytest.mark.parametrize("n", [4096, 33000])
def test_adjoint_step_default_params(nb, orc, gpu, n):
    """SimParam defaults, the library's shape (tile pairs at 4096, super-tile
    pairs at 33000), against the fp64 adjoint."""
    import torch

    pos, _ = nb.disk_ic(n)
    rng = np.random.default_rng(3)
    lxn = rng.normal(size=(3, n)).astype(np.float32)
    lvn = rng.normal(size=(3, n)).astype(np.float32)
    params = nb.default_params(num_particles=n)
    outs = [torch.empty(3, n, device=gpu) for _ in range(2)]
    nb.adjoint_step(params, to_dev(torch, pos, gpu), torch.from_numpy(lxn).to(gpu),
                    torch.from_numpy(lvn).to(gpu), *outs)
    torch.cuda.synchronize()
    ref_lx, ref_lv = orc.adjoint_step_f64(orc.params(), pos, lxn, lvn)
    # compare the pair-sum part: (lam_x - lam_x') / (dt G)
    c = float(np.float32(0.005)) * 2.0
    got = (outs[0].cpu().numpy().astype(np.float64) - lxn) / c
    ref = (np.stack(ref_lx) - lxn) / c
    print(f"default params: pair part normwise {rel(got, ref):.2e}")
    assert rel(got, ref) < 1e-4
    assert rel(outs[1].cpu().numpy(), np.stack(ref_lv)) < 1e-6


@pytest.mark.parametrize("shape", ["targets", "symmetric", "symmetric_tiles"])
@pytest.mark.parametrize("n", [1, 2, 257, 1500, 5000])
def test_fused_step_equals_separate(nb, gpu, n, shape):
    """Forward part bit-equal to the forward kernel of the same decomposition;
    adjoint part equal to the plain adjoint kernel's up to the rounding of
    its inverse-cube and 1/d (formed from rsqrt(d) there, from rsqrt(d^3)
    here); every shape of each scheme."""
    acfgs = [0] if shape == "targets" else [
        c for c in sym_adjoint_configs(nb) if nb.adjoint_config_info(c)["scheme"] == shape]
    for acfg in acfgs:
        fused_equals_separate(nb, gpu, n, acfg)


def fused_equals_separate(nb, gpu, n, acfg):
    import torch

    fcfg = matching_forward_config(nb, acfg)
    pos, vel = nb.plummer_ic(n, seed=n, g_mass=1.0)
    params = nb.default_params(num_particles=n, G=1.0 / n)
    rng = np.random.default_rng(n)
    lxn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(gpu)
    lvn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(gpu)
    P = to_dev(torch, pos, gpu)
    V1 = to_dev(torch, vel, gpu)
    V2 = V1.clone()
    N1 = torch.empty_like(P)
    N2 = torch.empty_like(P)
    lx1, lv1, lx2, lv2 = [torch.empty(3, n, device=gpu) for _ in range(4)]
    nb.forward_step(params, list(P), list(V1), list(N1), cfg=fcfg)
    nb.adjoint_step(params, P, lxn, lvn, lx1, lv1, cfg=acfg)
    nb.fwd_adjoint_step(params, P, V2, N2, lxn, lvn, lx2, lv2, cfg=acfg)
    torch.cuda.synchronize()
    for a, b in ((V1, V2), (N1, N2), (lv1, lv2)):
        assert torch.equal(a, b), acfg
    # the plain adjoint forms K and 1/d from one rsqrt(d), the fused one as
    # the forward kernel does: equal up to that rounding
    d = (lx1 - lx2).double().norm() / max(float(lx1.double().norm()), 1e-30)
    assert d < 1e-6, (acfg, float(d))


@pytest.mark.parametrize("unroll", [2, 4])
def test_banded_adjoint_past_one_million(nb, orc, gpu, unroll):
    """N = 1310720 + 777 (2562 super-tiles of 512 in five bands of source
    super-tiles with 2 per lane, the library's shape; 1281 super-tiles of
    1024 in two bands with 4 per lane), 4 super-tiles per item: the adjoint
    and its forward twin run in bands.  With dt = G = 1 and lam_x' = 0,
    lam_x is the pair sum J^T g: checked against the target-sharded adjoint
    (itself checked against the fp64 oracle at 3000) and by its total,
    sum_k sum_j M_kj (g_j - g_k) = 0, which a lost or mis-signed reaction
    breaks.  The fused step: forward part bit-equal to the banded forward
    step, whose forces net to zero and are as close to the oracle's fp64
    forces as the target-sharded step's."""
    import torch

    n = 1310720 + 777
    band = next(c for c in sym_adjoint_configs(nb)
                if (nb.adjoint_config_info(c)["scheme"], nb.adjoint_config_info(c)["unroll"],
                    nb.adjoint_config_info(c)["split"]) == ("symmetric", unroll, 4))
    targets = next(c for c in range(nb.adjoint_config_count())
                   if nb.adjoint_config_info(c)["scheme"] == "targets"
                   and nb.adjoint_config_info(c)["split"] > 1)
    ftargets = next(c for c in range(nb.forward_config_count())
                    if nb.forward_config_info(c) == nb.adjoint_config_info(targets))
    pos, _ = nb.disk_ic(n)
    params = nb.default_params(num_particles=n, dt=1.0, G=1.0, damping=0.5)
    rng = np.random.default_rng(7)
    P = to_dev(torch, pos, gpu)
    lxn = torch.zeros(3, n, device=gpu)
    lvn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(gpu)
    out = {}
    for name, cfg in (("default", -1), ("band", band), ("targets", targets)):
        lx, lv = torch.empty(3, n, device=gpu), torch.empty(3, n, device=gpu)
        print(f"adjoint N={n} {name} cfg {cfg}", flush=True)
        nb.adjoint_step(params, P, lxn, lvn, lx, lv, cfg=cfg)
        torch.cuda.synchronize()
        out[name] = lx
    assert torch.equal(out["default"], out["band"]) == (unroll == 2)
    got, ref = out["band"].double(), out["targets"].double()
    e = float((got - ref).norm() / ref.norm())
    scale = got.abs().sum(1)
    net = (got.sum(1).abs() / scale).cpu().numpy()
    net_t = (ref.sum(1).abs() / scale).cpu().numpy()
    print(f"banded vs targets {e:.2e}; net / sum|.| banded {net}, targets {net_t}")
    assert e < 1e-5
    assert np.all(net <= 1e-5) and np.all(net <= 2 * net_t + 1e-7)
    # fused step and the forward steps (zero velocities: vel' = dt G F)
    fcfg = matching_forward_config(nb, band)
    V = {k: torch.zeros(3, n, device=gpu) for k in ("fwd", "fused", "targets")}
    N = {k: torch.empty(3, n, device=gpu) for k in V}
    print(f"forward N={n} cfg {fcfg} / {ftargets}, fused cfg {band}", flush=True)
    nb.forward_step(params, list(P), list(V["fwd"]), list(N["fwd"]), cfg=fcfg)
    nb.forward_step(params, list(P), list(V["targets"]), list(N["targets"]), cfg=ftargets)
    lx2, lv2 = torch.empty(3, n, device=gpu), torch.empty(3, n, device=gpu)
    nb.fwd_adjoint_step(params, P, V["fused"], N["fused"], lxn, lvn, lx2, lv2, cfg=band)
    torch.cuda.synchronize()
    assert torch.equal(V["fwd"], V["fused"]) and torch.equal(N["fwd"], N["fused"])
    d = float((lx2.double() - got).norm() / got.norm())
    assert d < 1e-6, d
    # the disk's net forces cancel deeply (fp32 sums differ by 1e-3 of |F|
    # for some particles): both schemes against fp64 forces on 512 targets,
    # the first and the last (partial) super-tile's
    F = V["fwd"].double()
    fscale = F.abs().sum(1)
    fnet = (F.sum(1).abs() / fscale).cpu().numpy()
    idx = np.concatenate([np.arange(256), np.arange(n - 256, n)])
    f64 = np.stack(orc.forces_f64(pos, idx))
    err = {}
    for k in ("fwd", "targets"):
        g = V[k].cpu().numpy()[:, idx].astype(np.float64)
        err[k] = np.linalg.norm(g - f64, axis=0) / np.linalg.norm(f64, axis=0)
    print(f"force err vs fp64: banded p50 {np.median(err['fwd']):.2e} max "
          f"{err['fwd'].max():.2e}; targets p50 {np.median(err['targets']):.2e} max "
          f"{err['targets'].max():.2e}; net force / sum|F| {fnet}; fused lx {d:.2e}")
    assert np.all(fnet <= 1e-5)
    assert np.median(err["fwd"]) <= 2 * np.median(err["targets"]) + 1e-7
    assert err["fwd"].max() <= 2 * err["targets"].max() + 1e-6


def test_every_adjoint_shape(nb, gpu):
    import torch

    n = 5000
    pos, _ = nb.disk_ic(n)
    rng = np.random.default_rng(5)
    lxn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(gpu)
    lvn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(gpu)
    params = nb.default_params(num_particles=n)
    P = to_dev(torch, pos, gpu)
    ref = None
    for cfg in range(nb.adjoint_config_count()):
        lx, lv = torch.empty(3, n, device=gpu), torch.empty(3, n, device=gpu)
        nb.adjoint_step(params, P, lxn, lvn, lx, lv, cfg=cfg)
        torch.cuda.synchronize()
        got = (lx - lxn).double().cpu().numpy()
        if ref is None:
            ref = got
        assert rel(got, ref) < 1e-5, (cfg, nb.adjoint_config_info(cfg))


def disk_problem(nb, orc, torch, gpu, n, steps, every, eps=None, far=False):
    (pos, vel) = nb.disk_ic(n)
    x0 = to_dev(torch, pos, gpu)
    v0 = to_dev(torch, vel, gpu)
    params = nb.default_params(num_particles=n)
    if eps is not None:
        params.dist_eps = eps
    if far:  # residual ~10: far above fp32 rounding of the positions
        return x0, v0, params, (x0 + 10.0).contiguous()
    rng = np.random.default_rng(7)
    pert = torch.from_numpy(rng.normal(scale=1e-3, size=(3, n)).astype(np.float32)).to(gpu)
    with torch.no_grad():
        target, _ = nb.trajectory(x0 + pert, v0, params, steps, every)
    return x0, v0, params, target


def gpu_gradient(nb, torch, x0, v0, params, target, steps, every):
    x = x0.clone().requires_grad_(True)
    v = v0.clone().requires_grad_(True)
    xT, _ = nb.trajectory(x, v, params, steps, every)
    loss = ((xT.double() - target.double()) ** 2).sum(0).mean()
    loss.backward()
    return loss.item(), x.grad.double().cpu().numpy(), v.grad.double().cpu().numpy(), xT


def gpu_states(nb, torch, params, x0, v0, steps):
    """Positions x_0 .. x_T from the forward kernel, one launch per step (the
    launch shape the trajectory's forward pass uses)."""
    P, V = x0.clone(), v0.clone()
    states = [P.cpu().numpy()]
    for _ in range(steps):
        N = torch.empty_like(P)
        nb.forward_step(params, list(P), list(V), list(N))
        P = N
        states.append(P.cpu().numpy())
    return states


def fp64_adjoint_along(orc, p, states, target):
    """fp64 reverse sweep through the given (fp32) states."""
    n = states[0].shape[1]
    lx = 2.0 * (states[-1].astype(np.float64) - target) / n
    lv = np.zeros_like(lx)
    for t in range(len(states) - 2, -1, -1):
        lx, lv = orc.adjoint_step_f64(p, states[t], lx, lv)
        lx, lv = np.stack(lx), np.stack(lv)
    return lx, lv


def test_trajectory_gradient_vs_fp64_same_states(nb, orc, gpu):
    """The differentiable-trajectory loss at SimParam defaults: the GPU's
    checkpointed adjoint against the fp64 adjoint along the same states."""
    import torch

    n, T, K = 2048, 40, 10
    x0, v0, params, target = disk_problem(nb, orc, torch, gpu, n, T, K)
    L, gx, gv, xT = gpu_gradient(nb, torch, x0, v0, params, target, T, K)
    states = gpu_states(nb, torch, params, x0, v0, T)
    assert np.array_equal(states[-1], xT.detach().cpu().numpy())  # same trajectory
    gx64, gv64 = fp64_adjoint_along(orc, orc.params(), states, target.double().cpu().numpy())
    ex, ev = rel(gx, gx64), rel(gv, gv64)
    print(f"N={n} T={T} K={K}: loss {L:.6e}; grad vs fp64 adjoint on the same states: "
          f"x {ex:.2e} v {ev:.2e}")
    assert ex <= 1e-3 and ev <= 1e-3


def test_trajectory_gradient_zero_eps(nb, orc, gpu):
    """dist_eps = 0 through the whole differentiable trajectory (symmetric
    forward and adjoint shapes, self pairs masked): the checkpointed adjoint
    against the fp64 adjoint (which skips j == k) along the same states."""
    import torch

    n, T, K = 4096, 10, 5
    x0, v0, params, target = disk_problem(nb, orc, torch, gpu, n, T, K, eps=0.0)
    L, gx, gv, xT = gpu_gradient(nb, torch, x0, v0, params, target, T, K)
    states = gpu_states(nb, torch, params, x0, v0, T)
    assert np.array_equal(states[-1], xT.detach().cpu().numpy())
    assert np.isfinite(gx).all() and np.isfinite(gv).all()
    gx64, gv64 = fp64_adjoint_along(orc, orc.params(dist_eps=0.0), states,
                                    target.double().cpu().numpy())
    ex, ev = rel(gx, gx64), rel(gv, gv64)
    print(f"N={n} T={T} K={K} eps=0: loss {L:.6e}; grad vs fp64 adjoint on the same "
          f"states: x {ex:.2e} v {ev:.2e}")
    assert ex <= 1e-3 and ev <= 1e-3


def test_trajectory_gradient_vs_fp64_trajectory(nb, orc, gpu):
    """Independent fp64 trajectory + adjoint (every state kept), for a loss
    whose residual is far above fp32 rounding."""
    import torch

    n, T, K = 2048, 40, 10
    x0, v0, params, target = disk_problem(nb, orc, torch, gpu, n, T, K, eps=1e-2, far=True)
    L, gx, gv, _ = gpu_gradient(nb, torch, x0, v0, params, target, T, K)
    L64, gx64, gv64 = orc.traj_grad_f64(orc.params(dist_eps=1e-2), x0.cpu().numpy(),
                                        v0.cpu().numpy(), target.double().cpu().numpy(), T)
    ex, ev = rel(gx, gx64), rel(gv, gv64)
    print(f"N={n} T={T} K={K} eps=1e-2, far target: loss {L:.6e} (fp64 {L64:.6e}); "
          f"grad rel err x {ex:.2e} v {ev:.2e}")
    assert abs(L - L64) <= 1e-4 * L64
    assert ex <= 1e-3 and ev <= 1e-3


def test_checkpoint_interval_does_not_change_gradient(nb, orc, gpu):
    import torch

    n, T = 1024, 24
    x0, v0, params, target = disk_problem(nb, orc, torch, gpu, n, T, 5)
    grads = []
    for K in (1, 5, 7, 24, 100):
        x = x0.clone().requires_grad_(True)
        v = v0.clone().requires_grad_(True)
        xT, _ = nb.trajectory(x, v, params, T, K)
        ((xT - target) ** 2).sum(0).mean().backward()
        grads.append((x.grad.clone(), v.grad.clone()))
    for gx, gv in grads[1:]:  # replayed states are bit-identical
        assert torch.equal(gx, grads[0][0]) and torch.equal(gv, grads[0][1])


def test_trajectory_gradient_banded_size(nb, gpu):
    """N = 300001, T = 3, checkpoints every 2: the trajectory's reverse sweep
    runs the library's adjoint shape (2 per lane, four super-tiles per item,
    banded).  Its gradient of a random linear loss of (x_T, v_T) equals a
    manual reverse sweep through the same forward states with the
    target-sharded adjoint kernel, to fp32 rounding of the pair sums."""
    import torch

    n, T = 300001, 3
    pos, vel = nb.disk_ic(n)
    x0 = to_dev(torch, pos, gpu)
    v0 = to_dev(torch, vel, gpu)
    params = nb.default_params(num_particles=n)
    rng = np.random.default_rng(13)
    wx = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(gpu)
    wv = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(gpu)
    x = x0.clone().requires_grad_(True)
    v = v0.clone().requires_grad_(True)
    xT, vT = nb.trajectory(x, v, params, T, 2)
    ((xT * wx).sum() + (vT * wv).sum()).backward()
    torch.cuda.synchronize()
    targets = next(c for c in range(nb.adjoint_config_count())
                   if nb.adjoint_config_info(c)["scheme"] == "targets"
                   and nb.adjoint_config_info(c)["split"] > 1)
    P, V, states = x0.clone(), v0.clone(), [x0.clone()]
    for _ in range(T - 1):
        N = torch.empty_like(P)
        nb.forward_step(params, list(P), list(V), list(N))
        P = N
        states.append(P)

    def sweep(p):
        lx, lv = wx.clone(), wv.clone()
        for t in reversed(range(T)):
            lx2, lv2 = torch.empty_like(lx), torch.empty_like(lv)
            nb.adjoint_step(p, states[t], lx, lv, lx2, lv2, cfg=targets)
            lx, lv = lx2, lv2
        torch.cuda.synchronize()
        return lx.double(), lv.double()

    lx, lv = sweep(params)
    # the same sweep without the pair sums (G = 0): the errors are measured
    # against the part of the gradient the pair sums make
    lx0, lv0 = sweep(nb.default_params(num_particles=n, G=0.0))
    ex = float((x.grad.double() - lx).norm() / (lx - lx0).norm())
    ev = float((v.grad.double() - lv).norm() / (lv - lv0).norm())
    print(f"N={n} T={T}: banded vs target-sharded reverse sweep, relative to the "
          f"pair-sum part: x {ex:.2e} v {ev:.2e}")
    assert ex < 1e-4 and ev < 1e-4


def test_trajectory_central_differences(nb, orc, gpu):
    """Adjoint gradient vs central differences (h = 1e-3) of the fp64 loss
    on 32 random coordinates of x0 and v0 (far target, eps = 1e-2)."""
    import torch

    n, T, K = 1024, 30, 10
    x0, v0, params, target = disk_problem(nb, orc, torch, gpu, n, T, K, eps=1e-2, far=True)
    x = x0.clone().requires_grad_(True)
    v = v0.clone().requires_grad_(True)
    xT, _ = nb.trajectory(x, v, params, T, K)
    ((xT.double() - target.double()) ** 2).sum(0).mean().backward()
    gx = x.grad.double().cpu().numpy()
    gv = v.grad.double().cpu().numpy()
    X0 = x0.double().cpu().numpy()
    V0 = v0.double().cpu().numpy()
    tgt = target.double().cpu().numpy()
    p = orc.params(dist_eps=1e-2)
    rng = np.random.default_rng(11)
    h = 1e-3
    fd, an = [], []
    for k in range(32):
        c, i = int(rng.integers(3)), int(rng.integers(n))
        which = k % 2
        Xp, Xm, Vp, Vm = X0.copy(), X0.copy(), V0.copy(), V0.copy()
        if which == 0:
            Xp[c, i] += h
            Xm[c, i] -= h
        else:
            Vp[c, i] += h
            Vm[c, i] -= h
        Lp, _, _ = orc.traj_grad_f64(p, Xp, Vp, tgt, T, grad=False)
        Lm, _, _ = orc.traj_grad_f64(p, Xm, Vm, tgt, T, grad=False)
        fd.append((Lp - Lm) / (2 * h))
        an.append(gx[c, i] if which == 0 else gv[c, i])
    fd, an = np.array(fd), np.array(an)
    err = rel(an, fd)
    print(f"central differences on 32 coordinates: rel err {err:.2e}")
    assert err <= 1e-3
