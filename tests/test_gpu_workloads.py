"""The benchmark workloads at their full sizes, checked on the GPU.

Batched uneven systems: 64 independent Plummer systems (scale radius 1,
G M = 1 per system, seeds 1000 + s), N_s drawn uniformly from
{1024, 1280, ..., 32768} (seed 0), 200 steps of the autotuned persistent
launch replayed from a hipGraph (the batched benchmark's exact setup).
Each system is one particle_interaction problem (src/simulator.cu:186-229).
  * after the first step every system equals, bit for bit, the single-system
    step of the same decomposition and summation order;
  * the smallest, the largest and two middle systems match the
    reference-order fp32 oracle on 2048-target samples (per-step bar);
  * after 200 steps every state is finite and each system's momentum equals
    the damping-only prediction damping^200 P_0 (the pair terms cancel in
    pairs) to 1e-6 of its sum of |m v|.

Differentiable trajectory: N = 16384 reference disk, T = 200 steps,
checkpoints every 20, loss = mean over particles of |x_T - target|^2 with
the target a forward run from initial positions perturbed by Gaussian
sigma = 1e-3.  The checkpointed adjoint's gradient against the fp64 CPU
adjoint swept along the GPU's own fp32 states (bar 1e-3 normwise).  With
eps = 1e-7 the 200-step system is chaotic and no finite-difference or
independent-trajectory gradient is meaningful, so those two checks run on
the conditioned problem (eps = 1e-2, a target 10 units away, T = 40) at the
same N: the gradient against an independent fp64 CPU trajectory + adjoint
(bar 1e-3) and against central differences of fp32 GPU trajectories
(h = 1e-3, 32 coordinates; the fp32 rounding of x_T limits those to about
2e-2, bar 5e-2).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ----------------------------------------------------------------------------
# batched uneven systems


def plummer_batch(nb):
    sizes = nb.plummer_batch_sizes(64, seed=0)
    params, pos, vel = [], [[], [], []], [[], [], []]
    for k, m in enumerate(sizes):
        p, v = nb.plummer_ic(m, seed=1000 + k, a=1.0, g_mass=1.0)
        params.append(nb.default_params(num_particles=m, G=1.0 / m))
        for c in range(3):
            pos[c].append(p[c])
            vel[c].append(v[c])
    cat = lambda a: tuple(np.ascontiguousarray(np.concatenate(x), np.float32) for x in a)
    return sizes, params, cat(pos), cat(vel)


def forward_twin(nb, bcfg):
    """Single-system forward shape with the batch shape's decomposition and
    per-particle summation order: symmetric batch shapes take one super-tile
    pair per work item, target-sharded ones sum all sources in tile order
    (no source split)."""
    b = nb.batch_config_info(bcfg)
    for c in range(nb.forward_config_count()):
        f = nb.forward_config_info(c)
        if f["scheme"] != b["scheme"] or f["split"] != 1:
            continue
        if (f["unroll"], f["tile"]) != (b["unroll"], b["tile"]):
            continue
        if b["scheme"] == "targets" and f["block"] != b["block"]:
            continue
        return c
    raise AssertionError(f"no single-system twin of batch shape {b}")


def test_batched_64_plummer_systems(nb, orc, gpu):
    import torch

    sizes, params, pos, vel = plummer_batch(nb)
    assert len(sizes) == 64 and min(sizes) >= 1024 and max(sizes) <= 32768
    off = np.concatenate([[0], np.cumsum(sizes)])
    b = nb.Batch(params, pos, vel)
    best = b.autotune(2)
    info = nb.batch_config_info(best)
    print(f"batch of 64 systems ({off[-1]} particles): autotuned {info}", flush=True)
    b.step(1)
    bp, bv = b.positions(), b.velocities()

    # every system: bit-equal to its single-system step of the same shape
    fcfg = forward_twin(nb, best)
    for s, (p, m) in enumerate(zip(params, sizes)):
        sl = slice(off[s], off[s + 1])
        dp = [torch.from_numpy(a[sl].copy()).to(gpu) for a in pos]
        dv = [torch.from_numpy(a[sl].copy()).to(gpu) for a in vel]
        dn = [torch.empty_like(t) for t in dp]
        nb.forward_step(p, dp, dv, dn, cfg=fcfg)
        torch.cuda.synchronize()
        for c in range(3):
            assert np.array_equal(bp[c][sl], dn[c].cpu().numpy()), (s, m, c)
            assert np.array_equal(bv[c][sl], dv[c].cpu().numpy()), (s, m, c)

    # smallest, largest and two middle systems against the oracle (one step)
    order = np.argsort(sizes, kind="stable")
    for s in (order[0], order[-1], order[len(order) // 3], order[2 * len(order) // 3]):
        m = sizes[s]
        sl = slice(off[s], off[s + 1])
        p = orc.params(num_particles=m, G=params[s].G)
        sp = [a[sl] for a in pos]
        sv = [a[sl] for a in vel]
        i0 = max(0, (m - 2048) // 2)  # the middle 2048 targets (all of a smaller system)
        i1 = min(m, i0 + 2048)
        rp, _ = orc.step_f32(p, sp, sv, i0, i1)
        got = np.stack([a[sl][i0:i1] for a in bp]).astype(np.float64)
        ref = np.stack([a[i0:i1] for a in rp]).astype(np.float64)
        e = np.abs(got - ref).max(0) / np.maximum(np.abs(ref).max(0), 1.0)
        print(f"system {s} (N={m}) targets [{i0},{i1}) vs oracle: "
              f"max {e.max():.2e} p99 {np.percentile(e, 99):.2e}", flush=True)
        assert np.percentile(e, 99) <= 1e-5 and e.max() <= 1e-4, (s, m)

    # 199 more steps (200 in all) through the graph-replayed launch
    b.step(199)
    xp, xv = b.positions(), b.velocities()
    b.close()
    assert all(np.isfinite(a).all() for a in xp + xv)
    damp = float(np.float32(params[0].damping)) ** 200
    worst = 0.0
    for s, (p, m) in enumerate(zip(params, sizes)):
        sl = slice(off[s], off[s + 1])
        st = lambda arrs: torch.from_numpy(np.stack([a[sl] for a in arrs])).to(gpu)
        p0 = np.array(nb.energy_momentum(p, st(pos), st(vel))["momentum"])
        p1 = np.array(nb.energy_momentum(p, st(xp), st(xv))["momentum"])
        scale = float(np.abs(np.stack([a[sl] for a in xv]).astype(np.float64)).sum())
        drift = float(np.abs(p1 - damp * p0).max()) / scale
        worst = max(worst, drift)
        assert drift <= 1e-6, (s, m, drift)
    print(f"200 steps: largest momentum drift / sum|v| over the 64 systems {worst:.2e}")


# ----------------------------------------------------------------------------
# differentiable trajectory


def disk_problem(nb, torch, gpu, n, steps, every, eps=None, far=False):
    pos, vel = nb.disk_ic(n)
    x0 = torch.from_numpy(np.stack(pos)).to(gpu)
    v0 = torch.from_numpy(np.stack(vel)).to(gpu)
    params = nb.default_params(num_particles=n)
    if eps is not None:
        params.dist_eps = eps
    if far:
        return x0, v0, params, (x0 + 10.0).contiguous()
    rng = np.random.default_rng(0)
    pert = torch.from_numpy(rng.normal(scale=1e-3, size=(3, n)).astype(np.float32)).to(gpu)
    with torch.no_grad():
        target, _ = nb.trajectory(x0 + pert, v0, params, steps, every)
    return x0, v0, params, target


def loss_and_grad(nb, x0, v0, params, target, steps, every):
    x = x0.clone().requires_grad_(True)
    v = v0.clone().requires_grad_(True)
    xT, _ = nb.trajectory(x, v, params, steps, every)
    loss = ((xT.double() - target.double()) ** 2).sum(0).mean()
    loss.backward()
    return (loss.item(), x.grad.double().cpu().numpy(), v.grad.double().cpu().numpy(),
            xT.detach())


def test_trajectory_16384_t200_vs_fp64_adjoint(nb, orc, gpu):
    import torch

    n, T, K = 16384, 200, 20
    x0, v0, params, target = disk_problem(nb, torch, gpu, n, T, K)
    L, gx, gv, xT = loss_and_grad(nb, x0, v0, params, target, T, K)
    # the GPU's own states: one forward launch per step, the trajectory's shape
    P, V, states = x0.clone(), v0.clone(), [x0.cpu().numpy()]
    for _ in range(T):
        N = torch.empty_like(P)
        nb.forward_step(params, list(P), list(V), list(N))
        P = N
        states.append(P.cpu().numpy())
    assert np.array_equal(states[-1], xT.cpu().numpy())  # the same trajectory
    tgt = target.double().cpu().numpy()
    lx = 2.0 * (states[-1].astype(np.float64) - tgt) / n
    lv = np.zeros_like(lx)
    p = orc.params()
    for t in range(T - 1, -1, -1):
        lx, lv = orc.adjoint_step_f64(p, states[t], lx, lv)
        lx, lv = np.stack(lx), np.stack(lv)
    ex, ev = rel(gx, lx), rel(gv, lv)
    print(f"N={n} T={T} K={K}: loss {L:.6e}; gradient vs the fp64 adjoint along the "
          f"same states: x {ex:.2e} v {ev:.2e}")
    assert ex <= 1e-3 and ev <= 1e-3


def test_trajectory_16384_conditioned(nb, orc, gpu):
    import torch

    n, T, K = 16384, 40, 20
    x0, v0, params, target = disk_problem(nb, torch, gpu, n, T, K, eps=1e-2, far=True)
    L, gx, gv, _ = loss_and_grad(nb, x0, v0, params, target, T, K)
    tgt = target.double().cpu().numpy()
    L64, gx64, gv64 = orc.traj_grad_f64(orc.params(dist_eps=1e-2), x0.cpu().numpy(),
                                        v0.cpu().numpy(), tgt, T)
    ex, ev = rel(gx, gx64), rel(gv, gv64)
    print(f"N={n} T={T} eps=1e-2, far target: loss {L:.6e} (fp64 {L64:.6e}); gradient vs "
          f"an independent fp64 trajectory: x {ex:.2e} v {ev:.2e}", flush=True)
    assert abs(L - L64) <= 1e-4 * L64
    assert ex <= 1e-3 and ev <= 1e-3

    # central differences of the fp64 loss of fp32 GPU trajectories
    target64 = target.double()

    def loss_of(x, v):
        with torch.no_grad():
            xT, _ = nb.trajectory(x, v, params, T, K)
        return float(((xT.double() - target64) ** 2).sum(0).mean())

    rng = np.random.default_rng(3)
    h = 1e-3
    fd, an = [], []
    for k in range(32):
        c, i = int(rng.integers(3)), int(rng.integers(n))
        which = k % 2
        base = (x0, v0)
        plus = [t.clone() for t in base]
        minus = [t.clone() for t in base]
        plus[which][c, i] += h
        minus[which][c, i] -= h
        hp = float(plus[which][c, i] - base[which][c, i])
        hm = float(base[which][c, i] - minus[which][c, i])
        fd.append((loss_of(*plus) - loss_of(*minus)) / (hp + hm))
        an.append(gx[c, i] if which == 0 else gv[c, i])
    err = rel(np.array(an), np.array(fd))
    print(f"central differences (h={h}) on 32 coordinates: rel err {err:.2e}")
    assert err <= 5e-2
