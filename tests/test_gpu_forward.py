"""GPU parity of the forward step against the CPU oracle (reference-order fp32
restatement of particle_interaction, src/simulator.cu:186-229).

Error metric (per particle): max_c |x_gpu - x_ref| / max(max_c |x_ref|, 1),
reported as max and 99th percentile over particles.

One step from an identical state: the per-step target is 1e-5 against the
oracle; the GPU differs from it only by rounding (hardware rsqrt, fused
multiply-adds, tiled summation order).  Where a particle's force sum is
ill-conditioned (a close encounter with eps = 1e-7) two correct fp32
evaluations can differ by more than that, so every particle where the GPU
and the oracle differ by more than 1e-6 is refereed by the fp64 step from
the same fp32 state: the GPU must be no farther from it than the oracle is
(plus 1e-7 absolute); the 99th percentile of the GPU's own distance from
the fp64 step must stay within 1e-5.

Ten steps from the initial conditions: close encounters amplify rounding
chaotically (the oracle itself ends up to 7e-2 from fp64 at N = 25600), so
the trajectories are refereed by an fp64 trajectory of all particles from
the same initial state: the GPU's max and 99th percentile distance from it
must be no larger than the fp32 oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-5


def per_particle(got, ref):
    g = np.stack(got).astype(np.float64)
    r = np.stack(ref).astype(np.float64)
    return np.abs(g - r).max(0) / np.maximum(np.abs(r).max(0), 1.0)


def fp64_step(orc, p, pos, vel, idx):
    """One step in fp64 for targets idx from the float32 state."""
    F = np.stack(orc.forces_f64(pos, idx, eps=float(np.float32(p.dist_eps)),
                                skip_self=p.calc_method == 0))
    x = np.stack([a[idx] for a in pos]).astype(np.float64)
    v = np.stack([a[idx] for a in vel]).astype(np.float64)
    v1 = float(np.float32(p.damping)) * v + float(np.float32(p.dt)) * float(np.float32(p.G)) * F
    return x + float(np.float32(p.dt)) * v1


def assert_step_parity(orc, p, pos, vel, gpu_pos, ref_pos, idx, label):
    """gpu_pos/ref_pos: positions after one step at targets idx.

    Every particle where the GPU and the fp32 oracle differ by more than 1e-6
    is refereed by the fp64 step: the GPU may be no farther from it than the
    oracle is (+1e-7).  The 99th percentile of the GPU's own error against
    the fp64 step must stay within the per-step target (where the two fp32
    results agree to 1e-6, the oracle's distance is used)."""
    e_go = per_particle(gpu_pos, ref_pos)
    bad = np.nonzero(e_go > 1e-6)[0]
    msg = f"{label}: vs fp32 oracle max {e_go.max():.2e} p99 {np.percentile(e_go, 99):.2e}"
    e_gpu = e_go.copy()  # upper bound of the GPU's distance where they agree
    if bad.size:
        x64 = fp64_step(orc, p, pos, vel, idx[bad])
        e_g64 = per_particle([a[bad] for a in gpu_pos], x64)
        e_o64 = per_particle([a[bad] for a in ref_pos], x64)
        worse = int(np.count_nonzero(e_g64 > e_o64 + 1e-7))
        msg += (f"; {bad.size} over 1e-6: vs fp64 gpu max {e_g64.max():.2e} "
                f"p99 {np.percentile(e_g64, 99):.2e}, oracle max {e_o64.max():.2e} "
                f"p99 {np.percentile(e_o64, 99):.2e}; gpu farther at {worse}")
        assert worse == 0, msg
        e_gpu[bad] = e_g64
    msg += f"; gpu p99 vs fp64 <= {np.percentile(e_gpu, 99):.2e}"
    print(msg)
    assert np.percentile(e_gpu, 99) <= TOL_STEP, msg


def gpu_steps(nb, params, pos, vel, steps):
    params.sim_iterations_per_frame = steps
    sim = nb.Simulator(params, pos, vel)
    sim.step_sim()
    out = sim.positions(), sim.velocities()
    sim.close()
    return out


def dev3(torch, arrs, gpu):
    return [torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(gpu) for a in arrs]


def test_one_step_config1(nb, orc, gpu):
    n = 25600  # 100 x 256, the CPU configuration
    pos, vel = nb.disk_ic(n)
    (gp, gv) = gpu_steps(nb, nb.default_params(num_particles=n), pos, vel, 1)
    p = orc.params(num_particles=n)
    rp, rv = orc.step_f32(p, pos, vel)
    assert_step_parity(orc, p, pos, vel, gp, rp, np.arange(n), f"N={n} step 1")


def assert_trajectory_parity(gpu_pos, ref_pos, pos64, label):
    """Positions after the same number of steps from the same initial state:
    the GPU's max and 99th-percentile distance from the fp64 trajectory are
    no larger than the fp32 oracle's."""
    e_g = per_particle(gpu_pos, pos64)
    e_o = per_particle(ref_pos, pos64)
    e_go = per_particle(gpu_pos, ref_pos)
    msg = (f"{label}: vs fp64 gpu max {e_g.max():.2e} p99 {np.percentile(e_g, 99):.2e}, "
           f"oracle max {e_o.max():.2e} p99 {np.percentile(e_o, 99):.2e}; gpu vs oracle max "
           f"{e_go.max():.2e} p99 {np.percentile(e_go, 99):.2e}")
    print(msg)
    assert e_g.max() <= e_o.max(), msg
    assert np.percentile(e_g, 99) <= np.percentile(e_o, 99), msg


def test_tenth_step_and_trajectory_config1(nb, orc, gpu):
    n = 25600
    pos, vel = nb.disk_ic(n)
    p = orc.params(num_particles=n)
    rp, rv = pos, vel
    for _ in range(9):
        rp, rv = orc.step_f32(p, rp, rv)
    # step 10 from the identical (oracle) state at step 9
    gp, _ = gpu_steps(nb, nb.default_params(num_particles=n), rp, rv, 1)
    rp10, _ = orc.step_f32(p, rp, rv)
    assert_step_parity(orc, p, rp, rv, gp, rp10, np.arange(n), f"N={n} step 10")
    # ten steps from the initial conditions, refereed by the fp64 trajectory
    gp10, _ = gpu_steps(nb, nb.default_params(num_particles=n), pos, vel, 10)
    x64, v64 = pos, vel
    for _ in range(10):
        x64, v64 = orc.step_f64(p, x64, v64)
    assert_trajectory_parity(gp10, rp10, x64, f"N={n} 10-step trajectory")


@pytest.mark.parametrize("n", [1, 2, 3, 5, 63, 64, 65, 255, 257, 1023, 1025, 4097])
def test_ragged_sizes(nb, orc, gpu, n):
    pos, vel = nb.plummer_ic(n, seed=n, g_mass=1.0)
    params = nb.default_params(num_particles=n, G=1.0 / max(n, 1))
    gp, gv = gpu_steps(nb, params, pos, vel, 1)
    p = orc.params(num_particles=n, G=1.0 / max(n, 1))
    rp, rv = orc.step_f32(p, pos, vel)
    assert_step_parity(orc, p, pos, vel, gp, rp, np.arange(n), f"N={n}")
    if n == 1:  # no force: the integrator's rounding is reproduced exactly
        for a, b in zip(gp + gv, rp + rv):
            assert np.array_equal(a, b)


def test_every_launch_shape(nb, orc, gpu):
    import torch

    n = 5000  # not a multiple of any block or tile
    pos, vel = nb.disk_ic(n)
    p = orc.params(num_particles=n)
    rp, rv = orc.step_f32(p, pos, vel)
    params = nb.default_params(num_particles=n)
    results = []
    for cfg in range(nb.forward_config_count()):
        dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
        dn = [torch.empty_like(t) for t in dp]
        nb.forward_step(params, dp, dv, dn, cfg=cfg)
        torch.cuda.synchronize()
        got = [t.cpu().numpy() for t in dn]
        assert_step_parity(orc, p, pos, vel, got, rp, np.arange(n),
                           f"cfg {cfg} {nb.forward_config_info(cfg)}")
        results.append(got)


def test_handle_and_device_paths_agree(nb, gpu):
    import torch

    n = 3000
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n)
    gp, gv = gpu_steps(nb, params, pos, vel, 1)
    dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
    dn = [torch.empty_like(t) for t in dp]
    nb.forward_step(params, dp, dv, dn)
    torch.cuda.synchronize()
    for a, b in zip(gp + gv, [t.cpu().numpy() for t in dn + dv]):
        assert np.array_equal(a, b)


def test_zero_eps_skips_self(nb, orc, gpu):
    n = 700
    pos, vel = nb.plummer_ic(n, seed=3, g_mass=1.0)
    params = nb.default_params(num_particles=n, dist_eps=0.0, G=1.0 / n)
    gp, gv = gpu_steps(nb, params, pos, vel, 1)
    p = orc.params(num_particles=n, dist_eps=0.0, G=1.0 / n)
    rp, rv = orc.step_f32(p, pos, vel)
    assert np.isfinite(np.stack(gp)).all()
    assert_step_parity(orc, p, pos, vel, gp, rp, np.arange(n), f"eps=0 N={n}")


def test_predicated_matches_reference_semantics(nb, orc, gpu):
    n = 1000
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n, calc_method=nb.NB_CALC_PREDICATED)
    gp, gv = gpu_steps(nb, params, pos, vel, 1)
    rp, rv = orc.step_f32(orc.params(num_particles=n, calc_method=1), pos, vel)
    for a, b in zip(gp + gv, rp + rv):  # force is +0: bit-exact damping + drift
        assert np.array_equal(a, b)
    params.dist_eps = 0.0  # 0 * rsqrt(0) in the self term: NaN
    gp, gv = gpu_steps(nb, params, pos, vel, 1)
    assert np.isnan(np.stack(gv)).all()


def test_kahan_accumulation_is_closer_to_fp64(nb, orc, gpu):
    import torch

    n = 65536
    pos, vel = nb.disk_ic(n)
    idx = np.arange(0, n, 97)
    f64 = np.stack(orc.forces_f64(pos, idx))
    params = nb.default_params(num_particles=n, G=1.0, dt=1.0, damping=0.0)
    errs = {}
    for acc in (nb.NB_ACC_PLAIN, nb.NB_ACC_KAHAN):
        dp = dev3(torch, pos, gpu)
        dv = [torch.zeros(n, device=gpu) for _ in range(3)]
        dn = [torch.empty_like(t) for t in dp]
        nb.forward_step(params, dp, dv, dn, acc=acc)  # v' = F with these params
        torch.cuda.synchronize()
        F = np.stack([t.cpu().numpy()[idx] for t in dv]).astype(np.float64)
        errs[acc] = np.abs(F - f64).max(0) / np.abs(f64).max(0)
    print(f"force rel err vs fp64: plain p50 {np.median(errs[0]):.2e} max {errs[0].max():.2e}; "
          f"kahan p50 {np.median(errs[1]):.2e} max {errs[1].max():.2e}")
    assert np.median(errs[1]) < np.median(errs[0])
    assert errs[1].max() < 1e-5


def test_empty_system(nb, gpu):
    params = nb.default_params(num_particles=0)
    e = np.zeros(0, np.float32)
    sim = nb.Simulator(params, (e, e, e), (e, e, e))
    sim.step_sim()
    assert sim.num_particles == 0
    sim.close()


def test_frame_timing(nb, gpu):
    n = 8192
    pos, vel = nb.disk_ic(n)
    sim = nb.Simulator(nb.default_params(num_particles=n), pos, vel)
    sim.step_sim()
    assert 0 < sim.last_kernel_ms <= sim.last_step_ms * 1.05 + 0.05
    sim.close()
    # the library is built for gfx950 only: the name carries the part's ISA
    assert "gfx950" in nb.device_name(), nb.device_name()


def has_kahan(info):
    """The shape has a Kahan variant: every target-sharded shape, the
    symmetric ones with 2 or 4 particles per lane and one item per tile pair
    (the two-item tile-pair shape is plain sums only)."""
    if info["scheme"] == "targets":
        return True
    return info["unroll"] <= 4 and not (info["scheme"] == "symmetric_tiles"
                                        and info["split"] > 1)


def sym_configs(nb):
    """Symmetric shapes: super-tile items first, then the tile-pair form."""
    info = [nb.forward_config_info(c) for c in range(nb.forward_config_count())]
    return ([c for c, i in enumerate(info) if i["scheme"] == "symmetric"] +
            [c for c, i in enumerate(info) if i["scheme"] == "symmetric_tiles"])


def target_config(nb, n):
    """The library's target-sharded choice for large n (4 x 256, split 8)."""
    for c in range(nb.forward_config_count()):
        i = nb.forward_config_info(c)
        if i["scheme"] == "targets" and (i["block"], i["unroll"], i["split"]) == (256, 4, 8):
            return c
    raise AssertionError("no 256 x 4 split-8 shape")


@pytest.mark.parametrize("cfg", [-1, "targets"])
def test_full_size_sampled_steps(nb, orc, gpu, cfg):
    """N = 262144 (1024 x 256): step 1 and step 10 (from the GPU's own state
    at step 9) checked on two 2048-particle target ranges, for the library's
    default shape (the symmetric scheme) and the target-sharded one."""
    n = 262144
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n, sim_iterations_per_frame=1)
    sim = nb.Simulator(params, pos, vel)
    if cfg == "targets":
        sim.set_config(target_config(nb, n))
    else:
        assert nb.forward_config_info(sim.config)["scheme"] == "symmetric"
    p = orc.params(num_particles=n)
    ranges = [(0, 2048), (200000, 202048)]
    prev = (pos, vel)
    for step in range(1, 11):
        sim.step_sim()
        if step in (1, 10):
            gp = sim.positions()
            for (i0, i1) in ranges:
                rp, _ = orc.step_f32(p, prev[0], prev[1], i0, i1)
                sl = slice(i0, i1)
                assert_step_parity(orc, p, prev[0], prev[1], [a[sl] for a in gp],
                                   [a[sl] for a in rp], np.arange(i0, i1),
                                   f"N={n} step {step} targets [{i0},{i1})")
        if step == 9:
            prev = (sim.positions(), sim.velocities())
    sim.close()


def forces64_torch(torch, X, eps, chunk=512):
    """fp64 all-pairs force sum_j (x_j - x_i) (|r|^2 + eps)^-3/2 in plain
    PyTorch (float64 throughout, X: (3, n) float64 CUDA tensor), chunks of
    `chunk` targets.  The self term is 0 for eps > 0."""
    n = X.shape[1]
    F = torch.empty_like(X)
    for i0 in range(0, n, chunk):
        xi = X[:, i0:i0 + chunk]
        r = [X[c][None, :] - xi[c][:, None] for c in range(3)]
        d = r[0] * r[0]
        d.addcmul_(r[1], r[1]).addcmul_(r[2], r[2]).add_(eps)
        K = torch.rsqrt(d * d * d)
        for c in range(3):
            F[c, i0:i0 + chunk] = (r[c] * K).sum(1)
    return F


def fp64_trajectory_torch(torch, gpu, p, pos, vel, steps):
    """The integrator of src/simulator.cu:214-228 in fp64 (the parameters'
    float32 values) over `steps` steps of forces64_torch."""
    X = torch.from_numpy(np.stack(pos).astype(np.float64)).to(gpu)
    V = torch.from_numpy(np.stack(vel).astype(np.float64)).to(gpu)
    f = lambda v: float(np.float32(v))
    eps, damping, dt, G = f(p.dist_eps), f(p.damping), f(p.dt), f(p.G)
    for _ in range(steps):
        F = forces64_torch(torch, X, eps)
        V = damping * V + (dt * G) * F
        X = X + dt * V
    return [a for a in X.cpu().numpy()]


def test_full_size_trajectory(nb, orc, gpu):
    """N = 262144: ten steps from the initial conditions with the library's
    default shape (symmetric) and the target-sharded shape, against the fp32
    oracle's ten steps, both refereed by an fp64 trajectory of all particles
    (plain PyTorch float64 on the GPU, its force checked here against the
    oracle's fp64 force)."""
    import torch

    n = 262144
    pos, vel = nb.disk_ic(n)
    p = orc.params(num_particles=n)
    # the fp64 referee's force against the oracle's fp64 force on a sample
    idx = np.arange(0, n, 509)
    F = forces64_torch(torch, torch.from_numpy(np.stack(pos).astype(np.float64)).to(gpu),
                       float(np.float32(p.dist_eps)))
    Fs = F[:, torch.from_numpy(idx).to(gpu)].cpu().numpy()
    Fo = np.stack(orc.forces_f64(pos, idx))
    ef = np.abs(Fs - Fo).max(0) / np.abs(Fo).max(0)
    print(f"fp64 referee force vs oracle fp64 force on {idx.size} targets: max {ef.max():.2e}")
    assert ef.max() <= 1e-10
    del F
    x64 = fp64_trajectory_torch(torch, gpu, p, pos, vel, 10)
    rp, rv = pos, vel
    for _ in range(10):
        rp, rv = orc.step_f32(p, rp, rv)
    for cfg in (-1, target_config(nb, n)):
        params = nb.default_params(num_particles=n, sim_iterations_per_frame=10)
        sim = nb.Simulator(params, pos, vel)
        if cfg >= 0:
            sim.set_config(cfg)
        label = nb.forward_config_info(sim.config)["scheme"]
        sim.step_sim()
        gp = sim.positions()
        sim.close()
        assert_trajectory_parity(gp, rp, x64, f"N={n} 10-step trajectory ({label})")


@pytest.mark.parametrize("acc", ["plain", "kahan"])
def test_precision_size_sampled_step(nb, orc, gpu, acc):
    """N = 2^20, the precision study's mixture (half Plummer, half disk,
    seed 7): one step of the library's shape (plain: symmetric, 8 per lane,
    two super-tiles per item; Kahan: symmetric, 2 per lane, four super-tiles
    per item, in two bands of source super-tiles) on three 2048-target
    ranges against the oracle, and the momentum change of the step against
    the damping-only prediction (fp64 GPU reduction): the symmetric sums
    cancel to rounding."""
    import torch

    from nbody_amd.studies import mixed_ics

    n = 1 << 20
    pos, vel = mixed_ics(n)
    params = nb.default_params(num_particles=n, sim_iterations_per_frame=1)
    sim = nb.Simulator(params, pos, vel)
    if acc == "kahan":
        sim.set_accumulation(nb.NB_ACC_KAHAN)
    info = nb.forward_config_info(sim.config)
    # Kahan: 2 per lane, four super-tiles per item, in two bands
    assert info["scheme"] == "symmetric", info
    sim.step_sim()
    gp, gv = sim.positions(), sim.velocities()
    sim.close()
    p = orc.params(num_particles=n)
    for (i0, i1) in [(0, 2048), (n // 2 - 1024, n // 2 + 1024), (n - 2048, n)]:
        rp, _ = orc.step_f32(p, pos, vel, i0, i1)
        sl = slice(i0, i1)
        assert_step_parity(orc, p, pos, vel, [a[sl] for a in gp], [a[sl] for a in rp],
                           np.arange(i0, i1), f"N=2^20 {acc} targets [{i0},{i1})")
    # momentum: P' = damping P + dt G sum F, with sum F ~ 0 (antisymmetric)
    em0 = nb.energy_momentum(params, torch.from_numpy(np.stack(pos)).to(gpu),
                             torch.from_numpy(np.stack(vel)).to(gpu))
    em1 = nb.energy_momentum(params, torch.from_numpy(np.stack(gp)).to(gpu),
                             torch.from_numpy(np.stack(gv)).to(gpu))
    dv = np.stack(gv).astype(np.float64) - np.float64(params.damping) * np.stack(vel)
    scale = np.abs(dv).sum(1)  # ~ dt G sum |F|
    drift = np.abs(np.array(em1["momentum"]) - params.damping * np.array(em0["momentum"]))
    print(f"N=2^20 {acc}: momentum drift / sum|dv| {drift / scale}")
    # measured 1e-9 (plain) and 1e-11 (Kahan); a lost or mis-signed reaction
    # gives order 1
    assert np.all(drift <= (1e-7 if acc == "plain" else 1e-8) * scale), (drift, scale)


def test_two_million_sampled_step(nb, orc, gpu):
    """N = 2^21 disk, the library's shape (symmetric, 8 per lane, four
    super-tiles per item, 16 GB of partials): one step on two 2048-target
    ranges against the oracle."""
    n = 1 << 21
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n, sim_iterations_per_frame=1)
    sim = nb.Simulator(params, pos, vel)
    info = nb.forward_config_info(sim.config)
    assert (info["scheme"], info["unroll"], info["split"]) == ("symmetric", 8, 4), info
    sim.step_sim()
    gp = sim.positions()
    sim.close()
    p = orc.params(num_particles=n)
    for (i0, i1) in [(0, 2048), (n - 2048, n)]:
        rp, _ = orc.step_f32(p, pos, vel, i0, i1)
        sl = slice(i0, i1)
        assert_step_parity(orc, p, pos, vel, [a[sl] for a in gp], [a[sl] for a in rp],
                           np.arange(i0, i1), f"N=2^21 targets [{i0},{i1})")


def test_banded_step_past_two_million(nb, orc, gpu):
    """N = 2^21 + 4097: the plain step runs in bands of source super-tiles
    (two here, each with at most 8 GB of partials), summed band by band:
    one step against the oracle on the first targets and the last, partial
    super-tile."""
    n = (1 << 21) + 4097
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n, sim_iterations_per_frame=1)
    sim = nb.Simulator(params, pos, vel)
    info = nb.forward_config_info(sim.config)
    assert (info["scheme"], info["unroll"], info["split"]) == ("symmetric", 8, 4), info
    sim.step_sim()
    gp = sim.positions()
    sim.close()
    p = orc.params(num_particles=n)
    for (i0, i1) in [(0, 2048), (n - 2048, n)]:
        rp, _ = orc.step_f32(p, pos, vel, i0, i1)
        sl = slice(i0, i1)
        assert_step_parity(orc, p, pos, vel, [a[sl] for a in gp], [a[sl] for a in rp],
                           np.arange(i0, i1), f"N=2^21+4097 targets [{i0},{i1})")


def test_reference_driver_runs(nb, gpu):
    """The reference's own src/nbody.cpp, compiled unchanged against this
    project's simulator.hpp (oracle/_ref/nbody_dropin), drives the GPU."""
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "nbody_dropin"
    runs = [[str(nb.DRIVER_PATH)]]
    if exe.exists():
        runs.append([str(exe)])
    for cmd in runs:
        # 2048 particles (target-sharded) and 51200 (symmetric tile pairs),
        # frames of 2 steps (replayed hipGraphs)
        for batches in ("8", "200"):
            res = subprocess.run(cmd + [batches, "2", "0.999998", "0.005", "1e-7", "2", "5",
                                        "64", "BRANCH"],
                                 capture_output=True, text=True, timeout=120)
            assert res.returncode == 0, res.stderr
            lines = [l for l in res.stdout.splitlines() if l.startswith("At step")]
            assert len(lines) == 3 and lines[-1].startswith("At step 5 kernel time is ")


@pytest.mark.parametrize("n", [1, 3001, 300001])
def test_staged_readback_matches_plain_reads(nb, gpu, n):
    """The staging kernel (grid-stride loop over six rows into device-mapped
    pinned memory) returns what the plain copies return, bit for bit."""
    pos, vel = nb.disk_ic(n)
    sim = nb.Simulator(nb.default_params(num_particles=n, sim_iterations_per_frame=3), pos, vel)
    sim.step_sim()
    (sp, sv) = sim.staged()
    for a, b in zip(sp + sv, sim.positions() + sim.velocities()):
        assert np.array_equal(a, b)
    sim.close()


def test_vec4_pack(nb, gpu):
    import torch

    n = 4099
    pos, vel = nb.disk_ic(n)
    sim = nb.Simulator(nb.default_params(num_particles=n), pos, vel)
    p4 = sim.read_vec4(0, 1.0)
    v4 = sim.read_vec4(1, 0.0)
    assert np.array_equal(p4[:, :3], np.stack(pos, axis=1)) and np.all(p4[:, 3] == 1.0)
    assert np.array_equal(v4[:, :3], np.stack(vel, axis=1)) and np.all(v4[:, 3] == 0.0)
    sim.close()
    soa = torch.from_numpy(np.stack(pos)).to(gpu)
    out = nb.pack_vec4(soa, 2.5).cpu().numpy()
    assert np.array_equal(out[:, :3], np.stack(pos, axis=1)) and np.all(out[:, 3] == 2.5)
    with pytest.raises(nb.NbError):  # output overlapping the input
        import ctypes
        nb.check(nb.load().nb_pack_vec4_dev(ctypes.byref(nb.NbVec3(*[ctypes.c_void_p(
            soa[c].data_ptr()) for c in range(3)])), 1.0, ctypes.c_void_p(soa.data_ptr()), n, None))


@pytest.mark.parametrize("n", [1, 2, 255, 1024, 2047, 2048, 2049, 6145, 70001])
def test_symmetric_scheme_sizes(nb, orc, gpu, n):
    """Every symmetric shape at sizes around its tile and super-tile lengths
    (padding particles must add exact zeros), checked against the oracle on
    the first, a middle and the last 1500 targets."""
    import torch

    pos, vel = nb.plummer_ic(n, seed=n, g_mass=1.0)
    params = nb.default_params(num_particles=n, G=1.0 / n)
    p = orc.params(num_particles=n, G=1.0 / n)
    ranges = sorted({(0, min(n, 1500)), (max(0, n // 2 - 750), min(n, n // 2 + 750)),
                     (max(0, n - 1500), n)})
    refs = {r: orc.step_f32(p, pos, vel, *r)[0] for r in ranges}
    for cfg in sym_configs(nb):
        dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
        dn = [torch.empty_like(t) for t in dp]
        nb.forward_step(params, dp, dv, dn, cfg=cfg)
        torch.cuda.synchronize()
        got = [t.cpu().numpy() for t in dn]
        assert np.isfinite(np.stack(got)).all()
        for (i0, i1), rp in refs.items():
            sl = slice(i0, i1)
            assert_step_parity(orc, p, pos, vel, [a[sl] for a in got], [a[sl] for a in rp],
                               np.arange(i0, i1),
                               f"N={n} cfg {cfg} {nb.forward_config_info(cfg)} [{i0},{i1})")


@pytest.mark.parametrize("eps", [0.0, 1e-20])
@pytest.mark.parametrize("n", [255, 2049, 6145])
def test_symmetric_scheme_vanishing_eps(nb, orc, gpu, eps, n):
    """dist_eps = 0 (and eps^3 below the normal range): every symmetric
    shape, plain and Kahan sums, masks the self pairs as the reference's
    `continue` skips them; finite results in parity with the oracle step."""
    import torch

    pos, vel = nb.plummer_ic(n, seed=n + 7, g_mass=1.0)
    params = nb.default_params(num_particles=n, dist_eps=eps, G=1.0 / n)
    p = orc.params(num_particles=n, dist_eps=eps, G=1.0 / n)
    rp, _ = orc.step_f32(p, pos, vel)
    for cfg in sym_configs(nb):
        accs = [nb.NB_ACC_PLAIN]
        if has_kahan(nb.forward_config_info(cfg)):
            accs.append(nb.NB_ACC_KAHAN)
        for acc in accs:
            dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
            dn = [torch.empty_like(t) for t in dp]
            nb.forward_step(params, dp, dv, dn, cfg=cfg, acc=acc)
            torch.cuda.synchronize()
            got = [t.cpu().numpy() for t in dn]
            what = f"eps={eps} N={n} cfg {cfg} {nb.forward_config_info(cfg)} acc {acc}"
            assert np.isfinite(np.stack(got)).all(), what
            assert_step_parity(orc, p, pos, vel, got, rp, np.arange(n), what)


COINCIDENT = [(1, 0), (1500, 5), (3000, 2999), (2990, 100), (701, 700), (2000, 700)]


@pytest.mark.parametrize("eps", ["default", 0.0])
def test_coincident_particles(nb, orc, gpu, eps):
    """Distinct particles at one position (3001 particles; pairs within a
    tile, across tiles, in the last partial tile, and a triple).  With
    SimParam's eps the pair term r K is an exact zero: every launch shape,
    plain and Kahan sums, in parity with the oracle step.  With eps = 0 the
    reference's arithmetic gives 0 * rsqrt(0) = NaN for exactly those
    particles (its `continue` skips only j == i): every shape gives NaN for
    the same particles as the oracle and parity for all others."""
    import torch

    n = 3001
    pos, vel = nb.disk_ic(n)
    pos = [a.copy() for a in pos]
    for dst, src in COINCIDENT:
        for a in pos:
            a[dst] = a[src]
    kw = {} if eps == "default" else {"dist_eps": eps}
    params = nb.default_params(num_particles=n, **kw)
    p = orc.params(num_particles=n, **kw)
    rp, _ = orc.step_f32(p, pos, vel)
    dup = sorted({i for pr in COINCIDENT for i in pr})
    ref_nan = ~np.isfinite(np.stack(rp)).all(0)
    if eps == "default":
        assert not ref_nan.any()
    else:
        assert np.array_equal(np.nonzero(ref_nan)[0], dup)
    keep = np.nonzero(~ref_nan)[0]
    for cfg in range(nb.forward_config_count()):
        info = nb.forward_config_info(cfg)
        accs = [nb.NB_ACC_PLAIN]
        if has_kahan(info):
            accs.append(nb.NB_ACC_KAHAN)
        for acc in accs:
            dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
            dn = [torch.empty_like(t) for t in dp]
            nb.forward_step(params, dp, dv, dn, cfg=cfg, acc=acc)
            torch.cuda.synchronize()
            got = [t.cpu().numpy() for t in dn]
            what = f"eps={eps} cfg {cfg} {info} acc {acc}"
            assert np.array_equal(~np.isfinite(np.stack(got)).all(0), ref_nan), what
            assert_step_parity(orc, p, pos, vel, [a[keep] for a in got],
                               [a[keep] for a in rp], keep, what)


def test_symmetric_scheme_is_deterministic(nb, gpu):
    """Partials are summed in a fixed order: repeated steps are bit-identical."""
    import torch

    n = 100000
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n)
    outs = []
    for cfg in sym_configs(nb)[:2] * 2:
        dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
        dn = [torch.empty_like(t) for t in dp]
        nb.forward_step(params, dp, dv, dn, cfg=cfg)
        torch.cuda.synchronize()
        outs.append(np.stack([t.cpu().numpy() for t in dn + dv]))
    h = len(outs) // 2
    for a, b in zip(outs[:h], outs[h:]):
        assert np.array_equal(a, b)


def test_symmetric_scheme_momentum(nb, orc, gpu):
    """Each pair term enters both sums with opposite sign, so the total force
    of the symmetric scheme is a sum of rounding residues: it must be far
    below the total force magnitude and no larger than the fp32 oracle's."""
    import torch

    n = 20000
    pos, vel = nb.plummer_ic(n, seed=11, g_mass=1.0)
    params = nb.default_params(num_particles=n, G=1.0, dt=1.0, damping=0.0)
    cfg = sym_configs(nb)[0]
    dp = dev3(torch, pos, gpu)
    dv = [torch.zeros(n, device=gpu) for _ in range(3)]
    dn = [torch.empty_like(t) for t in dp]
    nb.forward_step(params, dp, dv, dn, cfg=cfg)  # v' = F with these params
    torch.cuda.synchronize()
    F = np.stack([t.cpu().numpy() for t in dv]).astype(np.float64)
    p = orc.params(num_particles=n, G=1.0, dt=1.0, damping=0.0)
    _, rv = orc.step_f32(p, pos, [np.zeros(n, np.float32)] * 3)
    Fo = np.stack(rv).astype(np.float64)
    scale = np.abs(F).sum(1)
    net, net_o = np.abs(F.sum(1)), np.abs(Fo.sum(1))
    print(f"net force / sum |F|: symmetric {net / scale}, oracle {net_o / scale}")
    assert np.all(net <= 1e-5 * scale)
    assert np.all(net <= 2 * net_o + 1e-7 * scale)


def test_symmetric_scheme_rejects_other_modes(nb, gpu):
    import torch

    cfg = sym_configs(nb)[0]
    n = 3000
    pos, vel = nb.disk_ic(n)
    for kw, acc in (({"calc_method": nb.NB_CALC_PREDICATED}, nb.NB_ACC_PLAIN),
                    ({"calc_method": nb.NB_CALC_PREDICATED}, nb.NB_ACC_KAHAN)):
        params = nb.default_params(num_particles=n, **kw)
        dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
        dn = [torch.empty_like(t) for t in dp]
        with pytest.raises(nb.NbError, match="symmetric"):
            nb.forward_step(params, dp, dv, dn, cfg=cfg, acc=acc)
    # the two-item tile-pair shape takes plain sums only
    two = [c for c in sym_configs(nb) if not has_kahan(nb.forward_config_info(c))
           and nb.forward_config_info(c)["scheme"] == "symmetric_tiles"]
    assert two
    params = nb.default_params(num_particles=n)
    dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
    dn = [torch.empty_like(t) for t in dp]
    with pytest.raises(nb.NbError, match="two items per tile pair"):
        nb.forward_step(params, dp, dv, dn, cfg=two[0], acc=nb.NB_ACC_KAHAN)
    # the simulator's default shape follows the accumulation mode
    sim = nb.Simulator(nb.default_params(num_particles=70000), *nb.disk_ic(70000))
    assert nb.forward_config_info(sim.config)["scheme"] != "targets"
    sim.set_accumulation(nb.NB_ACC_KAHAN)  # the symmetric Kahan shapes: 2 or 4 per lane
    assert nb.forward_config_info(sim.config)["unroll"] in (2, 4)
    sim.step_sim()
    sim.set_config(sym_configs(nb)[0])  # 8 per lane: no Kahan variant
    with pytest.raises(nb.NbError, match="4 particles per lane"):
        sim.step_sim()
    sim.set_config(-1)
    sim.set_accumulation(nb.NB_ACC_PLAIN)
    assert nb.forward_config_info(sim.config)["scheme"] != "targets"
    sim.close()
    # beyond 2^20 particles (2^17 for tile pairs, 2^18 with 2 per lane, 2^24
    # with plain sums, 4 or 8 per lane and 4 per item): refused before any
    # launch
    m = (1 << 20) + 1
    big = [torch.zeros(m, device=gpu) for _ in range(9)]
    with pytest.raises(nb.NbError, match="2\\^20"):
        nb.forward_step(nb.default_params(num_particles=m), big[:3], big[3:6], big[6:],
                        cfg=cfg)
    tiles = next(c for c in sym_configs(nb)
                 if nb.forward_config_info(c)["scheme"] == "symmetric_tiles")
    m = (1 << 17) + 1
    with pytest.raises(nb.NbError, match="2\\^17"):
        nb.forward_step(nb.default_params(num_particles=m), [t[:m] for t in big[:3]],
                        [t[:m] for t in big[3:6]], [t[:m] for t in big[6:]], cfg=tiles)
    large = next(c for c in sym_configs(nb)  # 8 per lane, 4 per item: banded to 2^24
                 if nb.forward_config_info(c)["unroll"] == 8
                 and nb.forward_config_info(c)["split"] >= 4)
    m2 = (1 << 24) + 1
    big2 = [torch.zeros(m2, device=gpu) for _ in range(9)]
    with pytest.raises(nb.NbError, match="2\\^24"):
        nb.forward_step(nb.default_params(num_particles=m2), big2[:3], big2[3:6], big2[6:],
                        cfg=large)
    m20 = (1 << 20) + 1
    kahan4 = next(c for c in sym_configs(nb)  # Kahan sums: 2^20
                  if nb.forward_config_info(c)["scheme"] == "symmetric"
                  and nb.forward_config_info(c)["unroll"] == 4)
    with pytest.raises(nb.NbError, match="2\\^20"):
        nb.forward_step(nb.default_params(num_particles=m20), big[:3], big[3:6], big[6:],
                        cfg=kahan4, acc=nb.NB_ACC_KAHAN)
    del big2
    small = next(c for c in sym_configs(nb)  # super-tiles of 2 per lane: 2^18
                 if nb.forward_config_info(c)["scheme"] == "symmetric"
                 and nb.forward_config_info(c)["unroll"] == 2)
    m = (1 << 18) + 1
    with pytest.raises(nb.NbError, match="2\\^18"):
        nb.forward_step(nb.default_params(num_particles=m), [t[:m] for t in big[:3]],
                        [t[:m] for t in big[3:6]], [t[:m] for t in big[6:]], cfg=small)


def test_frame_graph_replays_banded_steps(nb, gpu):
    """N = 800001 with Kahan sums: the library's shape (2 per lane, four
    super-tiles per item) runs each step as two bands of source super-tiles
    (band kernel and fold per band, then the finish kernel).  Frames of 2
    and 3 steps replayed from captured hipGraphs equal step-by-step launches
    bit for bit."""
    n = 800001
    pos, vel = nb.disk_ic(n)
    sims = []
    for graph in (True, False):
        s = nb.Simulator(nb.default_params(num_particles=n, sim_iterations_per_frame=2),
                         pos, vel)
        s.set_accumulation(nb.NB_ACC_KAHAN)
        s.set_graph(graph)
        sims.append(s)
    info = nb.forward_config_info(sims[0].config)
    assert (info["scheme"], info["unroll"], info["split"]) == ("symmetric", 2, 4), info
    for k, iters in enumerate((2, 3)):
        for s in sims:
            s.set_frame_iterations(iters)
            s.step_sim()
        g, d = sims
        for a, b in zip(g.positions() + g.velocities(), d.positions() + d.velocities()):
            assert np.array_equal(a, b), (k, iters)
    for s in sims:
        s.close()


@pytest.mark.parametrize("n", [3000, 25600])
def test_frame_graph_replays_match_direct_launches(nb, gpu, n):
    """Frames replayed from captured hipGraphs equal step-by-step launches bit
    for bit, over several frames with reads between them, an odd frame
    length (buffer parity) and a change of launch shape."""
    pos, vel = nb.disk_ic(n)
    g = nb.Simulator(nb.default_params(num_particles=n, sim_iterations_per_frame=4), pos, vel)
    d = nb.Simulator(nb.default_params(num_particles=n, sim_iterations_per_frame=4), pos, vel)
    d.set_graph(False)
    plan = [4, 4, 3, 3, 2, 5]
    for k, iters in enumerate(plan):
        for s in (g, d):
            s.set_frame_iterations(iters)
            if k == 4:
                s.set_config(0)  # 256 x 1, no split
            s.step_sim()
        for a, b in zip(g.positions() + g.velocities(), d.positions() + d.velocities()):
            assert np.array_equal(a, b), (k, iters)
    g.close()
    d.close()


@pytest.mark.parametrize("n", [3000, 70001])
def test_symmetric_kahan(nb, orc, gpu, n):
    """Kahan-compensated accumulation in the symmetric shapes (2 or 4 per lane):
    parity with the oracle step, and forces closer to fp64 than the plain
    symmetric sums (both sampled against the fp64 oracle forces)."""
    import torch

    pos, vel = nb.disk_ic(n)
    idx = np.arange(0, n, 7)
    f64 = np.stack(orc.forces_f64(pos, idx))
    unit = nb.default_params(num_particles=n, G=1.0, dt=1.0, damping=0.0)
    kahan_cfgs = [c for c in sym_configs(nb) if has_kahan(nb.forward_config_info(c))]
    assert kahan_cfgs
    for cfg in kahan_cfgs:
        errs = {}
        for acc in (nb.NB_ACC_PLAIN, nb.NB_ACC_KAHAN):
            dp = dev3(torch, pos, gpu)
            dv = [torch.zeros(n, device=gpu) for _ in range(3)]
            dn = [torch.empty_like(t) for t in dp]
            nb.forward_step(unit, dp, dv, dn, cfg=cfg, acc=acc)  # v' = F
            torch.cuda.synchronize()
            F = np.stack([t.cpu().numpy()[idx] for t in dv]).astype(np.float64)
            errs[acc] = np.abs(F - f64).max(0) / np.abs(f64).max(0)
        info = nb.forward_config_info(cfg)
        print(f"N={n} cfg {cfg} {info}: plain p50 {np.median(errs[0]):.2e} "
              f"max {errs[0].max():.2e}; kahan p50 {np.median(errs[1]):.2e} "
              f"max {errs[1].max():.2e}")
        assert np.median(errs[1]) < np.median(errs[0]) and errs[1].max() < 1e-5
        # one SimParam-default step against the oracle
        params = nb.default_params(num_particles=n)
        dp, dv = dev3(torch, pos, gpu), dev3(torch, vel, gpu)
        dn = [torch.empty_like(t) for t in dp]
        nb.forward_step(params, dp, dv, dn, cfg=cfg, acc=nb.NB_ACC_KAHAN)
        torch.cuda.synchronize()
        p = orc.params(num_particles=n)
        rp, _ = orc.step_f32(p, pos, vel, 0, min(n, 3000))
        sl = slice(0, min(n, 3000))
        assert_step_parity(orc, p, pos, vel, [t.cpu().numpy()[sl] for t in dn],
                           [a[sl] for a in rp], np.arange(sl.stop), f"kahan N={n} cfg {cfg}")
