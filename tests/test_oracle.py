"""Pins the CPU oracle (oracle/nbody_oracle.c) before it is trusted as the
checker of the HIP kernels.

The reference ships no tests, golden vectors or fixtures for this path, so
each oracle function is pinned here by something independent of it:
  * std::mt19937: the C++ standard's known answer and numpy's MT19937;
  * generate_canonical / disk ICs: exact rational recomputation from raw
    engine outputs, bit-equality with the libstdc++-based product generator,
    closed-form invariants of the disk;
  * the fp32 step: two-body and regular-polygon closed forms, the N=1
    damping/drift recurrence replayed in numpy float32, fp64 brute force,
    self-skip with eps = 0, the PREDICATED (i == id) semantics;
  * fp64 energy: two-body closed form;  fp64 adjoint: central differences.
"""
from fractions import Fraction

import numpy as np
import pytest

P_DEFAULT = dict(G=2.0, dt=0.005, damping=0.999998, dist_eps=1.0e-7)


def raw_mt(seed, n):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)  # init_genrand: the std::mt19937(seed) seeding
    return bg.random_raw(n).astype(np.uint64)


# ------------------------------------------------------------------ engine
def test_mt19937_known_answer(orc):
    # [rand.predef]: the 10000th invocation of a default-constructed
    # std::mt19937 produces 4123659995.
    assert orc.mt_nth(5489, 10000) == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 5489, 2**31 + 7])
def test_mt19937_matches_numpy(orc, seed):
    raw = raw_mt(seed, 1300)  # crosses two twists
    for k in (1, 2, 624, 625, 1248, 1300):
        assert orc.mt_nth(seed, k) == int(raw[k - 1])


def canonical_from_raw(u0, u1):
    """generate_canonical<double,53> from two 32-bit outputs, exactly:
    (u0 + u1 * 2^32) rounded to double, divided by 2^64."""
    s = float(Fraction(int(u0)) + Fraction(int(u1)) * 2**32)  # one rounding
    return s / 2.0**64


def test_disk_first_particles_from_raw_engine(orc):
    n = 8
    (px, py, pz), (vx, vy, vz) = orc.disk(n)
    raw = raw_mt(5489, 6 * n)
    u = [canonical_from_raw(raw[2 * k], raw[2 * k + 1]) for k in range(3 * n)]
    pi_f = float(np.float32(3.14159265358979323846))
    for i in range(n):
        # per particle: t then s; all z after the whole x-y loop
        t = np.float32(u[2 * i] * 2 * pi_f)
        s = np.float32(u[2 * i + 1] * 100)
        assert abs(float(px[i]) - np.cos(float(t)) * float(s)) <= 1e-6 * float(s) + 1e-30
        assert abs(float(py[i]) - np.sin(float(t)) * float(s)) <= 1e-6 * float(s) + 1e-30
        assert pz[i] == np.float32(4.0 * u[2 * n + i])


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_disk_matches_libstdcxx_generator(orc, nb, flags):
    """The oracle's C restatement equals the product generator, which draws
    from libstdc++'s std::mt19937 + uniform_real_distribution<double>."""
    a = orc.disk(4096, 5489, flags)
    b = nb.disk_ic(4096, 5489, flags)
    for u, v in zip(a[0] + a[1], b[0] + b[1]):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))


def test_disk_gen_order_matches_libstdcxx_generator(orc, nb):
    a = orc.disk_gen(2048)
    b = nb.disk_gen_ic(2048)
    for u, v in zip(a[0] + a[1], b[0] + b[1]):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    # gen.cpp draws z per particle, the simulator after all particles
    c = orc.disk(2048)
    assert not np.array_equal(a[0][2], c[0][2])


def test_disk_invariants(orc):
    (x, y, z), (vx, vy, vz) = orc.disk(20000)
    r = np.hypot(x.astype(np.float64), y.astype(np.float64))
    assert r.max() < 100.0 + 1e-4 and r.min() >= 0
    assert z.min() >= 0 and z.max() < 4.0
    assert np.all(vz == 0)
    # tangential: v . (x, y) = 0, |v| = sqrt(2 r)
    dot = vx.astype(np.float64) * x + vy.astype(np.float64) * y
    speed = np.sqrt(vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2)
    assert np.all(np.abs(dot) <= 1e-5 * (r * speed + 1))
    assert np.allclose(speed, np.sqrt(2 * r), rtol=2e-6, atol=1e-6)
    # radius is uniform on [0, 100): mean 50, not area-uniform (mean 66.7)
    assert abs(r.mean() - 50) < 1.0


def test_golden_disk_fixture(orc):
    """Committed hex-float fixture of the first 256 reference-disk particles."""
    from golden import load_disk_fixture

    want = load_disk_fixture()
    (px, py, pz), (vx, vy, vz) = orc.disk(256)
    got = np.stack([px, py, pz, vx, vy, vz])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# --------------------------------------------------------------- fp32 step
def f32(x):
    return np.float32(x)


def test_two_body_closed_form(orc):
    d, eps = 1.5, 1.0e-7
    pos = (np.array([0.0, d], np.float32), np.zeros(2, np.float32), np.zeros(2, np.float32))
    vel = (np.zeros(2, np.float32), np.array([0.3, -0.3], np.float32), np.zeros(2, np.float32))
    p = orc.params(**P_DEFAULT)
    (nx, ny, nz), (vx, vy, vz) = orc.step_f32(p, pos, vel)
    F = d / (d * d + eps) ** 1.5  # attraction toward the other body
    v0x = 0.0 * 0.999998 + F * 0.005 * 2.0
    assert np.isclose(vx[0], v0x, rtol=1e-6)
    assert np.isclose(vx[1], -v0x, rtol=1e-6)
    assert np.isclose(vy[0], 0.3 * 0.999998, rtol=1e-7)
    # position update uses the new velocity: x' = x + v' dt
    assert np.isclose(nx[0], 0.0 + v0x * 0.005, rtol=1e-6)
    assert np.isclose(ny[1], -0.3 * 0.999998 * 0.005, rtol=1e-6)
    assert np.all(nz == 0)


def test_regular_polygon_closed_form(orc):
    k, R, eps = 7, 3.0, 1.0e-7
    th = 2 * np.pi * np.arange(k) / k
    pos = (np.float32(R) * np.cos(th).astype(np.float32),
           np.float32(R) * np.sin(th).astype(np.float32), np.zeros(k, np.float32))
    zeros = tuple(np.zeros(k, np.float32) for _ in range(3))
    p = orc.params(G=1.0, dt=1.0, damping=1.0, dist_eps=eps)
    _, (vx, vy, _) = orc.step_f32(p, pos, zeros)
    m = np.arange(1, k)
    Fx = R * np.sum((np.cos(th[m]) - 1) / (4 * R * R * np.sin(th[m] / 2) ** 2 + eps) ** 1.5)
    # particle 0 sits on +x: net force points to the centre
    assert np.isclose(vx[0], Fx, rtol=2e-6)
    assert abs(vy[0]) < 1e-6 * abs(Fx)
    radial = vx * np.cos(th) + vy * np.sin(th)
    assert np.allclose(radial, Fx, rtol=5e-6)
    f64 = orc.forces_f64(pos, eps=eps)
    assert np.isclose(f64[0][0], Fx, rtol=1e-6)


def test_single_particle_recurrence(orc):
    """N = 1: no force, so v_n = v * damping each step and x += v_n * dt,
    replayed in numpy float32 with the same rounding: bit-equal."""
    p = orc.params(**P_DEFAULT)
    pos = (np.array([1.25], np.float32), np.array([-3.5], np.float32), np.array([0.5], np.float32))
    vel = (np.array([0.7], np.float32), np.array([0.1], np.float32), np.array([-2.0], np.float32))
    x = [a.copy() for a in pos]
    v = [a.copy() for a in vel]
    for _ in range(25):
        pos, vel = orc.step_f32(p, pos, vel)
        for c in range(3):
            v[c] = v[c] * f32(0.999998) + (f32(0.0) * f32(0.005)) * f32(2.0)
            x[c] = x[c] + v[c] * f32(0.005)
    for c in range(3):
        assert np.array_equal(pos[c], x[c]) and np.array_equal(vel[c], v[c])


def test_fp32_against_fp64_brute_force(orc, nb):
    (pos, vel) = nb.disk_ic(1024)
    idx = np.arange(1024)
    f = orc.forces_f64(pos, idx)
    p = orc.params(G=1.0, dt=1.0, damping=0.0, dist_eps=1e-7)
    zeros = tuple(np.zeros(1024, np.float32) for _ in range(3))
    _, (fx, fy, fz) = orc.step_f32(p, pos, zeros)  # v' = F
    F64 = np.stack(f)
    F32 = np.stack([fx, fy, fz]).astype(np.float64)
    scale = np.abs(F64).max(0)
    assert np.max(np.abs(F32 - F64).max(0) / scale) < 5e-5


def test_momentum_antisymmetry(orc, nb):
    pos, _ = nb.plummer_ic(2000, seed=3)
    F = np.stack(orc.forces_f64(pos))
    assert np.all(np.abs(F.sum(1)) < 1e-9 * np.abs(F).sum(1))


def test_self_skip_with_zero_eps(orc):
    pos = (np.array([0.0, 2.0], np.float32), np.zeros(2, np.float32), np.zeros(2, np.float32))
    vel = tuple(np.zeros(2, np.float32) for _ in range(3))
    p = orc.params(G=1.0, dt=1.0, damping=1.0, dist_eps=0.0)
    _, (vx, _, _) = orc.step_f32(p, pos, vel)
    assert np.isfinite(vx).all() and np.isclose(vx[0], 0.25, rtol=1e-6)
    # PREDICATED keeps only the self term: 0 * rsqrt(0) = NaN for eps = 0
    p.calc_method = 1
    _, (vx, _, _) = orc.step_f32(p, pos, vel)
    assert np.isnan(vx).all()


def test_predicated_is_pure_damping(orc, nb):
    pos, vel = nb.disk_ic(512)
    p = orc.params(**P_DEFAULT, calc_method=1)
    (nx, ny, nz), (vx, vy, vz) = orc.step_f32(p, pos, vel)
    for c, (v0, v1) in enumerate(zip(vel, (vx, vy, vz))):
        assert np.array_equal(v1, v0 * f32(0.999998))
    assert np.array_equal(nx, pos[0] + vx * f32(0.005))


def test_step_target_range(orc, nb):
    pos, vel = nb.disk_ic(3000)
    p = orc.params(**P_DEFAULT)
    full = orc.step_f32(p, pos, vel)
    part = orc.step_f32(p, pos, vel, 1000, 1777)
    for a, b in zip(full[0] + full[1], part[0] + part[1]):
        assert np.array_equal(a[1000:1777], b[1000:1777])


# ------------------------------------------------------------------ fp64
def test_two_body_energy(orc):
    d, eps, G = 2.0, 1e-3, 2.0
    pos = (np.array([0.0, d], np.float32), np.zeros(2, np.float32), np.zeros(2, np.float32))
    vel = (np.array([0.5, -0.25], np.float32), np.array([1.0, 0.0], np.float32),
           np.zeros(2, np.float32))
    p = orc.params(G=G, dist_eps=eps)
    e = orc.energy_momentum_f64(p, pos, vel)
    ke = 0.5 * (0.25 + 1.0 + 0.0625)
    pe = -G / np.sqrt(d * d + float(np.float32(eps)))
    assert np.isclose(e[0], ke, rtol=1e-12) and np.isclose(e[1], pe, rtol=1e-7)
    assert np.allclose(e[2:], [0.25, 1.0, 0.0])


def test_force_is_minus_energy_gradient(orc, nb):
    pos, _ = nb.plummer_ic(64, seed=11)
    p = orc.params(G=1.0, dist_eps=1e-2)
    F = orc.forces_f64(pos, eps=float(np.float32(1e-2)))
    zero = tuple(np.zeros(64, np.float32) for _ in range(3))
    i, h = 5, 1e-3
    for c in range(3):
        up = [a.copy() for a in pos]
        dn = [a.copy() for a in pos]
        up[c][i] += np.float32(h)
        dn[c][i] -= np.float32(h)
        hu = float(up[c][i]) - float(pos[c][i])
        hd = float(pos[c][i]) - float(dn[c][i])
        eu = orc.energy_momentum_f64(p, up, zero)[1]
        ed = orc.energy_momentum_f64(p, dn, zero)[1]
        assert np.isclose(-(eu - ed) / (hu + hd), F[c][i], rtol=1e-4, atol=1e-6)


def test_adjoint_step_matches_central_differences(orc):
    rng = np.random.default_rng(0)
    n = 48
    x = rng.normal(size=(3, n))
    v = rng.normal(size=(3, n))
    lx1 = rng.normal(size=(3, n))
    lv1 = rng.normal(size=(3, n))
    dx = rng.normal(size=(3, n))
    dv = rng.normal(size=(3, n))
    p = orc.params(G=1.5, dt=0.01, damping=0.99, dist_eps=0.05)
    lx, lv = orc.adjoint_step_f64(p, x, lx1, lv1)

    def J(h):
        xn, vn = orc.step_f64(p, x + h * dx, v + h * dv)
        return np.sum(np.stack(xn) * lx1) + np.sum(np.stack(vn) * lv1)

    h = 1e-5
    fd = (J(h) - J(-h)) / (2 * h)
    an = np.sum(np.stack(lx) * dx) + np.sum(np.stack(lv) * dv)
    assert np.isclose(an, fd, rtol=1e-7)


def test_trajectory_gradient_matches_central_differences(orc, nb):
    n, T = 96, 12
    pos, vel = nb.plummer_ic(n, seed=5, g_mass=1.0)
    x0 = np.stack(pos).astype(np.float64)
    v0 = np.stack(vel).astype(np.float64)
    rng = np.random.default_rng(1)
    target = x0 + rng.normal(scale=0.05, size=x0.shape)
    p = orc.params(G=1.0 / n, dt=0.01, damping=0.999, dist_eps=1e-3)
    L, gx, gv = orc.traj_grad_f64(p, x0, v0, target, T)
    h = 1e-6
    for (c, i) in [(0, 3), (1, 50), (2, 77)]:
        for which, g in (("x", gx), ("v", gv)):
            xp, xm, vp, vm = x0.copy(), x0.copy(), v0.copy(), v0.copy()
            if which == "x":
                xp[c, i] += h
                xm[c, i] -= h
            else:
                vp[c, i] += h
                vm[c, i] -= h
            Lp, _, _ = orc.traj_grad_f64(p, xp, vp, target, T, grad=False)
            Lm, _, _ = orc.traj_grad_f64(p, xm, vm, target, T, grad=False)
            assert np.isclose((Lp - Lm) / (2 * h), g[c, i], rtol=1e-5, atol=1e-10)


def test_golden_step_fixture(orc):
    """The oracle's one-step output for the 256-particle disk is frozen."""
    from golden import load_disk_fixture, load_step1_fixture

    d = load_disk_fixture()
    want = load_step1_fixture()
    nxt, vel = orc.step_f32(orc.params(), tuple(d[:3]), tuple(d[3:]))
    got = np.stack(nxt + vel)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
