"""C-ABI contract tests that need no GPU: the libraries load, every function
the headers in include/ declare is exported, parameter handling matches the
reference's compiled SimParam, and errors come back as status codes."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"
REF_SIMPARAM = ROOT / "oracle" / "_ref" / "libsimparam_ref.so"


def exported(lib: Path, demangle: bool = False) -> set:
    flags = ["-D", "--defined-only"] + (["-C"] if demangle else [])
    out = subprocess.run(["nm", *flags, str(lib)], capture_output=True, text=True, check=True)
    names = set()
    for line in out.stdout.splitlines():
        parts = line.split(maxsplit=2)
        if len(parts) == 3 and parts[1] in ("T", "W"):
            names.add(parts[2])
    return names


def c_functions(header: Path) -> list:
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return re.findall(r"^[A-Za-z_][\w \*]*?\b(nb_\w+)\s*\(", text, flags=re.M)


def test_every_c_declaration_is_exported(nb):
    have = exported(nb.LIB_PATH) | exported(nb.SIM_LIB_PATH)
    decls = []
    for h in sorted(INCLUDE.glob("*.h")):
        decls += c_functions(h)
    assert len(decls) >= 20
    missing = [d for d in decls if d not in have]
    assert not missing, missing


def test_cpp_surface_is_exported(nb):
    have = exported(nb.SIM_LIB_PATH, demangle=True)
    want = [
        "SimParam::SimParam()",
        "SimParam::parseArgs(int, char**)",
        "simulation::DiskGalaxySimulator::DiskGalaxySimulator(SimParam)",
        "simulation::DiskGalaxySimulator::stepSim()",
        "simulation::DiskGalaxySimulator::getParticlePos()",
        "simulation::DiskGalaxySimulator::getParticleVel()",
        "simulation::DiskGalaxySimulator::getDeviceName[abi:cxx11]()",
        "simulation::length(simulation::vec3)",
        "simulation::cross(simulation::vec3, simulation::vec3)",
        "simulation::normalize(simulation::vec3)",
    ]
    missing = [w for w in want if w not in have]
    assert not missing, missing
    assert any(s.startswith("getCalculationMethod(") for s in have)
    assert any(s.startswith("simulation::DiskGalaxySimulator::DiskGalaxySimulator(SimParam, "
                            "simulation::ParticleData") for s in have)


def _sim_lib():
    import nbody_amd as nb

    nb.load()
    lib = ctypes.CDLL(str(nb.SIM_LIB_PATH))
    lib.nb_simparam_parse.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                      ctypes.POINTER(nb.NbParams)]
    return lib


def _argv(args):
    arr = (ctypes.c_char_p * (len(args) + 1))(*[a.encode() for a in args], None)
    return len(args), arr


def _ours(args):
    import nbody_amd as nb

    p = nb.NbParams()
    argc, argv = _argv(args)
    st = _sim_lib().nb_simparam_parse(argc, argv, ctypes.byref(p))
    return st, p


def _reference(args):
    lib = ctypes.CDLL(str(REF_SIMPARAM))
    f, i, u = ctypes.c_float, ctypes.c_int, ctypes.c_ulonglong
    vals = [f(), f(), u(), u(), i(), f(), f(), i(), i()]
    argc, argv = _argv(args)
    st = lib.ref_simparam(argc, argv, *[ctypes.byref(v) for v in vals])
    return st, [v.value for v in vals]


ARGV_CASES = [
    ["nbody"],
    ["nbody", "100"],
    ["nbody", "1024", "5", "0.999", "0.001", "1.0e-3", "2.0"],
    ["nbody", "50", "4", "0.999998", "0.005", "1e-7", "2", "10", "128", "PREDICATED"],
    ["nbody", "7x", "3.9", "abc", "-1", "0", "1e3", "3", "0", "BRANCH"],
]


REFERENCE_STYLE_SIMULATOR = r"""
// A simulator written against the reference interface (src/simulator.cuh:
// 107-116): it overrides exactly the seven pure virtuals and nothing else.
#include <type_traits>
#include "simulator.hpp"

namespace {
class OwnSimulator : public simulation::Simulator {
public:
  void stepSim() override { ++steps; }
  size_t getNumParticles() override { return 3; }
  const simulation::ParticleData &getParticlePos() override { return pos; }
  const simulation::ParticleData &getParticleVel() override { return vel; }
  float getLastStepTime() override { return 1.5f; }
  const std::string *getDeviceName() override { return &name; }
  int getGwSize() override { return 64; }
  int steps = 0;

private:
  simulation::ParticleData pos{3}, vel{3};
  std::string name{"own"};
};

template <class T, class = void> struct has_get_cm : std::false_type {};
template <class T>
struct has_get_cm<T, decltype(void(&T::getCM))> : std::true_type {};
} // namespace

// getCM is DiskGalaxySimulator's own non-virtual member (src/simulator.cuh:140)
static_assert(!has_get_cm<simulation::Simulator>::value, "getCM is not in the interface");
static_assert(has_get_cm<simulation::DiskGalaxySimulator>::value, "getCM on the simulator");
static_assert(!std::is_polymorphic<SimParam>::value, "SimParam is a plain struct");

int main() {
  OwnSimulator own;
  simulation::Simulator &s = own;  // usable through the interface, as a renderer does
  s.stepSim();
  s.stepSim();
  return (own.steps == 2 && s.getNumParticles() == 3 && s.getGwSize() == 64 &&
          s.getParticlePos().x.size() == 3 && *s.getDeviceName() == "own") ? 0 : 1;
}
"""


def test_reference_style_simulator_compiles(tmp_path):
    """A class that implements exactly the reference interface's seven pure
    virtuals compiles against include/simulator.hpp and runs through a
    Simulator reference."""
    src = tmp_path / "own.cpp"
    src.write_text(REFERENCE_STYLE_SIMULATOR)
    exe = tmp_path / "own"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{INCLUDE}", str(src), "-o",
                    str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


@pytest.mark.parametrize("args", ARGV_CASES)
def test_simparam_matches_reference_build(nb, args):
    if not REF_SIMPARAM.exists():
        pytest.skip("reference SimParam not built (needs the reference checkout)")
    st_ref, ref = _reference(args)
    st, p = _ours(args)
    assert st == st_ref == 0
    ours = [p.G, p.dt, p.num_particles, p.num_frames, p.sim_iterations_per_frame, p.damping,
            p.dist_eps, p.gw_size, p.calc_method]
    assert ours == ref


def test_simparam_rejects_unknown_method(nb):
    args = ["nbody", "1", "1", "1", "1", "1", "1", "1", "64", "FAST"]
    st, _ = _ours(args)
    assert st == nb.NB_ERR_INVALID_ARG
    if REF_SIMPARAM.exists():
        assert _reference(args)[0] == 1  # std::invalid_argument in the reference


def test_defaults_match_c_abi(nb):
    st, p = _ours(["nbody"])
    q = nb.default_params()
    assert st == 0
    for f, _ in nb.NbParams._fields_:
        assert getattr(p, f) == getattr(q, f), f


def test_invalid_arguments_are_status_codes(nb):
    lib = nb.load()
    p = nb.default_params(calc_method=7)
    h = ctypes.c_void_p()
    z = np.zeros(4, np.float32)
    ptr = ctypes.c_void_p(z.ctypes.data)
    st = lib.nb_sim_create(ctypes.byref(p), *[ptr] * 6, ctypes.byref(h))
    assert st == nb.NB_ERR_INVALID_ARG
    assert "BRANCH or PREDICATED" in lib.nb_last_error().decode()
    assert lib.nb_forward_config_info(10**6, None, None, None, None) == nb.NB_ERR_INVALID_ARG
    p = nb.default_params(num_particles=4)
    st = lib.nb_sim_create(ctypes.byref(p), None, None, None, None, None, None, ctypes.byref(h))
    assert st == nb.NB_ERR_INVALID_ARG
    st = lib.nb_forward_step_dev(ctypes.byref(p), *[None] * 9, 4, -1, 0, None)
    assert st == nb.NB_ERR_INVALID_ARG
    st = lib.nb_forward_step_dev(ctypes.byref(p), *[ptr] * 9, 4, -1, 9, None)
    assert st == nb.NB_ERR_INVALID_ARG  # unknown accumulation mode
    assert lib.nb_ic_disk(4, 1, 0, None, None, None, None, None, None) == nb.NB_ERR_INVALID_ARG
    assert lib.nb_ic_disk(0, 1, 0, None, None, None, None, None, None) == nb.NB_OK


def test_no_device_is_reported(nb):
    if nb.device_count() > 0:
        pytest.skip("a GPU is visible")
    pos, vel = nb.disk_ic(64)
    with pytest.raises(nb.NbError) as e:
        nb.Simulator(nb.default_params(num_particles=64), pos, vel)
    assert e.value.code == nb.NB_ERR_NO_DEVICE
    assert nb.device_name() == "Unknown Device"


def test_forward_configs_are_described(nb):
    n = nb.forward_config_count()
    assert n >= 4
    schemes = set()
    for c in range(n):
        info = nb.forward_config_info(c)
        assert info["split"] >= 1
        if info["scheme"] == "targets":  # the block stages whole LDS tiles
            assert info["tile"] % info["block"] == 0
        schemes.add(info["scheme"])
        if info["scheme"] == "symmetric":  # 4 waves, super-tile = 4 x 64 x unroll
            assert info["block"] == 256 and info["tile"] == 256 * info["unroll"]
        if info["scheme"] == "symmetric_tiles":  # one tile of 64 x unroll per wave
            assert info["block"] == 256 and info["tile"] == 64 * info["unroll"]
    assert nb.forward_config_info(0)["scheme"] == "targets"
    assert schemes == {"targets", "symmetric", "symmetric_tiles"}
    lib = nb.load()
    assert lib.nb_forward_config_scheme(n) == -1 and lib.nb_forward_config_scheme(-1) == -1


def test_default_shapes_are_reported(nb):
    """The shapes cfg = -1 picks (host logic, no GPU): symmetric 8 per lane at
    N = 262144, tile pairs of 2 per lane at the reference's default 12800,
    target-sharded below 4096, for PREDICATED, and for a NULL params."""
    info = lambda c: nb.forward_config_info(c)
    big = info(nb.forward_default_config(262144))
    assert (big["scheme"], big["unroll"], big["split"]) == ("symmetric", 8, 1)
    small = info(nb.forward_default_config(12800))
    assert (small["scheme"], small["unroll"]) == ("symmetric_tiles", 2)
    # two items per tile pair below 6656 particles, one from there
    for n, split in ((4096, 2), (6655, 2), (6656, 1), (8192, 1)):
        tp = info(nb.forward_default_config(n))
        assert (tp["scheme"], tp["unroll"], tp["split"]) == ("symmetric_tiles", 2, split), n
    assert info(nb.forward_default_config(1000))["scheme"] == "targets"
    pred = nb.default_params(num_particles=262144, calc_method=nb.NB_CALC_PREDICATED)
    assert info(nb.forward_default_config(262144, pred))["scheme"] == "targets"
    kahan = info(nb.forward_default_config(262144, acc=nb.NB_ACC_KAHAN))
    assert (kahan["scheme"], kahan["unroll"]) == ("symmetric", 2)
    adj = nb.adjoint_config_info(nb.adjoint_default_config(262144))
    assert (adj["scheme"], adj["unroll"]) == ("symmetric", 2)
    lib = nb.load()
    assert lib.nb_forward_default_config(None, 100, 0) == -1
    assert lib.nb_adjoint_default_config(None, 100) == -1
    with pytest.raises(nb.NbError):
        nb.forward_default_config(100, acc=7)


def test_adjoint_configs_are_described(nb):
    n = nb.adjoint_config_count()
    schemes = [nb.adjoint_config_info(c)["scheme"] for c in range(n)]
    assert schemes[0] == "targets" and "symmetric" in schemes
    for c in range(n):  # every symmetric adjoint shape has a forward twin
        info = nb.adjoint_config_info(c)
        if info["scheme"] != "targets":
            assert any(nb.forward_config_info(f) == info
                       for f in range(nb.forward_config_count()))
    assert nb.load().nb_adjoint_config_scheme(n) == -1


def test_driver_rejects_bad_method(nb):
    res = subprocess.run([str(nb.DRIVER_PATH), "1", "1", "1", "1", "1", "1", "1", "64", "X"],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode == 1 and "BRANCH or PREDICATED" in res.stderr


def test_plummer_invariants(nb):
    (x, y, z), (vx, vy, vz) = nb.plummer_ic(20000, seed=9, a=1.0, g_mass=1.0)
    pos = np.stack([x, y, z]).astype(np.float64)
    vel = np.stack([vx, vy, vz]).astype(np.float64)
    assert np.allclose(pos.mean(1), 0, atol=1e-6) and np.allclose(vel.mean(1), 0, atol=1e-6)
    r = np.linalg.norm(pos, axis=0)
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1) = 1.305 a
    assert abs(np.median(r) - 1.305) < 0.05
    # virial equilibrium: 2K + W = 0 with W = -(3 pi / 32) G M^2 / a (G M = 1, M = 1)
    ke = 0.5 * np.mean(np.sum(vel * vel, 0))
    assert abs(2 * ke - 3 * np.pi / 32) < 0.02
    # bounded: every particle below the local escape speed
    vesc = np.sqrt(2.0) * (r * r + 1.0) ** -0.25
    assert np.all(np.linalg.norm(vel, axis=0) <= vesc * (1 + 1e-5) + 1e-6)
