"""ctypes binding of libnbody_amd (include/nbody_amd.h).

The binding is thin: every call goes to the C ABI, device buffers are passed
as raw pointers (PyTorch tensors' data_ptr()), streams as hipStream_t
handles.  If the shared library is missing the import fails loudly; there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_float, c_int, c_size_t, c_uint32, c_void_p
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_DIR = PKG_ROOT / "lib"
LIB_PATH = LIB_DIR / "libnbody_amd.so"
SIM_LIB_PATH = LIB_DIR / "libnbody_sim.so"
DRIVER_PATH = LIB_DIR / "nbody_amd"

NB_OK = 0
NB_ERR_INVALID_ARG = 1
NB_ERR_NO_DEVICE = 2
NB_ERR_HIP = 3
NB_ERR_OUT_OF_MEMORY = 4
NB_ERR_UNSUPPORTED = 5

NB_CALC_BRANCH = 0
NB_CALC_PREDICATED = 1
NB_ACC_PLAIN = 0
NB_ACC_KAHAN = 1
NB_IC_TRIG_DOUBLE = 1
NB_IC_LENGTH_F32 = 2

DEFAULT_SEED = 5489  # std::mt19937's default seed (the reference's `std::mt19937 gen;`)


class NbParams(ctypes.Structure):
    """nb_params: SimParam field for field (src/sim_param.hpp:30-39)."""

    _fields_ = [
        ("G", c_float),
        ("dt", c_float),
        ("num_particles", c_size_t),
        ("num_frames", c_size_t),
        ("sim_iterations_per_frame", c_int),
        ("damping", c_float),
        ("dist_eps", c_float),
        ("gw_size", c_int),
        ("calc_method", c_int),
    ]


class NbVec3(ctypes.Structure):
    """nb_vec3: three device arrays of one SoA vector field."""

    _fields_ = [("x", c_void_p), ("y", c_void_p), ("z", c_void_p)]


class NbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nbody_amd error {code}: {msg}")
        self.code = code


_FP = POINTER(c_float)
_lib: ctypes.CDLL | None = None


def _sig(lib: ctypes.CDLL, name: str, restype, argtypes) -> None:
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes


def load() -> ctypes.CDLL:
    """Loads libnbody_amd.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: PyTorch loads its bundled libamdhip64 by
    # path, and a second runtime loaded first (this library's /opt/rocm one)
    # would not be shared with it.  Importing torch first makes this
    # library's libamdhip64.so.7 dependency resolve to torch's copy.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python cuda-to-sycl-nbody_amd/build.py` "
            "(the HIP library is required; there is no fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp = c_void_p
    _sig(lib, "nb_params_default", None, [POINTER(NbParams)])
    _sig(lib, "nb_last_error", c_char_p, [])
    _sig(lib, "nb_device_count", c_int, [])
    _sig(lib, "nb_device_name", c_int, [c_char_p, c_size_t])
    _sig(lib, "nb_ic_disk", c_int, [c_size_t, c_uint32, c_int] + [vp] * 6)
    _sig(lib, "nb_ic_disk_gen", c_int, [c_size_t, c_uint32] + [vp] * 6)
    _sig(lib, "nb_ic_plummer", c_int, [c_size_t, c_uint32, c_float, c_float] + [vp] * 6)
    _sig(lib, "nb_sim_create", c_int, [POINTER(NbParams)] + [vp] * 6 + [POINTER(c_void_p)])
    _sig(lib, "nb_sim_step", c_int, [vp])
    _sig(lib, "nb_sim_advance", c_int, [vp, c_int])
    _sig(lib, "nb_sim_sync", c_int, [vp])
    _sig(lib, "nb_sim_read_pos", c_int, [vp, vp, vp, vp])
    _sig(lib, "nb_sim_read_vel", c_int, [vp, vp, vp, vp])
    _sig(lib, "nb_sim_last_step_ms", c_float, [vp])
    _sig(lib, "nb_sim_last_kernel_ms", c_float, [vp])
    _sig(lib, "nb_sim_num_particles", c_size_t, [vp])
    _sig(lib, "nb_sim_calc_method", c_int, [vp])
    _sig(lib, "nb_sim_gw_size", c_int, [vp])
    _sig(lib, "nb_sim_set_accumulation", c_int, [vp, c_int])
    _sig(lib, "nb_sim_set_frame_iterations", c_int, [vp, c_int])
    _sig(lib, "nb_sim_set_config", c_int, [vp, c_int])
    _sig(lib, "nb_sim_set_graph", c_int, [vp, c_int])
    _sig(lib, "nb_sim_config", c_int, [vp])
    _sig(lib, "nb_sim_stage_readback", c_int, [vp])
    _sig(lib, "nb_sim_staged", c_int, [vp, POINTER(c_void_p)])
    _sig(lib, "nb_sim_read_vec4", c_int, [vp, c_int, c_float, vp])
    _sig(lib, "nb_sim_destroy", c_int, [vp])
    _sig(lib, "nb_forward_step_dev", c_int,
         [POINTER(NbParams)] + [vp] * 9 + [c_size_t, c_int, c_int, vp])
    _sig(lib, "nb_forward_step_noalloc_dev", c_int,
         [POINTER(NbParams)] + [vp] * 9 + [c_size_t, c_int, c_int, vp])
    V3 = POINTER(NbVec3)
    _sig(lib, "nb_adjoint_step_dev", c_int,
         [POINTER(NbParams), V3, V3, V3, V3, V3, c_size_t, c_int, vp])
    _sig(lib, "nb_fwd_adjoint_step_dev", c_int,
         [POINTER(NbParams), V3, V3, V3, V3, V3, V3, V3, c_size_t, c_int, vp])
    _sig(lib, "nb_adjoint_config_count", c_int, [])
    _sig(lib, "nb_adjoint_config_info", c_int, [c_int] + [POINTER(c_int)] * 4)
    _sig(lib, "nb_adjoint_config_scheme", c_int, [c_int])
    _sig(lib, "nb_traj_checkpoint_floats", c_size_t, [c_size_t, c_int, c_int])
    _sig(lib, "nb_traj_workspace_floats", c_size_t, [c_size_t, c_int])
    _sig(lib, "nb_traj_forward_dev", c_int,
         [POINTER(NbParams), vp, vp, c_int, c_int, vp, vp, c_size_t, vp])
    _sig(lib, "nb_traj_backward_dev", c_int,
         [POINTER(NbParams), vp, vp, vp, c_int, c_int, vp, c_size_t, vp])
    _sig(lib, "nb_batch_create", c_int,
         [c_int, POINTER(NbParams)] + [vp] * 6 + [POINTER(c_void_p)])
    for name in ("nb_batch_step", "nb_batch_step_direct", "nb_batch_step_separate",
                 "nb_batch_set_config"):
        _sig(lib, name, c_int, [vp, c_int])
    _sig(lib, "nb_batch_autotune", c_int, [vp, c_int, POINTER(c_int)])
    _sig(lib, "nb_batch_config", c_int, [vp])
    _sig(lib, "nb_batch_config_count", c_int, [])
    _sig(lib, "nb_batch_config_info", c_int, [c_int] + [POINTER(c_int)] * 3)
    _sig(lib, "nb_batch_config_scheme", c_int, [c_int])
    _sig(lib, "nb_batch_last_ms", c_float, [vp])
    _sig(lib, "nb_batch_num_particles", c_size_t, [vp])
    _sig(lib, "nb_batch_read_pos", c_int, [vp, vp, vp, vp])
    _sig(lib, "nb_batch_read_vel", c_int, [vp, vp, vp, vp])
    _sig(lib, "nb_batch_destroy", c_int, [vp])
    _sig(lib, "nb_pack_vec4_dev", c_int, [V3, c_float, vp, c_size_t, vp])
    _sig(lib, "nb_forces_f64_dev", c_int,
         [POINTER(NbParams), V3, vp, c_size_t, vp, vp, vp, c_size_t, vp])
    _sig(lib, "nb_energy_momentum_dev", c_int,
         [POINTER(NbParams), V3, V3, c_size_t, vp, vp])
    _sig(lib, "nb_forward_config_count", c_int, [])
    _sig(lib, "nb_forward_config_info", c_int, [c_int] + [POINTER(c_int)] * 4)
    _sig(lib, "nb_forward_config_scheme", c_int, [c_int])
    _sig(lib, "nb_forward_default_config", c_int, [POINTER(NbParams), c_size_t, c_int])
    _sig(lib, "nb_adjoint_default_config", c_int, [POINTER(NbParams), c_size_t])
    _sig(lib, "nb_scratch_release", c_int, [])
    _sig(lib, "nb_scratch_held", c_int, [POINTER(c_size_t)])
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != NB_OK:
        msg = load().nb_last_error().decode(errors="replace")
        raise NbError(status, msg)


def default_params(**overrides) -> NbParams:
    p = NbParams()
    load().nb_params_default(ctypes.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise AttributeError(f"nb_params has no field {k}")
        setattr(p, k, v)
    return p
