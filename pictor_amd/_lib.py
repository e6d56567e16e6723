"""ctypes binding of libpictor_hip.so (the C ABI declared in include/pictor_hip.h).

The library is built in-tree (pictor_amd/libpictor_hip.so).  Loading it never
falls back to anything else: a missing or unloadable library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libpictor_hip.so")

PICTOR_OK = 0
PICTOR_ERR_ARG = 1
PICTOR_ERR_HIP = 2
PICTOR_ERR_BOUNDS = 3
PICTOR_ERR_CAPACITY = 12

FORM_FP64 = 0
FORM_FP32_SPLIT = 1

FDTD_YEE = 0
FDTD_LEHE = 1
GEOM_CARTESIAN = 0
GEOM_CYL = 1
GEOM_CYL_MODAL = 2

# boundary face types (PICTOR_BC_*) by the reference's names (bc_face%type_fld / type_prtl)
BC_TYPES = {"prdc": 0, "cond": 1, "open": 2, "pml": 3, "refl": 4, "remv": 5}

# every entry point of include/pictor_hip.h
EXPORTS = (
    "pictor_init", "pictor_finalize", "pictor_last_error", "pictor_get_scalars",
    "pictor_upload_fields", "pictor_download_fields", "pictor_upload_particles",
    "pictor_particle_count", "pictor_download_particles", "pictor_update_b_half",
    "pictor_move_deposit", "pictor_exchange_particles", "pictor_update_e",
    "pictor_current_halo_add", "pictor_smoothen_current", "pictor_add_current", "pictor_field_halo",
    "pictor_reset_current", "pictor_sort_tiles", "pictor_step", "pictor_synchronize",
    "pictor_diagnostics", "pictor_timer_reset", "pictor_timer_get", "pictor_memory_bytes",
    "pictor_setup_weibel", "pictor_adjoint_forward", "pictor_adjoint_backward", "pictor_adjoint_stats",
    "pictor_adjoint_engine", "pictor_adjoint_release", "pictor_cyl_plasma", "pictor_cyl_laser",
    "pictor_moving_window", "pictor_probe_atomics", "pictor_cyl_push_scheme", "pictor_cyl_sort_period",
    "pictor_save_state", "pictor_load_state", "pictor_set_bc_fld", "pictor_set_bc_prtl",
    "pictor_set_boundary_speed", "pictor_pml", "pictor_enforce_bc_final", "pictor_save_energy",
    "pictor_save_fields", "pictor_energy_noalloc", "pictor_adjoint_forward_device",
    "pictor_adjoint_backward_device", "pictor_probe_stream", "pictor_upload_particles_tagged",
    "pictor_download_particles_tagged", "pictor_filtered_efield",
)


class PictorConfig(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("epc", ctypes.c_int32), ("form", ctypes.c_int32), ("tile", ctypes.c_int32), ("fdtd", ctypes.c_int32),
        ("geometry", ctypes.c_int32), ("curr_filter", ctypes.c_int32),
        ("c", ctypes.c_double), ("c_ompe", ctypes.c_double), ("g0", ctypes.c_double), ("mi_me", ctypes.c_double),
        ("capacity_factor", ctypes.c_double), ("dr", ctypes.c_double),
    ]


class PictorError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pictor error {code}: {msg}")
        self.code = code


c_dp = ctypes.POINTER(ctypes.c_double)
c_ip = ctypes.POINTER(ctypes.c_int32)
c_vp = ctypes.c_void_p

_lib = None


def lib():
    """Load libpictor_hip.so and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is not built; run __graft_entry__.build() or make -C pictor_amd/csrc")
    L = ctypes.CDLL(LIBPATH)
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    L.pictor_init.argtypes = [ctypes.POINTER(PictorConfig), ctypes.POINTER(c_vp)]
    L.pictor_finalize.argtypes = [c_vp]
    L.pictor_last_error.argtypes = []
    L.pictor_last_error.restype = ctypes.c_char_p
    L.pictor_get_scalars.argtypes = [c_vp, c_dp]
    L.pictor_upload_fields.argtypes = [c_vp] + [c_dp] * 6
    L.pictor_download_fields.argtypes = [c_vp] + [c_dp] * 9
    L.pictor_upload_particles.argtypes = [c_vp, i64] + [c_dp] * 7 + [c_ip]
    L.pictor_particle_count.argtypes = [c_vp, ctypes.POINTER(i64)]
    L.pictor_download_particles.argtypes = [c_vp, i64] + [c_dp] * 7 + [c_ip, c_ip, ctypes.POINTER(i64)]
    L.pictor_filtered_efield.argtypes = [c_vp, i32, c_dp, c_dp, c_dp]
    L.pictor_upload_particles_tagged.argtypes = [c_vp, i64] + [c_dp] * 7 + [c_ip, c_ip]
    L.pictor_download_particles_tagged.argtypes = [c_vp, i64] + [c_dp] * 7 + [c_ip, c_ip, c_ip, ctypes.POINTER(i64)]
    for name in ("pictor_update_b_half", "pictor_move_deposit", "pictor_exchange_particles", "pictor_update_e",
                 "pictor_current_halo_add", "pictor_smoothen_current", "pictor_add_current", "pictor_reset_current", "pictor_sort_tiles",
                 "pictor_synchronize", "pictor_timer_reset"):
        getattr(L, name).argtypes = [c_vp]
    L.pictor_field_halo.argtypes = [c_vp, i32]
    L.pictor_step.argtypes = [c_vp, i32]
    L.pictor_diagnostics.argtypes = [c_vp, c_dp]
    L.pictor_timer_get.argtypes = [c_vp, i32, c_dp, ctypes.POINTER(i64)]
    L.pictor_memory_bytes.argtypes = [c_vp, ctypes.POINTER(i64)]
    L.pictor_setup_weibel.argtypes = [i32, i32, i32, i32, i32, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                      i64] + [c_dp] * 7 + [c_ip, ctypes.POINTER(i64)]
    L.pictor_adjoint_forward.argtypes = [c_vp, i64] + [c_dp] * 7 + [c_ip, i32, i32, c_dp]
    L.pictor_adjoint_backward.argtypes = [c_vp, ctypes.c_double, c_dp, c_dp, c_dp]
    # device-memory variants: raw device addresses (torch.Tensor.data_ptr())
    L.pictor_adjoint_forward_device.argtypes = [c_vp, i64] + [c_vp] * 8 + [i32, i32, c_dp]
    L.pictor_adjoint_backward_device.argtypes = [c_vp, ctypes.c_double, c_vp, c_vp, c_vp]
    L.pictor_adjoint_stats.argtypes = [c_vp, c_dp]
    L.pictor_adjoint_release.argtypes = [c_vp]
    L.pictor_adjoint_engine.argtypes = [c_vp, i32]
    L.pictor_probe_atomics.argtypes = [i32, c_dp]
    L.pictor_probe_stream.argtypes = [i64, c_dp]
    L.pictor_cyl_plasma.argtypes = [c_vp, i32, i32, i32, i32, ctypes.c_uint64]
    L.pictor_cyl_push_scheme.argtypes = [c_vp, i32]
    L.pictor_cyl_sort_period.argtypes = [c_vp, i32]
    L.pictor_cyl_laser.argtypes = [c_vp] + [ctypes.c_double] * 5
    L.pictor_moving_window.argtypes = [c_vp, i32, i32, i32, ctypes.c_uint64]
    L.pictor_save_state.argtypes = [c_vp, ctypes.c_char_p, i64]
    L.pictor_load_state.argtypes = [c_vp, ctypes.c_char_p, ctypes.POINTER(i64)]
    L.pictor_set_bc_fld.argtypes = [c_vp, i32, i32, i32, ctypes.c_double, ctypes.c_double]
    L.pictor_set_bc_prtl.argtypes = [c_vp, i32, i32, i32, ctypes.c_double]
    L.pictor_set_boundary_speed.argtypes = [c_vp, i32, i32, ctypes.c_double]
    L.pictor_pml.argtypes = [c_vp, i32]
    L.pictor_enforce_bc_final.argtypes = [c_vp]
    L.pictor_save_energy.argtypes = [c_vp, ctypes.c_char_p]
    L.pictor_save_fields.argtypes = [c_vp, ctypes.c_char_p, i32, i32]
    L.pictor_energy_noalloc.argtypes = [c_vp, c_dp]
    for name in EXPORTS:
        if name != "pictor_last_error":
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != PICTOR_OK:
        msg = lib().pictor_last_error()
        raise PictorError(rc, msg.decode() if msg else "")
    return rc
