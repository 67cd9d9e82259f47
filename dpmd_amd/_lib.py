"""ctypes binding of libdpmd_amd.so (the C ABI declared in include/dpmd/dpmd.h).

The shared library is built in-tree by ``dpmd_amd/build.py``.  Loading fails loudly when it
is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libdpmd_amd.so"

# ---- enums (mirror include/dpmd/dpmd.h) -------------------------------------------------
DISK, RIGID_BUMPY = 0, 1
NEIGHBOR_NAIVE, NEIGHBOR_CELL = 0, 1
FORCE_PBC, FORCE_WALL, FORCE_PER_SYSTEM = 0, 1, 2
ORDER_CURRENT, ORDER_CANONICAL = 0, 1
FIRE_PRED_REFERENCE, FIRE_PRED_MAX_FORCE, FIRE_PRED_FORCE_OR_PE = 0, 1, 2
OPT_COOP_REBUILD, OPT_COOP_IN_GRAPH, OPT_FUSED_VV, OPT_BLOCK, OPT_MAX_BATCH = 0, 1, 2, 3, 4
OPT_PERSISTENT_FIRE, OPT_PERSIST_BLOCK, OPT_PERSIST_PHASE, OPT_DEVICE_REBUILD, OPT_PERSIST_LDS = 5, 6, 7, 8, 9
OPT_REBUILD_ENGINE, OPT_JAM_RETRIES, OPT_FUSED_FIRE = 10, 11, 12
TUNE_FIRE, TUNE_NVE = 0, 1

STATUS = {0: "OK", 1: "ARGUMENT", 2: "STATE", 3: "HIP", 4: "OVERFLOW"}

F = dict(
    POS_X=0, POS_Y=1, VEL_X=2, VEL_Y=3, FORCE_X=4, FORCE_Y=5, PE=6, KE=7, RAD=8, MASS=9, AREA=10,
    LAST_POS_X=11, LAST_POS_Y=12, ANGLE=13, ANGULAR_VEL=14, TORQUE=15, MOMENT_INERTIA=16,
    STRESS_XX=17, STRESS_XY=18, STRESS_YX=19, STRESS_YY=20, OVERLAP=21, CONTACT_COUNT=22,
    BOX_SIZE_X=32, BOX_SIZE_Y=33, BOX_INV_X=34, BOX_INV_Y=35, E_INTERACTION=36, NEIGHBOR_CUTOFF=37,
    THRESH2=38, PE_TOTAL=39, KE_TOTAL=40, FPOWER_TOTAL=41, PRESSURE=42, TEMPERATURE=43,
    PACKING_FRACTION=44, MAX_FORCE2=45, STRESS_TOTAL_XX=46, STRESS_TOTAL_XY=47, STRESS_TOTAL_YX=48,
    STRESS_TOTAL_YY=49,
    SYSTEM_OFFSET=64, SYSTEM_SIZE=65, CELL_DIM_X=66, CELL_DIM_Y=67, CELL_SYSTEM_START=68, BOUNDARY=69,
    N_DOF=70, REBUILD_FLAG=71, FIRE_STEPS=72, N_CONTACTS_TOTAL=73,
    SYSTEM_ID=80, STATIC_INDEX=81, CELL_ID=82, NEIGHBOR_COUNT=83, CONTACTS=84,
    VERTEX_POS_X=96, VERTEX_POS_Y=97, VERTEX_FORCE_X=98, VERTEX_FORCE_Y=99, VERTEX_PE=100,
    VERTEX_RAD=101, VERTEX_MASS=102,
    VERTEX_PARTICLE_ID=112, PARTICLE_OFFSET=113, N_VERTICES_PER_PARTICLE=114, VERTEX_SYSTEM_ID=115,
    VERTEX_SYSTEM_OFFSET=116,
)
INT_FIELDS = {k for k, v in F.items() if v >= 64 and v < 96 or v >= 112}

vp, i32, i64, f64, u64 = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_uint64
pi32 = C.POINTER(C.c_int)
pi64 = C.POINTER(C.c_int64)


class TuneConfig(C.Structure):
    """struct dpmd_tune_config"""
    _fields_ = [("skin", C.c_double), ("cell_factor", C.c_double), ("block", C.c_int), ("max_batch", C.c_int),
                ("fused_vv", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneTrial(C.Structure):
    """struct dpmd_tune_trial"""
    _fields_ = [("config", TuneConfig), ("rate", C.c_double), ("seconds", C.c_double)]


# name: (restype, argtypes)
PROTOS = {
    "dpmd_ctx_create": (i32, [C.POINTER(vp), i32, vp]),
    "dpmd_ctx_destroy": (i32, [vp]),
    "dpmd_last_error": (C.c_char_p, [vp]),
    "dpmd_allocate_systems": (i32, [vp, i32]),
    "dpmd_allocate_particles": (i32, [vp, i32]),
    "dpmd_allocate_vertices": (i32, [vp, i32]),
    "dpmd_set_neighbor_method": (i32, [vp, i32]),
    "dpmd_ctx_info": (i32, [vp, pi32, pi32]),
    "dpmd_sizes": (i32, [vp, pi32, pi32, pi32, pi32]),
    "dpmd_field": (i32, [vp, i32, C.POINTER(vp), pi64]),
    "dpmd_set_field": (i32, [vp, i32, vp, i64]),
    "dpmd_get_field": (i32, [vp, i32, vp, i64, i32]),
    "dpmd_sync": (i32, [vp]),
    "dpmd_init_neighbors": (i32, [vp]),
    "dpmd_update_neighbors": (i32, [vp]),
    "dpmd_check_neighbors": (i32, [vp, i32]),
    "dpmd_neighbor_info": (i32, [vp, pi32, pi32, pi64]),
    "dpmd_get_neighbors_csr": (i32, [vp, vp, vp, i64]),
    "dpmd_set_neighbor_capacity": (i32, [vp, i32]),
    "dpmd_compute_forces": (i32, [vp, i32]),
    "dpmd_update_positions": (i32, [vp, vp, f64]),
    "dpmd_update_velocities": (i32, [vp, vp, f64]),
    "dpmd_scale_velocities": (i32, [vp, vp]),
    "dpmd_scale_positions": (i32, [vp, vp]),
    "dpmd_mix_velocities_and_forces": (i32, [vp, vp]),
    "dpmd_zero_velocities": (i32, [vp]),
    "dpmd_compute_ke": (i32, [vp]),
    "dpmd_compute_ke_total": (i32, [vp]),
    "dpmd_compute_pe_total": (i32, [vp]),
    "dpmd_compute_fpower_total": (i32, [vp]),
    "dpmd_compute_temperature": (i32, [vp]),
    "dpmd_compute_packing_fraction": (i32, [vp]),
    "dpmd_compute_stress_tensor": (i32, [vp]),
    "dpmd_compute_pressure": (i32, [vp]),
    "dpmd_compute_max_force": (i32, [vp]),
    "dpmd_set_average_velocity": (i32, [vp, vp, vp]),
    "dpmd_set_temperature": (i32, [vp, vp, i32]),
    "dpmd_increment_packing_fraction": (i32, [vp, vp]),
    "dpmd_save_state": (i32, [vp, vp, i32]),
    "dpmd_restore_state": (i32, [vp, vp, i32]),
    "dpmd_rand_uniform_2d": (i32, [vp, i32, i32, u64, u64, f64, f64, f64, f64]),
    "dpmd_rand_uniform_1d": (i32, [vp, i32, u64, u64, f64, f64]),
    "dpmd_rand_normal_2d": (i32, [vp, i32, i32, u64, u64, f64, f64, f64, f64]),
    "dpmd_vv_init": (i32, [vp, vp, i32]),
    "dpmd_vv_step": (i32, [vp, i32]),
    "dpmd_vv_run": (i32, [vp, i32, i32]),
    "dpmd_damped_vv_init": (i32, [vp, vp, vp, i32]),
    "dpmd_compute_damping_forces": (i32, [vp, vp]),
    "dpmd_compute_contacts": (i32, [vp]),
    "dpmd_compute_overlaps": (i32, [vp]),
    "dpmd_compute_n_contacts_total": (i32, [vp]),
    "dpmd_compute_pair_observables": (i32, [vp, vp, vp, vp, vp, vp, vp, i64]),
    "dpmd_particle_neighbors": (i32, [vp, vp, vp, i64, pi64]),
    "dpmd_compute_rigid_hessian": (i32, [vp, vp, i64]),
    "dpmd_compute_rigid_pair_observables": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i64]),
    "dpmd_nvt_rescale_run": (i32, [vp, i32, i32, vp, i32, vp, i32, i32, i32]),
    "dpmd_jam_binary_search": (i32, [vp, vp, i32, i32, f64, f64, f64, f64, i32, i32, vp, vp, pi32]),
    "dpmd_fire_init": (i32, [vp, vp, i32]),
    "dpmd_fire_step": (i32, [vp]),
    "dpmd_fire_converged": (i32, [vp, f64, f64, pi32]),
    "dpmd_fire_converged_max_force": (i32, [vp, f64, pi32]),
    "dpmd_fire_get_dt": (i32, [vp, vp]),
    "dpmd_fire_state": (i32, [vp, vp, vp, vp, vp]),
    "dpmd_fire_minimize": (i32, [vp, vp, i32, i32, f64, f64, i32, i32, i32, pi32, pi32]),
    "dpmd_adjoint_hvp": (i32, [vp, i32, vp, vp, vp, vp]),
    "dpmd_adjoint_hvp_noalloc": (i32, [vp, i32, vp, vp, vp, vp]),
    "dpmd_adjoint_vjp": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "dpmd_adjoint_solve": (i32, [vp, i32, vp, vp, vp, vp, f64, i32, pi32, C.POINTER(C.c_double)]),
    "dpmd_adjoint_pressure": (i32, [vp, i32, f64, i32, vp, vp, vp, pi32, C.POINTER(C.c_double)]),
    "dpmd_newton_minimize": (i32, [vp, i32, i32, f64, f64, i32, C.POINTER(C.c_double)]),
    "dpmd_vv_run_checkpointed": (i32, [vp, i32, i32, i32]),
    "dpmd_vv_backprop": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32]),
    "dpmd_jam_compress": (i32, [vp, vp, f64, f64, i32, i32, i32, f64, f64, i32, i32, pi32, pi32,
                                C.POINTER(C.c_int64)]),
    "dpmd_malloc": (i32, [C.POINTER(vp), i64]),
    "dpmd_free": (i32, [vp]),
    "dpmd_memcpy": (i32, [vp, vp, i64, i32]),
    "dpmd_fill_f64": (i32, [vp, i64, f64]),
    "dpmd_fill_i32": (i32, [vp, i64, C.c_int32]),
    "dpmd_synchronize": (i32, [vp]),
    "dpmd_set_option": (i32, [vp, i32, i32]),
    "dpmd_profile_durations": (i32, [vp, vp, i32, pi32]),
    "dpmd_get_option": (i32, [vp, i32, pi32]),
    "dpmd_profile": (i32, [vp, i32]),
    "dpmd_profile_read": (i32, [vp, C.POINTER(f64), pi32]),
    "dpmd_stream_copy": (i32, [vp, vp, i64, i32, C.POINTER(f64)]),
    "dpmd_cell_grid": (i32, [f64, f64, f64, f64, f64, pi32, pi32]),
    "dpmd_tune_get": (i32, [vp, C.POINTER(TuneConfig)]),
    "dpmd_tune_apply": (i32, [vp, C.POINTER(TuneConfig)]),
    "dpmd_tune_signature": (i32, [vp, i32, i32, C.c_char_p, i32]),
    "dpmd_autotune": (i32, [vp, i32, vp, i32, i32, C.c_char_p, C.POINTER(TuneConfig), C.POINTER(f64),
                            C.POINTER(TuneTrial), i32, pi32]),
    "dpmd_tune_cache_load": (i32, [C.c_char_p, C.c_char_p, C.POINTER(TuneConfig), C.POINTER(f64), pi32]),
    "dpmd_tune_cache_store": (i32, [C.c_char_p, C.c_char_p, C.POINTER(TuneConfig), f64]),
    "dpmd_version": (C.c_char_p, []),
}


class DpmdError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[{STATUS.get(status, status)}] {msg}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    """Load libdpmd_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python dpmd_amd/build.py` (no CPU fallback exists)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, ctx=None) -> None:
    if status != 0:
        msg = lib().dpmd_last_error(ctx)
        raise DpmdError(status, msg.decode() if msg else "")
