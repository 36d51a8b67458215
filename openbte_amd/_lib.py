"""ctypes binding of libopenbte_amd.so (the C ABI declared in include/openbte_amd.h).

The library is built in-tree by ``openbte_amd.build``; loading never builds.
A missing library raises ``RuntimeError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libopenbte_amd.so')

_i32 = C.c_int32
_i64 = C.c_int64
_f64 = C.c_double
_p = C.c_void_p


class BtePlan(C.Structure):
    """Mirror of obte_bte_plan_t (field order matters)."""
    _fields_ = [
        ('n', _i32), ('n_dirs', _i32), ('n_mfp', _i32), ('dim', _i32), ('mc', _i32), ('n_chunks', _i32),
        ('block', _i32), ('n_cut', _i32), ('n_wall', _i32), ('n_contact', _i32), ('n_faces', _i32),
        ('contact_kind', _i32), ('normalization', _f64), ('contact_jump', _f64),
        ('order', _p), ('pos', _p), ('lvl_off', _p), ('lvl_ptr', _p), ('kin', _i32), ('ncut_max', _i32),
        ('lvl_pair_max', _i32), ('sweep_window', _i32),
        ('nb_src', _p), ('nb_face', _p), ('cut_off', _p), ('cut_exit', _p),
        ('dirs', _p), ('k_face', _p), ('jump_face', _p), ('cf_ptr', _p), ('cf_face', _p),
        ('wall_ptr', _p), ('wall_side', _p), ('wall_db', _p), ('wall_temp', _p), ('wall_cell', _p),
        ('wall_rinv', _p), ('wall_dbs', _p), ('ce1', _p), ('ce2', _p), ('c_na', _p), ('c_t', _p), ('t_face', _p),
        ('mfps', _p),
        ('t_coeff', _p), ('sigma', _p), ('heat', _p), ('nano', _f64),
        ('pert', _p), ('rec', _p),
        ('delta_t', _p), ('tb', _p), ('t', _p), ('part_tj', _p), ('part_tb', _p),
        ('kappa', _p), ('supp', _p), ('scratch', _p),
        ('part_chunk', _p), ('keep_t', _i32), ('sig_a', _p), ('sig_dir', _p),
    ]


class Gcr(C.Structure):
    """Mirror of obte_gcr_t."""
    _fields_ = [
        ('kmax', _i32), ('rtol', _f64), ('y', _p), ('r', _p), ('x', _p), ('u', _p), ('c', _p),
        ('k', _p), ('active', _p), ('lane_active', _p), ('w', _p), ('n_active', _p), ('rnorm', _p),
    ]


class SweepIO(C.Structure):
    """Mirror of obte_sweep_io_t."""
    _fields_ = [('y', _p), ('x', _p), ('active', _p), ('lane_active', _p)]


class ReduceIO(C.Structure):
    """Mirror of obte_reduce_io_t."""
    _fields_ = [('dt_new', _p), ('j_new', _p), ('tb_new', _p), ('norm2', _p)]


class Batch(C.Structure):
    """Mirror of obte_bte_batch_t."""
    _fields_ = [
        ('n_problems', _i32), ('n_tasks', _i32), ('mc', _i32), ('block', _i32), ('kin', _i32), ('groups', _i32),
        ('max_elems', _i32), ('max_dirs', _i32), ('max_cells', _i32), ('max_walls', _i32),
        ('lvl_pair_max', _i32), ('sweep_window', _i32), ('part_sums', _i32),
        ('plans', _p), ('gcrs', _p), ('io', _p), ('rio', _p), ('live', _p), ('task_pid', _p), ('task_local', _p),
        ('queue_off', _p), ('queue_cnt', _p), ('wave_off', _p), ('wave_cnt', _p), ('wave_pair_max', _i32),
    ]


class FourierPlan(C.Structure):
    """Mirror of obte_fourier_plan_t."""
    _fields_ = [
        ('n', _i32), ('dim', _i32), ('nv', _i32), ('n_int', _i32), ('pin', _i32),
        ('row_ptr', _p), ('col', _p), ('w', _p), ('diag', _p), ('vol', _p),
        ('slot_cell', _p), ('slot_mode', _p), ('slot_val', _p), ('gweights', _p),
        ('side_e', _p), ('side_w', _p), ('side_vn', _p), ('slot_side', _p), ('slot_sgn', _p),
        ('n_contact', _i32), ('contact_kind', _i32), ('ce1', _p), ('ce2', _p), ('c_coef', _p),
        ('c_value', _f64), ('normalization', _f64),
        ('n_levels', _i32), ('levels', _p), ('n_coarse', _i32), ('coarse_inv', _p), ('coarse_b', _p),
        ('coarse_x', _p),
    ]


class AmgLevel(C.Structure):
    """Mirror of obte_amg_level_t."""
    _fields_ = [
        ('n', _i32), ('a_entries', _i64), ('r_entries', _i64), ('a_ptr', _p), ('a_col', _p), ('a_w', _p),
        ('a_diag', _p),
        ('p_ptr', _p), ('p_col', _p), ('p_val', _p), ('r_ptr', _p), ('r_col', _p), ('r_val', _p),
        ('b', _p), ('x', _p), ('res', _p), ('z', _p),
    ]


# name -> (argtypes)
_SIGNATURES = {
    'obte_last_error': [],
    'obte_sched_create': [_i32, _i64, _i32, _p, _p, _p, _p, _i32, _p, _i32, C.POINTER(_p)],
    'obte_sched_sizes': [_p, _p],
    'obte_sched_export': [_p, _p, _p, _p, _p, _p, _p, _p, _p],
    'obte_sched_export_slots': [_p, _i32, _p, _p],
    'obte_sched_destroy': [_p],
    'obte_set_host_threads': [_i32],
    'obte_amg_create': [_i32, _p, _p, _p, _p, _i32, _i32, _f64, _i32, C.POINTER(_p)],
    'obte_amg_sizes': [_p, C.POINTER(_i32), _p],
    'obte_amg_export': [_p, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    'obte_amg_export_dense': [_p, _p],
    'obte_amg_destroy': [_p],
    'obte_bte_assemble': [C.POINTER(BtePlan), _p],
    'obte_bte_source': [C.POINTER(BtePlan), _p],
    'obte_bte_sweep': [C.POINTER(BtePlan), _i32, _p, _p, _p, _p, _p],
    'obte_gcr_project': [C.POINTER(BtePlan), C.POINTER(Gcr), _p],
    'obte_gcr_begin_extend': [C.POINTER(BtePlan), C.POINTER(Gcr), _p],
    'obte_gcr_end_extend': [C.POINTER(BtePlan), C.POINTER(Gcr), _p],
    'obte_gcr_extend': [C.POINTER(BtePlan), C.POINTER(Gcr), _i32, _p, _p],
    'obte_gcr_reset': [C.POINTER(BtePlan), C.POINTER(Gcr), _i32, _p],
    'obte_sweep_window': [_i32, _i32, _i32],
    'obte_batch_assemble': [C.POINTER(Batch), _p],
    'obte_batch_source': [C.POINTER(Batch), _p],
    'obte_batch_sweep': [C.POINTER(Batch), _i32, _p],
    'obte_batch_gcr': [C.POINTER(Batch), _i32, _p],
    'obte_batch_gcr_reset': [C.POINTER(Batch), _i32, _p],
    'obte_batch_extend': [C.POINTER(Batch), _i32, _p, _p],
    'obte_batch_reduce': [C.POINTER(Batch), _p],
    'obte_batch_log_active': [C.POINTER(Batch), _p, _p],
    'obte_tim_face_grad': [C.POINTER(BtePlan), _p, _p, _p, _p, _p, _p, _p, _p, _p, C.c_int32, _p, _p],
    'obte_tim_face_grad_dirs': [C.POINTER(BtePlan), _p, _i32, _i32, _i32, _p, _p, _p, _p, _p, _p, _p, _p,
                                C.c_int32, _p, _p],
    'obte_rta_scatter': [C.c_int32, C.c_int32, _p, _p, _p, C.c_int32, _p, C.c_int32, _p, C.c_int32, _p, _p, _p, _p],
    'obte_bte_reduce': [C.POINTER(BtePlan), _p, _p, _p, _p, _p],
    'obte_fourier_solve': [C.POINTER(FourierPlan), _p, _p, _p, _p, _p, _f64, _i32, _i32,
                           C.POINTER(_i32), C.POINTER(_f64), _p],
    'obte_fourier_solve_noalloc': [C.POINTER(FourierPlan), _p, _p, _p, _f64, _i32, _i32,
                                   C.POINTER(_i32), C.POINTER(_f64), _p],
    'obte_fourier_gradient': [C.POINTER(FourierPlan), _p, _p, _p],
    'obte_fourier_correction': [C.POINTER(FourierPlan), _p, _p, _p, _p, _p, _p],
    'obte_fourier_kappa': [C.POINTER(FourierPlan), _p, _p, _p],
}

EXPORTED = tuple(_SIGNATURES)

_lib = None


def load() -> C.CDLL:
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f'{LIB_PATH} is missing: run `python -m openbte_amd.build` (no CPU fallback exists)')
    # load torch's HIP runtime first so that the library binds to the same one
    import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == 'obte_last_error' else C.c_int
    _lib = lib
    return lib


class ObteError(RuntimeError):
    pass


def call(name: str, *args) -> None:
    """Call a C-ABI function and raise ObteError on a negative return code."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.obte_last_error().decode(errors='replace')
        raise ObteError(f'{name} failed ({rc}): {msg}')


def ptr(t) -> int:
    """Device/host address of a torch tensor or numpy array (None -> 0)."""
    if t is None:
        return 0
    if hasattr(t, 'data_ptr'):
        return t.data_ptr()
    return t.ctypes.data
