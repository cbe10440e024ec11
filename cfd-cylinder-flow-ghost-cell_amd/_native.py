"""ctypes binding of libghostflow.so (the HIP/gfx950 C ABI declared in include/ghostflow.h).

The library is built in-tree by ``build.py``.  There is no fallback: if the
shared object is missing or the HIP runtime cannot be loaded, importing the
binding raises.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / 'libghostflow.so'

STATUS_EXCEPTIONS = {1: ValueError, 2: RuntimeError, 3: FloatingPointError, 4: RuntimeError}

# keep in sync with include/ghostflow.h
PRESSURE, U, V = 0, 1, 2
FIELD_U, FIELD_V, FIELD_P = 0, 1, 2
OP_GRAD_U, OP_GRAD_V, OP_DIV, OP_HELM_U, OP_HELM_V, OP_PRESSURE = range(6)
APPLY_ADVECTION, APPLY_DIFFUSION, APPLY_GRADIENT, APPLY_DIVERGENCE, APPLY_HMATVEC = range(5)
DIAG_NAMES = ['time', 'dt', 'cd', 'cl', 'cd_pressure', 'cd_viscous', 'cl_pressure', 'cl_viscous', 'fx', 'fy',
              'max_divergence', 'cfl', 'momentum_residual_linf', 'coupled_momentum_residual_linf',
              'pressure_residual_linf', 'pressure_residual_first_linf', 'pressure_residual_relative_l2',
              'unused_rel1', 'pressure_refinement_iterations', 'pressure_residual_target_linf', 'unused_t1',
              'pressure_roundoff_bound_linf', 'unused_r1', 'pressure_backward_error_linf', 'unused_b1',
              'helmholtz_refinement_iterations', 'status', 'rhs_max', 'pressure_solve_count', 'probe0', 'probe1']
DIAG_SOLVE0 = 32                     # per-solve records, MCF_SOLVE_LEN doubles each, in call order
SOLVE_NAMES = ['iterations', 'residual_linf', 'residual_relative_l2', 'refinement_iterations',
               'residual_target_linf', 'roundoff_bound_linf', 'backward_error_linf', 'status', 'seconds']
SOLVE_LEN = 10
MAX_SOLVES = 3
DIAG_LEN = 64
PB_DIRECT, PB_GMRES = 0, 1

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_U8 = ctypes.POINTER(ctypes.c_ubyte)
_H = ctypes.c_void_p

_SIGNATURES = {
    'mcf_create': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.POINTER(_H)]),
    'mcf_destroy': (None, [_H]),
    'mcf_last_error': (ctypes.c_char_p, [_H]),
    'mcf_device_name': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    'mcf_set_inlet': (ctypes.c_int, [_H, _D]),
    'mcf_set_inlet_perturbation': (ctypes.c_int, [_H, _D, _D, ctypes.c_double]),
    'mcf_set_solid_intervals': (ctypes.c_int, [_H, ctypes.c_int, _I, _I]),
    'mcf_set_ghost': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _I, _I, _I, _D]),
    'mcf_set_override': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _I, _I, _I, _D]),
    'mcf_set_rectangle': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, _D, _D, _D, _D, _D]),
    'mcf_set_schur': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int, _I, _D]),
    'mcf_set_surface': (ctypes.c_int, [_H, ctypes.c_int, _I, _I, _D, _I, _I, _D, _I, _I, _D, _D, _D,
                                       ctypes.c_double, ctypes.c_double]),
    'mcf_set_probes': (ctypes.c_int, [_H, ctypes.c_int, _I, _I, _D]),
    'mcf_set_state': (ctypes.c_int, [_H, _D, _D, _D, _D, ctypes.c_double, ctypes.c_double]),
    'mcf_get_state': (ctypes.c_int, [_H, _D, _D, _D, _D, _D, _D, _I]),
    'mcf_step': (ctypes.c_int, [_H, ctypes.c_double, _D]),
    'mcf_run': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_double, _D]),
    'mcf_project': (ctypes.c_int, [_H, _D, ctypes.c_double, _D, _D, _D]),
    'mcf_extend': (ctypes.c_int, [_H, _D, _D]),
    'mcf_forces': (ctypes.c_int, [_H, _D, _D, ctypes.c_double, _D]),
    'mcf_rect_inverse': (ctypes.c_int, [_H, ctypes.c_int, _D, _D]),
    'mcf_rect_inverse_noalloc': (ctypes.c_int, [_H, ctypes.c_int, _D, _D]),
    'mcf_schur_solve': (ctypes.c_int, [_H, ctypes.c_int, _D, _D, _D]),
    'mcf_set_pressure_backend': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                ctypes.c_double]),
    'mcf_get_pressure_backend': (ctypes.c_int, [_H, _I]),
    'mcf_pressure_solve': (ctypes.c_int, [_H, _D, _D, _D]),
    'mcf_apply_operator': (ctypes.c_int, [_H, ctypes.c_int, _D, _D, ctypes.c_double]),
    'mcf_divergence_stats': (ctypes.c_int, [_H, _D, _U8, _D]),
    'mcf_profile': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_double, _D]),
    'mcf_set_schur_system': (ctypes.c_int, [_H, ctypes.c_int, _D, _I, _I, _D, _I]),
    'mcf_set_dump': (ctypes.c_int, [_H, ctypes.c_int]),
    'mcf_get_dump': (ctypes.c_int, [_H, ctypes.c_int, _D, _D, _I, _I]),
    'mcf_schur_gmres': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int, _D, _I, _D]),
    'mcf_net_create': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, ctypes.c_ulonglong,
                                      ctypes.POINTER(_H)]),
    'mcf_net_destroy': (None, [_H]),
    'mcf_net_last_error': (ctypes.c_char_p, [_H]),
    'mcf_net_num_params': (ctypes.c_int, [_H]),
    'mcf_net_get_params': (ctypes.c_int, [_H, _F]),
    'mcf_net_set_params': (ctypes.c_int, [_H, _F]),
    'mcf_net_clear_groups': (ctypes.c_int, [_H]),
    'mcf_net_add_group': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _F, _F, _F, _F]),
    'mcf_net_train': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_ulonglong, _F]),
    'mcf_net_predict': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _F, _F]),
    'mcf_net_loss_grad': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _F, _F, _F, _F, _F]),
    'mcf_net_profile': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, ctypes.c_int, _F]),
    'mcf_set_launch': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int]),
    'mcf_get_launch': (ctypes.c_int, [_H, _I]),
    'mcf_set_launch_classes': (ctypes.c_int, [_H, _I]),
    'mcf_set_speculation': (ctypes.c_int, [_H, ctypes.c_int]),
    'mcf_get_speculation': (ctypes.c_int, [_H, _I]),
    'mcf_batch_set_speculation': (ctypes.c_int, [_H, ctypes.c_int]),
    'mcf_batch_get_speculation': (ctypes.c_int, [_H, _I]),
    'mcf_batch_set_launch': (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int]),
    'mcf_batch_get_launch': (ctypes.c_int, [_H, _I]),
    'mcf_batch_create': (ctypes.c_int, [ctypes.POINTER(_H), _D, ctypes.c_int, ctypes.POINTER(_H)]),
    'mcf_batch_destroy': (None, [_H]),
    'mcf_batch_last_error': (ctypes.c_char_p, [_H]),
    'mcf_batch_run': (ctypes.c_int, [_H, ctypes.c_int, _D]),
}

_lib = None


def library():
    """Load libghostflow.so once; raise if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f'{LIB_PATH} is not built; run build.py (hipcc --offload-arch=gfx950)')
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def as_arg(a):
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError('array arguments must be C-contiguous')
        if a.dtype == np.float64:
            return a.ctypes.data_as(_D)
        if a.dtype == np.float32:
            return a.ctypes.data_as(_F)
        if a.dtype == np.int32:
            return a.ctypes.data_as(_I)
        if a.dtype == np.uint8:
            return a.ctypes.data_as(_U8)
        raise TypeError(f'unsupported array dtype {a.dtype}')
    return a


class Recorder:
    """Stands in for a Handle while a case's tables are prepared away from the device: records the
    table uploads (entry point name and array arguments) for ``replay`` on a real handle."""

    def __init__(self):
        self.calls = []

    def call(self, name, *args):
        self.calls.append((name, tuple(np.array(a, copy=True) if isinstance(a, np.ndarray) else a for a in args)))

    def replay(self, handle):
        for name, args in self.calls:
            handle.call(name, *args)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Handle:
    """Owner of one mcf_handle; methods raise the reference's exception types."""

    def __init__(self, nx, ny, dx, dy, nu, rho):
        self.lib = library()
        h = _H()
        rc = self.lib.mcf_create(nx, ny, dx, dy, nu, rho, ctypes.byref(h))
        if rc:
            msg = self.lib.mcf_last_error(h).decode() if h.value else 'mcf_create failed'
            if h.value:
                self.lib.mcf_destroy(h)
            raise STATUS_EXCEPTIONS.get(rc, RuntimeError)(msg)
        self.h = h

    def check(self, rc):
        if rc:
            raise STATUS_EXCEPTIONS.get(rc, RuntimeError)(self.lib.mcf_last_error(self.h).decode())

    def call(self, name, *args):
        """Call an entry point; float64 / int32 / uint8 numpy arrays are passed as pointers."""
        self.check(getattr(self.lib, name)(self.h, *[as_arg(a) for a in args]))

    def close(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            self.lib.mcf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_name():
    buf = ctypes.create_string_buffer(256)
    rc = library().mcf_device_name(buf, 256)
    return buf.value.decode() if rc == 0 else None
