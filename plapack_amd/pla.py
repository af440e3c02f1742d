"""ctypes binding of libpla_mi355x.so, the PLAPACK C API on a 1x1 template.

Every entry point of include/PLA.h is bound with its C signature; handles are
opaque pointers, MPI_Comm / MPI_Datatype are MPICH ABI ints
(include/PLA_mpi_abi.h).  The helpers below (Session, from_numpy, to_numpy)
only move host arrays in and out of device-resident PLA_Obj objects; all
arithmetic happens in the library's HIP kernels.  There is no fallback: if
the library is missing or no GPU is visible, calls fail with an error.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpla_mi355x.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "PLA.h")

# MPICH ABI handle values (include/PLA_mpi_abi.h)
MPI_COMM_WORLD = 0x44000000
MPI_COMM_SELF = 0x44000001
MPI_INT = 0x4C000405
MPI_FLOAT = 0x4C00040A
MPI_DOUBLE = 0x4C00080B
MPI_DOUBLE_COMPLEX = 1275072546
MPI_COMPLEX = 1275070494

# constants (INCLUDE/PLA_define.h:11-102)
PLA_MVECTOR = 12
PLA_MATRIX = 13
PLA_MSCALAR = 14
PLA_NO_TRANSPOSE = 21
PLA_TRANSPOSE = 22
PLA_CONJUGATE_TRANSPOSE = 23
PLA_ALL_ROWS = -31
PLA_ALL_COLS = -32
PLA_SIDE_LEFT = 43
PLA_SIDE_RIGHT = 44
PLA_UPPER_TRIANGULAR = 51
PLA_LOWER_TRIANGULAR = 52
PLA_OP_PAN_PAN = 83
PLA_OP_ALL_ALG = 87
PLA_UNIT_DIAG = 101
PLA_NONUNIT_DIAG = 102
PLA_SUCCESS = 0
# error codes (negative; PLA_Last_error() holds the message)
PLA_ERR_NOT_INITIALIZED = -1001
PLA_ERR_INVALID_OBJECT = -1002
PLA_ERR_INVALID_PARAM = -1003
PLA_ERR_UNSUPPORTED = -1004
PLA_ERR_DIMENSIONS = -1005
PLA_ERR_DEVICE = -1006
PLA_ERR_NO_MEMORY = -1007
PLA_ERR_MULTI_PROCESS = -1008
PLA_ERR_NOT_CONVERGED = -1009
PLA_MODE_CLOSED = 71
PLA_MODE_OPEN = 72
PLA_STATE_API_ACTIVE = 73
PLA_STATE_NORMAL = 74
PLA_ALIGN_FIRST = -200
PLA_MIXED_FP16 = 1
PLA_MIXED_BF16 = 2
PLA_MIXED_FP16X3 = 3
PLA_MIXED_BF16X3 = 4
PLA_MIXED_FP8 = 5
PLA_PARAM_LOOKAHEAD = 1
PLA_PARAM_LEAF_ROWS = 2
PLA_PARAM_PANEL_CUS = 3
PLA_PARAM_LEAF_WIDTH = 4
PLA_PARAM_LEAF_KIND = 5
PLA_PARAM_NB_TAIL = 6
PLA_PARAM_NB_SWITCH = 7
PLA_PARAM_CHOL_PANEL_CUS = 8
PLA_NB_LU = 0
PLA_NB_CHOL = 1
PLA_NB_LU_MIXED = 2
PLA_BATCHED_FP64 = 1
PLA_BATCHED_FP32_IR = 2
PLA_BATCHED_AUTOTUNE = 0x100

_obj = ctypes.c_void_p
_pobj = ctypes.POINTER(ctypes.c_void_p)
_i = ctypes.c_int
_pi = ctypes.POINTER(ctypes.c_int)
_p = ctypes.c_void_p
_d = ctypes.c_double

# name -> argtypes (all return int unless listed in _RESTYPES)
_SIGS = {
    "PLA_Comm_1D_to_2D": [_i, _i, _i, _pi],
    "PLA_Comm_1D_to_2D_ratio": [_i, _d, _pi],
    "PLA_Abort": [ctypes.c_char_p, _i, ctypes.c_char_p],
    "PLA_Gemm_x": [_i, _i, _i, _obj, _obj, _obj, _obj, _obj],
    "PLA_Axpy": [_obj, _obj, _obj],
    "PLA_Nrm2": [_obj, _obj],
    "PLA_Symv": [_i, _obj, _obj, _obj, _obj, _obj],
    "PLA_Hemv": [_i, _obj, _obj, _obj, _obj, _obj],
    "PLA_Shift": [_obj, _i, _p],
    "PLA_Init": [_i],
    "PLA_Finalize": [],
    "PLA_Initialized": [_pi],
    "PLA_Temp_create": [_i, _i, _pobj],
    "PLA_Temp_free": [_pobj],
    "PLA_Temp_nb": [_obj, _pi],
    "PLA_Temp_zero_or_one": [_obj, _pi],
    "PLA_Temp_comm_all_size": [_obj, _pi],
    "PLA_Temp_comm_all": [_obj, _pi],
    "PLA_Temp_comm_all_rank": [_obj, _pi],
    "pla_Environ_set_nb_alg": [_i, _i],
    "PLA_Environ_nb_alg": [_i, _obj, _pi],
    "PLA_Set_error_checking": [_i, _i, _i, _i],
    "PLA_Get_error_checking": [_pi, _pi, _pi, _pi],
    "PLA_Matrix_create": [_i, _i, _i, _obj, _i, _i, _pobj],
    "PLA_Mvector_create": [_i, _i, _i, _obj, _i, _pobj],
    "PLA_Mscalar_create": [_i, _i, _i, _i, _i, _obj, _pobj],
    "PLA_Matrix_create_conf_to": [_obj, _pobj],
    "PLA_Mvector_create_conf_to": [_obj, _i, _pobj],
    "PLA_Mscalar_create_conf_to": [_obj, _i, _i, _pobj],
    "PLA_Create_constants_conf_to": [_obj, _pobj, _pobj, _pobj],
    "PLA_Obj_free": [_pobj],
    "PLA_Obj_objtype": [_obj, _pi],
    "PLA_Obj_datatype": [_obj, _pi],
    "PLA_Obj_template": [_obj, _pobj],
    "PLA_Obj_global_length": [_obj, _pi],
    "PLA_Obj_global_width": [_obj, _pi],
    "PLA_Obj_owner_row": [_obj, _pi],
    "PLA_Obj_owner_col": [_obj, _pi],
    "PLA_Obj_local_length": [_obj, _pi],
    "PLA_Obj_local_width": [_obj, _pi],
    "PLA_Obj_local_ldim": [_obj, _pi],
    "PLA_Obj_local_stride": [_obj, _pi],
    "PLA_Obj_local_buffer": [_obj, _pobj],
    "PLA_Obj_get_local_contents": [_obj, _i, _pi, _pi, _p, _i, _i],
    "PLA_Obj_set_local_contents": [_i, _i, _i, _p, _i, _i, _obj],
    "PLA_Obj_set_to_zero": [_obj],
    "PLA_Obj_set_to_one": [_obj],
    "PLA_Obj_set_to_minus_one": [_obj],
    "PLA_Obj_view_all": [_obj, _pobj],
    "PLA_Obj_split_4": [_obj, _i, _i, _pobj, _pobj, _pobj, _pobj],
    "PLA_Obj_horz_split_2": [_obj, _i, _pobj, _pobj],
    "PLA_Obj_vert_split_2": [_obj, _i, _pobj, _pobj],
    "PLA_Copy": [_obj, _obj],
    "PLA_LU": [_obj, _obj],
    "PLA_LU_linpack": [_obj, _obj],
    "PLA_Chol_simple": [_i, _obj],
    "PLA_Chol": [_i, _obj],
    "PLA_Chol_no_alloc": [_i, _obj],
    "PLA_Apply_pivots_to_rows": [_obj, _obj],
    "PLA_Trsv": [_i, _i, _i, _obj, _obj],
    "PLA_Trsm": [_i, _i, _i, _i, _obj, _obj, _obj],
    "PLA_Gemv": [_i, _obj, _obj, _obj, _obj, _obj],
    "PLA_Gemm": [_i, _i, _obj, _obj, _obj, _obj, _obj],
    "PLA_Syrk": [_i, _i, _obj, _obj, _obj, _obj],
    "PLA_API_begin": [],
    "PLA_API_end": [],
    "PLA_API_state": [_pi],
    "PLA_Obj_API_open": [_obj],
    "PLA_Obj_API_close": [_obj],
    "PLA_Obj_API_mode": [_obj, _pi],
    "PLA_Obj_API_sync": [_obj],
    "PLA_API_axpy_vector_to_global": [_i, _p, _p, _i, _obj, _i],
    "PLA_API_axpy_global_to_vector": [_i, _p, _obj, _i, _p, _i],
    "PLA_API_multi_axpy_vector_to_global": [_i, _p, _p, _p, _i, _obj, _p],
    "PLA_API_multi_axpy_global_to_vector": [_i, _p, _p, _obj, _p, _p, _i],
    "PLA_API_axpy_matrix_to_global": [_i, _i, _p, _p, _i, _obj, _i, _i],
    "PLA_API_axpy_global_to_matrix": [_i, _i, _p, _obj, _i, _i, _p, _i],
    "PLA_API_multi_axpy_matrix_to_global": [_i, _i, _p, _p, _p, _p, _i, _obj, _p, _p],
    "PLA_API_multi_axpy_global_to_matrix": [_i, _i, _p, _p, _p, _obj, _p, _p, _p, _i],
    "PLA_Matrix_infinity_norm": [_obj, _obj],
    "PLA_Matrix_one_norm": [_obj, _obj],
    "PLA_General_solve": [_obj, _obj, _obj],
    "PLA_Pos_def_solve": [_obj, _obj],
    "PLA_LU_mixed": [_obj, _obj, _obj, _obj, _i, _d, _i, _pi, ctypes.POINTER(_d)],
    "PLA_Chol_mixed": [_obj, _obj, _obj, _i, _d, _i, _pi, ctypes.POINTER(_d)],
    "PLA_Mixed_stats": [_pi, _pi, _pi, ctypes.POINTER(_d)],
    "PLA_Batched_plan_create": [_i, _pi, _p, _p, _i, _i, _d, _i, _p, _p, _p, _p, _p, _pobj],
    "PLA_Batched_plan_execute": [_obj],
    "PLA_Batched_plan_autotune": [_obj, ctypes.POINTER(_d)],
    "PLA_Batched_plan_config": [_obj, _pi, _pi],
    "PLA_Batched_plan_free": [_pobj],
    "PLA_Chol_batched": [_i, _pi, _p, _p, _i, _i, _d, _i, _p, _p, _p, _p, _p],
    "PLA_Last_error": [],
    "PLA_Obj_device_buffer": [_obj, _pobj],
    "PLA_Obj_attach_buffer": [_obj, _p, _i, _i],
    "PLA_Obj_fill_drand48": [_obj, ctypes.c_long, ctypes.c_int64],
    "PLA_Device_synchronize": [],
    "PLA_Set_device_nb": [_i],
    "PLA_Set_device_param": [_i, _i],
    "PLA_Get_device_nb": [_i, _pi],
    "PLA_Get_device_nb_for": [_i, _i, _pi],
    "PLA_Autotune_nb": [_i, _i, _pi, _p],
    "PLA_Kernel_timing": [_i],
    "PLA_Kernel_timing_query": [_i, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(_d),
                                ctypes.POINTER(_d)],
}
_RESTYPES = {"PLA_Last_error": ctypes.c_char_p}

_lib = None


class PLAError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} returned {code}: {msg}")
        self.code = code


def header_functions():
    """Names of the functions include/PLA.h declares."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|char\s*\*|const char\s*\*)\s*\*?\s*(\w+)\s*\(",
                                 txt, flags=re.M)))


def lib():
    """Load libpla_mi355x.so (built by `make -C plapack_amd`); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built; run `make -C plapack_amd` "
                          "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def last_error():
    return lib().PLA_Last_error().decode(errors="replace")


def call(name, *args, ok=(PLA_SUCCESS,)):
    """Call an entry point; raise PLAError on a return code outside `ok`."""
    rc = getattr(lib(), name)(*args)
    if rc not in ok:
        raise PLAError(name, rc, last_error())
    return rc


class Session:
    """PLA_Init on a 1x1 template; PLA_Finalize on exit."""

    def __init__(self, nb_distr=128, zero_or_one=0):
        self.nb_distr = nb_distr
        self.zero_or_one = zero_or_one
        self.templ = ctypes.c_void_p()

    def __enter__(self):
        L = lib()
        init = ctypes.c_int(0)
        L.PLA_Initialized(ctypes.byref(init))
        if not init.value:
            call("PLA_Init", MPI_COMM_WORLD)
        call("PLA_Temp_create", self.nb_distr, self.zero_or_one, ctypes.byref(self.templ))
        return self

    def __exit__(self, *exc):
        lib().PLA_Temp_free(ctypes.byref(self.templ))
        lib().PLA_Finalize()
        return False

    # ---- object helpers ----
    def matrix(self, m, n, dtype=MPI_DOUBLE):
        o = ctypes.c_void_p()
        call("PLA_Matrix_create", dtype, m, n, self.templ, PLA_ALIGN_FIRST, PLA_ALIGN_FIRST,
             ctypes.byref(o))
        return o

    def mvector(self, m, w=1, dtype=MPI_DOUBLE):
        o = ctypes.c_void_p()
        call("PLA_Mvector_create", dtype, m, w, self.templ, PLA_ALIGN_FIRST, ctypes.byref(o))
        return o

    def mscalar(self, m=1, n=1, dtype=MPI_DOUBLE):
        o = ctypes.c_void_p()
        call("PLA_Mscalar_create", dtype, PLA_ALL_ROWS, PLA_ALL_COLS, m, n, self.templ,
             ctypes.byref(o))
        return o

    def scalar(self, value):
        o = self.mscalar()
        set_contents(o, np.array([[value]], dtype=np.float64))
        return o


def free(*objs):
    for o in objs:
        lib().PLA_Obj_free(ctypes.byref(o))


def dims(o):
    m, n = ctypes.c_int(), ctypes.c_int()
    call("PLA_Obj_global_length", o, ctypes.byref(m))
    call("PLA_Obj_global_width", o, ctypes.byref(n))
    return m.value, n.value


def datatype(o):
    d = ctypes.c_int()
    call("PLA_Obj_datatype", o, ctypes.byref(d))
    return d.value


def _np_dtype(dt):
    return {MPI_DOUBLE: np.float64, MPI_INT: np.int32, MPI_FLOAT: np.float32,
            MPI_DOUBLE_COMPLEX: np.complex128, MPI_COMPLEX: np.complex64}[dt]


def set_contents(o, arr):
    dt = datatype(o)
    a = np.asfortranarray(np.asarray(arr, dtype=_np_dtype(dt)))
    if a.ndim == 1:
        a = a.reshape(-1, 1, order="F")
    m, n = a.shape
    call("PLA_Obj_set_local_contents", PLA_NO_TRANSPOSE, m, n,
         a.ctypes.data_as(ctypes.c_void_p), max(m, 1), 1, o)


def get_contents(o):
    dt = datatype(o)
    m, n = dims(o)
    a = np.zeros((m, n), dtype=_np_dtype(dt), order="F")
    rows, cols = ctypes.c_int(), ctypes.c_int()
    call("PLA_Obj_get_local_contents", o, PLA_NO_TRANSPOSE, ctypes.byref(rows), ctypes.byref(cols),
         a.ctypes.data_as(ctypes.c_void_p), max(m, 1), 1)
    return a


def from_numpy(session, arr):
    a = np.asarray(arr)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    dt = MPI_INT if a.dtype.kind in "iu" else (MPI_DOUBLE_COMPLEX if a.dtype.kind == "c" else MPI_DOUBLE)
    o = session.matrix(a.shape[0], a.shape[1], dt)
    set_contents(o, a)
    return o
