"""ctypes binding of libzmpo (include/zmpo.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``python -m zmpo_dmrg_amd.build``).  Nothing here falls back to a CPU
implementation: a missing library or a failing call raises ZmpoError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ZMPO_LIB: load another build of the same sources (e.g. with other
# compile-time tile constants) instead of the in-tree library
LIB_PATH = os.environ.get('ZMPO_LIB') or os.path.join(_HERE, 'libzmpo.so')

F64 = 0
F32 = 1
ZMPO_BF16 = 2        # BF16 operands and outputs, FP32 accumulation
ZMPO_BF16_F32 = 3    # BF16 operands, FP32 outputs and split-K workspace
ZMPO_ERR_ARG = -1
ZMPO_ERR_HIP = -2

TASK_DTYPE = np.dtype([('c_off', '<i8'), ('tile_base', '<i8'), ('M', '<i4'), ('N', '<i4'),
                       ('c_mlo', '<i4'), ('c_smhi', '<i4'), ('c_smlo', '<i4'), ('c_sn', '<i4'),
                       ('p_begin', '<i4'), ('p_end', '<i4'), ('tiles_n', '<i4'), ('pad', '<i4')])
PAIR_DTYPE = np.dtype([('a_off', '<i8'), ('b_off', '<i8'), ('scale', '<f8'), ('K', '<i4'),
                       ('a_mlo', '<i4'), ('a_smhi', '<i4'), ('a_smlo', '<i4'), ('a_sk', '<i4'),
                       ('b_nlo', '<i4'), ('b_snhi', '<i4'), ('b_snlo', '<i4'), ('b_sk', '<i4'),
                       ('pad', '<i4')])
UNIT_DTYPE = np.dtype([('task', '<i4'), ('m0', '<i4'), ('n0', '<i4'), ('p_begin', '<i4'),
                       ('p_end', '<i4'), ('slot', '<i4'), ('pad0', '<i4'), ('pad1', '<i4')])
RGROUP_DTYPE = np.dtype([('task', '<i4'), ('m0', '<i4'), ('n0', '<i4'), ('slot_begin', '<i4'),
                         ('nslots', '<i4'), ('pad', '<i4')])
CGROUP_DTYPE = np.dtype([('M', '<i4'), ('ncols', '<i4'), ('kcols', '<i4'), ('t1_begin', '<i4'),
                         ('t1_end', '<i4'), ('t2_begin', '<i4'), ('t2_end', '<i4'),
                         ('ab_begin', '<i4'), ('ab_end', '<i4'), ('rsplit', '<i4')])
CBLOCK_DTYPE = np.dtype([('a_off', '<i8'), ('a_smhi', '<i8'), ('a_smlo', '<i8'), ('a_sk', '<i8'),
                         ('K', '<i4'), ('col', '<i4'), ('a_mlo', '<i4'), ('pad', '<i4')])
CUNIT_DTYPE = np.dtype([('a_base', '<i8'), ('b1_base', '<i8'), ('b2_base', '<i8'),
                        ('c_base', '<i8'), ('group', '<i4'), ('m0', '<i4')])
SUNIT_DTYPE = np.dtype([('c_off', '<i8'), ('c_smhi', '<i8'), ('c_mlo', '<i4'), ('c_smlo', '<i4'),
                        ('c_sn', '<i4'), ('m0', '<i4'), ('n0', '<i4'), ('M', '<i4'), ('N', '<i4'),
                        ('p_begin', '<i4'), ('p_end', '<i4'), ('slot', '<i4'), ('pad0', '<i4'),
                        ('pad1', '<i4')])
SEQ_OP_DTYPE = np.dtype([('kind', '<i4'), ('shape', '<i4'), ('kt', '<i4'), ('upw', '<i4'),
                         ('rb', '<i4'), ('atile', '<i4'), ('kmax', '<i4'), ('nmax', '<i4'),
                         ('tasks', '<u8'), ('units', '<u8'), ('nunits', '<i8'), ('groups', '<u8'),
                         ('ngroups', '<i8'), ('pairs', '<u8'), ('blocks', '<u8'),
                         ('slot_a', '<i4'), ('slot_b', '<i4'), ('slot_b2', '<i4'),
                         ('slot_c', '<i4'), ('slot_work', '<i4'), ('beta', '<i4'),
                         ('out_f32', '<i4'), ('pad', '<i4')])
SEQ_SHAPED, SEQ_STREAM, SEQ_CHAIN, SEQ_CHAIN16, SEQ_GATHER = 0, 1, 2, 3, 4
H2D, D2H, D2D = 0, 1, 2
TILE = 64

EXPORTS = ['zmpo_last_error', 'zmpo_abi_version', 'zmpo_struct_sizes', 'zmpo_symAllowed',
           'zmpo_symAllowedInt', 'zmpo_blockgemm', 'zmpo_wavegemm', 'zmpo_shapedgemm', 'zmpo_streamgemm', 'zmpo_chaingemm', 'zmpo_chaingemm16', 'zmpo_skinnygemm', 'zmpo_dotgemm', 'zmpo_dots', 'zmpo_dots_workspace',
           'zmpo_lincomb', 'zmpo_precond', 'zmpo_gather', 'zmpo_scatter', 'zmpo_convert',
           'zmpo_scratch_bytes', 'zmpo_streamgemm_static', 'zmpo_alloc', 'zmpo_free',
           'zmpo_memcpy', 'zmpo_syevd_sectors', 'zmpo_schedule_build', 'zmpo_seq_create', 'zmpo_seq_destroy', 'zmpo_seq_add',
           'zmpo_seq_length', 'zmpo_seq_run', 'zmpo_sigma_two_site', 'zmpo_renorm',
           'zmpo_hdiag', 'zmpo_davidson_residual']


class ZmpoError(RuntimeError):
    pass


_lib = None


def lib():
    """Load (once) and return the ctypes handle; raises if the library is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ZmpoError('libzmpo.so not built at %s: run __graft_entry__.build()' % LIB_PATH)
    # Bring up torch's HIP runtime before libzmpo pulls in the system one:
    # the package shares device memory and streams with torch, and a process
    # that loaded libzmpo first has seen its kernel launches report no device.
    from . import device as _dv
    _dv.torch().cuda.is_available()
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    L.zmpo_last_error.restype = ctypes.c_char_p
    L.zmpo_abi_version.restype = i32
    L.zmpo_struct_sizes.argtypes = [ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.zmpo_symAllowed.argtypes = [i32, i32, vp, vp, vp] + [vp] * 8
    L.zmpo_symAllowedInt.argtypes = [i32, i32, vp, vp, vp, vp]
    L.zmpo_blockgemm.argtypes = [i32, vp, i64, vp, i64, vp, vp, vp, i32, vp]
    L.zmpo_wavegemm.argtypes = [i32, i32, vp, vp, i64, vp, i64, vp, vp, vp, vp, vp, i32, vp]
    L.zmpo_shapedgemm.argtypes = [i32, i32, vp, vp, i64, vp, i64, vp, vp, vp, vp, vp, i32, vp]
    L.zmpo_dotgemm.argtypes = [i32, vp, vp, i64, vp, i64, vp, vp, vp, vp, vp, i32, vp]
    L.zmpo_skinnygemm.argtypes = [i32, i32, vp, vp, i64, vp, vp, vp, vp, i32, vp]
    L.zmpo_chaingemm.argtypes = [i32, i32, i32, i32, i32, vp, i64, vp, vp, vp, vp, vp, vp, vp,
                                 vp, i32, vp]
    L.zmpo_chaingemm16.argtypes = L.zmpo_chaingemm.argtypes
    L.zmpo_streamgemm.argtypes = [i32, i32, i32, vp, i64, i32, vp, vp, i64, vp, vp, vp, vp, vp,
                                  i32, vp]
    L.zmpo_streamgemm_static.argtypes = [i32, i32, i32, vp, i64, i32, vp, vp, i64, vp, vp, vp,
                                         vp, i64, i32, vp]
    L.zmpo_scratch_bytes.restype = i64
    L.zmpo_alloc.argtypes = [ctypes.POINTER(vp), i64]
    L.zmpo_free.argtypes = [vp]
    L.zmpo_memcpy.argtypes = [vp, vp, i64, i32, vp]
    L.zmpo_syevd_sectors.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp]
    L.zmpo_sigma_two_site.argtypes = [vp, i32] + [vp] * 10
    L.zmpo_renorm.argtypes = [vp, i32] + [vp] * 8
    L.zmpo_hdiag.argtypes = [vp] * 10
    L.zmpo_davidson_residual.argtypes = [i64, vp, vp, i64, i32, vp, ctypes.c_double, vp,
                                         ctypes.c_double, vp, vp]
    L.zmpo_schedule_build.argtypes = [i64, vp, vp, i32, i32, i64, vp, vp, i64, vp, i64, vp]
    L.zmpo_seq_create.argtypes = [ctypes.POINTER(vp)]
    L.zmpo_seq_destroy.argtypes = [vp]
    L.zmpo_seq_add.argtypes = [vp, vp]
    L.zmpo_seq_length.argtypes = [vp]
    L.zmpo_seq_length.restype = i64
    L.zmpo_seq_run.argtypes = [vp, i32, vp, i32, vp]
    L.zmpo_dots.argtypes = [i32, vp, i64, i32, vp, i64, vp, vp, vp]
    L.zmpo_dots_workspace.argtypes = [i32]
    L.zmpo_dots_workspace.restype = i32
    L.zmpo_lincomb.argtypes = [i32, vp, vp, i64, i32, vp, dbl, dbl, i64, vp]
    L.zmpo_precond.argtypes = [i32, vp, vp, dbl, dbl, i64, vp]
    L.zmpo_gather.argtypes = [i32, vp, i32, vp, vp, i64, vp]
    L.zmpo_scatter.argtypes = [i32, vp, i32, vp, vp, i64, vp]
    L.zmpo_convert.argtypes = [i32, vp, i32, vp, i64, vp]
    ts, ps = ctypes.c_int(), ctypes.c_int()
    L.zmpo_struct_sizes(ctypes.byref(ts), ctypes.byref(ps))
    if ts.value != TASK_DTYPE.itemsize or ps.value != PAIR_DTYPE.itemsize:
        raise ZmpoError('ABI struct size mismatch: %d/%d vs %d/%d' % (
            ts.value, ps.value, TASK_DTYPE.itemsize, PAIR_DTYPE.itemsize))
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise ZmpoError(lib().zmpo_last_error().decode())
    return rc


def sym_allowed(nsyms, status, qsyms):
    """Allowed-block mask (int32, C order) through zmpo_symAllowed (qsym.c:5-22)."""
    rank = len(nsyms)
    if rank == 1:
        q = np.asarray(qsyms[0], dtype=np.float64)
        sgn = -1.0 if status[0] else 1.0
        return (np.abs(sgn * q).sum(axis=1) < 1e-10).astype(np.int32)
    if rank > 8:
        raise ZmpoError('block rule supports rank <= 8')
    nq = int(np.asarray(qsyms[0]).shape[1])
    nd = np.ascontiguousarray(nsyms, dtype=np.int32)
    st = np.ascontiguousarray([1 if s else 0 for s in status], dtype=np.int32)
    out = np.zeros(int(np.prod(nd)), dtype=np.int32)
    qs = [np.ascontiguousarray(q, dtype=np.float64) for q in qsyms]
    ptrs = [q.ctypes.data for q in qs] + [None] * (8 - rank)
    check(lib().zmpo_symAllowed(rank, nq, nd.ctypes.data, st.ctypes.data, out.ctypes.data, *ptrs))
    return out


class Seq:
    """A launch sequence (zmpo_seq_*): ops recorded once, run with operand
    slots bound to device pointers.  The recorded tables must outlive it."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().zmpo_seq_create(ctypes.byref(h)))
        self._h = h
        self._lib = lib()
        self.nops = 0

    def add(self, rec):
        r = np.ascontiguousarray(np.asarray(rec, dtype=SEQ_OP_DTYPE).reshape(1))
        check(self._lib.zmpo_seq_add(self._h, r.ctypes.data))
        self.nops += 1

    def run(self, dtype, slots, stream):
        arr = (ctypes.c_void_p * len(slots))(*[int(x) if x else None for x in slots])
        check(self._lib.zmpo_seq_run(self._h, dtype, arr, len(slots), stream))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            self._lib.zmpo_seq_destroy(h)
            self._h = None
