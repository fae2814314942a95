"""Block rule and C-ABI surface (CPU only: no device calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import zmpo_dmrg_amd._lib as L
from oracle.qtensor import blks_allowed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, 'oracle', '_ref', 'libqsym_ref.so')


def _random_qsyms(rng, rank):
    nsyms = rng.integers(1, 5, rank)
    pool = np.array([[0., 0.], [1., .5], [1., -.5], [2., 0.], [-1., -.5], [3., .5]])
    return [pool[rng.integers(0, len(pool), n)] for n in nsyms], nsyms


def test_library_exports_every_header_symbol():
    lib = L.lib()
    hdr = open(os.path.join(ROOT, 'include', 'zmpo.h')).read()
    names = set(re.findall(r'^(?:int64_t|int|const char\*)\s+(zmpo_\w+)\s*\(', hdr, re.M))
    assert names, 'no declarations parsed'
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(L.EXPORTS) == names
    assert lib.zmpo_abi_version() == 1


def test_struct_layout():
    ts, ps = ctypes.c_int(), ctypes.c_int()
    L.lib().zmpo_struct_sizes(ctypes.byref(ts), ctypes.byref(ps))
    assert (ts.value, ps.value) == (L.TASK_DTYPE.itemsize, L.PAIR_DTYPE.itemsize)


@pytest.mark.parametrize('rank', [2, 3, 4, 5, 6])
def test_sym_allowed_matches_closed_form(rank):
    rng = np.random.default_rng(rank)
    for _ in range(20):
        qs, ns = _random_qsyms(rng, rank)
        st = rng.integers(0, 2, rank).astype(bool)
        assert np.array_equal(L.sym_allowed(ns, st, qs), blks_allowed(ns, st, qs))


def _ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip('oracle/_ref/libqsym_ref.so not built (needs the reference checkout)')
    return ctypes.CDLL(REF_SO)


@pytest.mark.parametrize('rank', [2, 3, 4, 5, 6, 7])
def test_sym_allowed_matches_reference_qsym(rank):
    """Bit-identical masks against the reference's own ctypes/qsym.c."""
    ref = _ref_lib()
    rng = np.random.default_rng(100 + rank)
    for _ in range(10):
        qs, ns = _random_qsyms(rng, rank)
        st = rng.integers(0, 2, rank).astype(np.int32)
        nd = np.ascontiguousarray(ns, dtype=np.int32)
        qs = [np.ascontiguousarray(q) for q in qs]
        want = np.zeros(int(np.prod(nd)), np.int32)
        ref.symAllowed(ctypes.c_int(rank), ctypes.c_int(2), nd.ctypes.data_as(ctypes.c_void_p),
                       st.ctypes.data_as(ctypes.c_void_p), want.ctypes.data_as(ctypes.c_void_p),
                       *[q.ctypes.data_as(ctypes.c_void_p) for q in qs],
                       *[None] * (8 - rank))
        assert np.array_equal(L.sym_allowed(ns, st.astype(bool), qs), want)


def test_sym_allowed_rank8_reference_loop_bug():
    """qsym.c:436 runs the 8th index over ndims[6]; both agree when ndims[6]==ndims[7]."""
    ref = _ref_lib()
    rng = np.random.default_rng(8)
    qs, ns = _random_qsyms(rng, 8)
    ns[7] = ns[6]
    qs[7] = qs[7][:1].repeat(ns[6], axis=0)
    st = rng.integers(0, 2, 8).astype(np.int32)
    nd = np.ascontiguousarray(ns, dtype=np.int32)
    qs = [np.ascontiguousarray(q) for q in qs]
    want = np.zeros(int(np.prod(nd)), np.int32)
    ref.symAllowed(ctypes.c_int(8), ctypes.c_int(2), nd.ctypes.data_as(ctypes.c_void_p),
                   st.ctypes.data_as(ctypes.c_void_p), want.ctypes.data_as(ctypes.c_void_p),
                   *[q.ctypes.data_as(ctypes.c_void_p) for q in qs])
    assert np.array_equal(L.sym_allowed(ns, st.astype(bool), qs), want)


def test_sym_allowed_rejects_bad_rank():
    lib = L.lib()
    nd = np.ones(9, np.int32)
    rc = lib.zmpo_symAllowed(9, 2, nd.ctypes.data, nd.ctypes.data, nd.ctypes.data,
                             *[None] * 8)
    assert rc != 0 and b'rank' in lib.zmpo_last_error()


def test_sym_allowed_int_variant():
    rng = np.random.default_rng(5)
    lib = L.lib()
    for rank in (2, 3, 4):
        qs, ns = _random_qsyms(rng, rank)
        st = rng.integers(0, 2, rank).astype(np.int32)
        qi = [np.ascontiguousarray(np.stack([q[:, 0], 2 * q[:, 1]], 1).astype(np.int32)) for q in qs]
        arr = (ctypes.c_void_p * rank)(*[q.ctypes.data for q in qi])
        out = np.zeros(int(np.prod(ns)), np.int32)
        nd = np.ascontiguousarray(ns, np.int32)
        assert lib.zmpo_symAllowedInt(rank, 2, nd.ctypes.data, st.ctypes.data, out.ctypes.data,
                                      ctypes.cast(arr, ctypes.c_void_p)) == 0
        assert np.array_equal(out, blks_allowed(ns, st.astype(bool), qs))
