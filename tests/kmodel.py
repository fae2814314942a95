"""CPU model of the libzmpo grouped block GEMM, used to check plan lowering.

Executes a Program's task/pair arrays with numpy exactly as the HIP kernel
defines them (include/zmpo.h).  Test helper only; the product never calls it.
"""
import contextlib

import numpy as np


def _np(x):
    return x.numpy() if hasattr(x, 'numpy') else x


def emulate(prog, A, B, C, dtype=0, beta=0, stream=None):
    A, B, C = _np(A), _np(B), _np(C)
    if isinstance(prog, dict):
        from zmpo_dmrg_amd.qtensor import Program
        prog = Program(prog['tasks'], prog['pairs'])
    t, p = prog.tasks_host, prog.pairs_host
    for ti in range(len(t)):
        M, N = int(t['M'][ti]), int(t['N'][ti])
        m = np.arange(M)
        n = np.arange(N)
        cidx = (t['c_off'][ti] + (m // t['c_mlo'][ti]) * t['c_smhi'][ti]
                + (m % t['c_mlo'][ti]) * t['c_smlo'][ti])[:, None] + n[None, :] * t['c_sn'][ti]
        acc = np.zeros((M, N))
        for pi in range(int(t['p_begin'][ti]), int(t['p_end'][ti])):
            q = p[pi]
            k = np.arange(int(q['K']))
            aidx = (q['a_off'] + (m // q['a_mlo']) * q['a_smhi']
                    + (m % q['a_mlo']) * q['a_smlo'])[:, None] + k[None, :] * q['a_sk']
            bidx = q['b_off'] + k[:, None] * q['b_sk'] + (
                (n // q['b_nlo']) * q['b_snhi'] + (n % q['b_nlo']) * q['b_snlo'])[None, :]
            acc += q['scale'] * (A[aidx] @ B[bidx])
        C[cidx] = acc if beta == 0 else C[cidx] + acc


def _gather(dst, src, idx, *a):
    d, s_, i = _np(dst), _np(src), _np(idx)
    d[:] = np.where(i >= 0, s_[np.maximum(i, 0)], 0.0)
    return dst


class _CpuVecOps:
    """numpy stand-in for pdvdson._VecOps (same fixed-order semantics)."""

    def __init__(self, n):
        self.n = n

    def dots(self, X, nx, y):
        return np.array([float(_np(X[j]) @ _np(y)) for j in range(nx)])

    def lincomb(self, out, X, coef, alpha=1.0, beta=0.0):
        o = _np(out)
        acc = np.zeros(self.n)
        for j, c in enumerate(coef):
            acc += c * _np(X[j])
        o[:] = alpha * acc if beta == 0.0 else beta * o + alpha * acc
        return out

    def norm(self, x):
        return float(np.linalg.norm(_np(x)))

    def precond(self, r, diag, e, shift):
        rr = _np(r)
        rr[:] = rr / (np.abs(_np(diag) - e) + shift)


@contextlib.contextmanager
def host_model():
    """Run the zmpo_dmrg_amd engine on numpy buffers with the CPU kernel model."""
    import torch
    import zmpo_dmrg_amd.device as dv
    import zmpo_dmrg_amd.kernelQt as kq
    import zmpo_dmrg_amd.pdvdson as pd
    import zmpo_dmrg_amd.qtensor as qq
    saved = {}

    def patch(mod, name, fn):
        saved[(mod, name)] = getattr(mod, name)
        setattr(mod, name, fn)

    def upload(a, code=None):
        a = np.asarray(a)
        return torch.from_numpy(np.array(a, dtype=np.float64 if a.dtype.kind == 'f' else a.dtype))

    patch(dv, 'empty', lambda n, code=0: torch.zeros(int(max(n, 1)), dtype=torch.float64))
    patch(dv, 'zeros', lambda n, code=0: torch.zeros(int(max(n, 1)), dtype=torch.float64))
    patch(dv, 'upload', upload)
    patch(dv, 'download', lambda x: np.array(_np(x)))
    patch(dv, 'device', lambda: torch.device('cpu'))
    patch(dv, 'require', lambda: None)
    patch(pd, '_VecOps', _CpuVecOps)
    patch(qq, 'run_program', emulate)
    patch(kq, 'run_program', emulate)
    patch(qq, 'gather', _gather)
    patch(kq, 'gather', _gather)
    try:
        yield
    finally:
        for (mod, name), fn in saved.items():
            setattr(mod, name, fn)
