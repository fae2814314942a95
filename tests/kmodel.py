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


def emulate_wave(prog, A, B, C, dtype=0, beta=0, stream=None, kernel=None):
    """Model of zmpo_wavegemm: per-unit tiles over pair chunks, split tiles
    summed per group in slot order (the kernel's own schedule)."""
    from zmpo_dmrg_amd.qtensor import Program, make_schedule
    if isinstance(prog, dict):
        prog = Program(prog['tasks'], prog['pairs'])
    from zmpo_dmrg_amd.qtensor import KERNEL, SHAPES
    if kernel is None:
        kernel = getattr(prog, 'prefer', None) or KERNEL
    if kernel in ('tile', 'skinny'):
        return emulate(prog, A, B, C, dtype, beta)
    A, B, C = _np(A), _np(B), _np(C)
    rb = prog.rb
    tile = None if kernel == 'wave' else (32, 32) if kernel == 'dot' else SHAPES[prog.shape]
    TMs, TNs = (16 * rb, 16) if tile is None else tile
    units, groups, nslots = make_schedule(prog.tasks_host, prog.pairs_host, rb,
                                          reuse_key=getattr(prog, 'reuse_key', None), tile=tile)
    t, p = prog.tasks_host, prog.pairs_host
    work = np.zeros((max(nslots, 1), TMs, TNs))

    def cidx(tk, m, n):
        return (tk['c_off'] + (m // tk['c_mlo']) * tk['c_smhi']
                + (m % tk['c_mlo']) * tk['c_smlo'])[:, None] + n[None, :] * tk['c_sn']

    for u in units:
        tk = t[u['task']]
        m = u['m0'] + np.arange(TMs)
        n = u['n0'] + np.arange(TNs)
        mv, nv = m < tk['M'], n < tk['N']
        acc = np.zeros((TMs, TNs))
        for pi in range(int(u['p_begin']), int(u['p_end'])):
            q = p[pi]
            k = np.arange(int(q['K']))
            ai = q['a_off'] + (m // q['a_mlo']) * q['a_smhi'] + (m % q['a_mlo']) * q['a_smlo']
            bi = q['b_off'] + (n // q['b_nlo']) * q['b_snhi'] + (n % q['b_nlo']) * q['b_snlo']
            Am = A[np.where(mv, ai, 0)[:, None] + k[None, :] * q['a_sk']] * mv[:, None]
            Bm = B[k[:, None] * q['b_sk'] + np.where(nv, bi, 0)[None, :]] * nv[None, :]
            acc += q['scale'] * (Am @ Bm)
        if u['slot'] >= 0:
            work[u['slot']] = acc
        else:
            ci = cidx(tk, m[mv], n[nv])
            v = acc[np.ix_(mv, nv)]
            C[ci] = v if beta == 0 else C[ci] + v
    for g in groups:
        tk = t[g['task']]
        s_ = work[g['slot_begin']:g['slot_begin'] + g['nslots']].sum(0)
        m = g['m0'] + np.arange(TMs)
        n = g['n0'] + np.arange(TNs)
        mv, nv = m < tk['M'], n < tk['N']
        ci = cidx(tk, m[mv], n[nv])
        v = s_[np.ix_(mv, nv)]
        C[ci] = v if beta == 0 else C[ci] + v


def emulate_stream(prog, A, B, C, kt, upw, beta=0):
    """Model of zmpo_streamgemm: the per-workgroup cursor over (unit, pair,
    k-tile) items of the stream kernel, with its fetch/compute/epilogue order,
    on the zmpo_sunit records; split tiles summed per group in slot order."""
    from zmpo_dmrg_amd.qtensor import SHAPES, make_schedule, stream_units
    A, B, C = _np(A), _np(B), _np(C)
    TMs, TNs = SHAPES[prog.shape]
    units, groups, nslots = make_schedule(prog.tasks_host, prog.pairs_host, prog.rb,
                                          reuse_key=getattr(prog, 'reuse_key', None),
                                          tile=(TMs, TNs))
    su = stream_units(prog.tasks_host, units)
    p = prog.pairs_host
    work = np.zeros((max(nslots, 1), TMs, TNs))
    nitems = 0
    for u0 in range(0, len(su), upw):
        rng_ = su[u0:u0 + upw]
        n = len(rng_)
        st = {}

        def setup_pair():
            q = p[st['lp']]
            un = rng_[st['lu']]
            m = un['m0'] + np.arange(TMs)
            nn = un['n0'] + np.arange(TNs)
            mv, nv = m < un['M'], nn < un['N']
            st['ai'] = np.where(mv, q['a_off'] + (m // q['a_mlo']) * q['a_smhi']
                                + (m % q['a_mlo']) * q['a_smlo'], -1)
            st['bi'] = np.where(nv, q['b_off'] + (nn // q['b_nlo']) * q['b_snhi']
                                + (nn % q['b_nlo']) * q['b_snlo'], -1)
            st['q'] = q
            st['lK'] = int(q['K'])

        def begin_unit():
            un = rng_[st['lu']]
            st['lp'], st['lpe'], st['lkt'], st['lK'] = int(un['p_begin']), int(un['p_end']), 0, 0
            if st['lp'] < st['lpe']:
                setup_pair()

        def fetch():
            unit = st['lu']
            lK, lkt = st['lK'], st['lkt']
            if lK > 0:
                q = st['q']
                k = lkt * kt + np.arange(kt)
                kv = k < lK
                ai, bi = st['ai'], st['bi']
                Am = np.where((ai[:, None] >= 0) & kv[None, :],
                              A[np.maximum(ai[:, None] + k[None, :] * q['a_sk'], 0)], 0.0) * q['scale']
                Bm = np.where(kv[:, None] & (bi[None, :] >= 0),
                              B[np.maximum(k[:, None] * q['b_sk'] + bi[None, :], 0)], 0.0)
                tile = (Am, Bm)
            else:
                tile = None
            more_k = lK > 0 and (lkt + 1) * kt < lK
            more_p = lK > 0 and st['lp'] + 1 < st['lpe']
            last = not more_k and not more_p
            if more_k:
                st['lkt'] += 1
            elif more_p:
                st['lp'] += 1
                st['lkt'] = 0
                setup_pair()
            else:
                st['lu'] += 1
                if st['lu'] < n:
                    begin_unit()
            return tile, last, unit

        st['lu'] = 0
        begin_unit()
        cur = fetch()
        acc = np.zeros((TMs, TNs))
        while True:
            nvalid = st['lu'] < n
            nxt = fetch() if nvalid else None
            tile, last, unit = cur
            nitems += 1
            if tile is not None:
                acc += tile[0] @ tile[1]
            if last:
                un = rng_[unit]
                if un['slot'] >= 0:
                    work[un['slot']] = acc
                else:
                    m = un['m0'] + np.arange(TMs)
                    nn = un['n0'] + np.arange(TNs)
                    mv, nv = m < un['M'], nn < un['N']
                    ci = (un['c_off'] + (m[mv] // un['c_mlo']) * un['c_smhi']
                          + (m[mv] % un['c_mlo']) * un['c_smlo'])[:, None] + nn[nv][None, :] * un['c_sn']
                    v = acc[np.ix_(mv, nv)]
                    C[ci] = v if beta == 0 else C[ci] + v
                acc = np.zeros((TMs, TNs))
            if not nvalid:
                break
            cur = nxt
    t = prog.tasks_host
    for g in groups:
        tk = t[g['task']]
        s_ = work[g['slot_begin']:g['slot_begin'] + g['nslots']].sum(0)
        m = g['m0'] + np.arange(TMs)
        nn = g['n0'] + np.arange(TNs)
        mv, nv = m < tk['M'], nn < tk['N']
        ci = (tk['c_off'] + (m[mv] // tk['c_mlo']) * tk['c_smhi']
              + (m[mv] % tk['c_mlo']) * tk['c_smlo'])[:, None] + nn[nv][None, :] * tk['c_sn']
        v = s_[np.ix_(mv, nv)]
        C[ci] = v if beta == 0 else C[ci] + v
    return nitems


def emulate_chain(prog, A, B1, B2, C, beta=0):
    """Model of zmpo_chaingemm: per unit, the group's A tile (staged A
    blocks), the intermediate columns from the first-stage tasks, then the
    output blocks from it."""
    A, B1, B2, C = _np(A), _np(B1), _np(B2), _np(C)
    t, p, g, bl = prog.tasks_host, prog.pairs_host, prog.groups_host, prog.blocks_host
    tm = 16 * prog.rb
    for l_ in prog.launches:
        for u in l_['units']:
            gg = g[u['group']]
            assert gg['kcols'] <= l_['kmax'] and gg['ncols'] <= l_['nmax']
            mv = min(tm, int(gg['M']) - int(u['m0']))
            m = int(u['m0']) + np.arange(mv)
            atile = np.zeros((mv, max(int(gg['kcols']), 1)))
            for bk in bl[int(gg['ab_begin']):int(gg['ab_end'])]:
                k = np.arange(int(bk['K']))
                ai = u['a_base'] + bk['a_off'] + (m // bk['a_mlo']) * bk['a_smhi'] + (m % bk['a_mlo']) * bk['a_smlo']
                atile[:, int(bk['col']) + k] = A[ai[:, None] + k[None, :] * bk['a_sk']]
            mid = np.zeros((mv, max(int(gg['ncols']), 1)))
            for lo, hi, src, B, base, to_mid in (
                    (gg['t1_begin'], gg['t1_end'], atile, B1, u['b1_base'], True),
                    (gg['t2_begin'], gg['t2_end'], mid, B2, u['b2_base'], False)):
                for ti in range(int(lo), int(hi)):
                    tk = t[ti]
                    n = np.arange(int(tk['N']))
                    acc = np.zeros((mv, len(n)))
                    for q in p[int(tk['p_begin']):int(tk['p_end'])]:
                        k = np.arange(int(q['K']))
                        bi = base + q['b_off'] + (n // q['b_nlo']) * q['b_snhi'] + (n % q['b_nlo']) * q['b_snlo']
                        acc += q['scale'] * (src[:, int(q['a_off']) + k] @ B[k[:, None] * q['b_sk'] + bi[None, :]])
                    if to_mid:
                        mid[:, int(tk['c_off']):int(tk['c_off']) + len(n)] = acc
                    else:
                        ci = (u['c_base'] + tk['c_off'] + (m // tk['c_mlo']) * tk['c_smhi']
                              + (m % tk['c_mlo']) * tk['c_smlo'])[:, None] + n[None, :] * tk['c_sn']
                        C[ci] = acc if beta == 0 else C[ci] + acc


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

    def gram(self, X, nx, Y, ny):
        return np.array([self.dots(X, nx, Y[i]) for i in range(ny)])

    def project_out(self, r, X, nx):
        return self.lincomb(r, X, self.dots(X, nx, r), -1.0, 1.0)

    def residual(self, r, W, V, nb, coef, theta):
        self.lincomb(r, W, coef[:nb], 1.0, 0.0)
        self.lincomb(r, V, coef[:nb], -theta, 1.0)
        return r

    def norm(self, x):
        return float(np.linalg.norm(_np(x)))

    def precond(self, r, diag, e, shift):
        rr = _np(r)
        rr[:] = rr / (np.abs(_np(diag) - e) + shift)

    def copy(self, out, src):
        _np(out)[:] = _np(src).reshape(-1)
        return out

    def scale(self, x, s):
        xx = _np(x)
        xx[:] = 0.0 if s == 0.0 else s * xx
        return x

    def axpy(self, out, a, x):
        o = _np(out)
        o[:] = o + a * _np(x).reshape(-1)
        return out


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
    patch(dv, 'clear', lambda x, code=None: x.zero_())
    patch(dv, 'upload', upload)
    patch(dv, 'download', lambda x: np.array(_np(x)))
    patch(dv, 'device', lambda: torch.device('cpu'))
    patch(dv, 'require', lambda: None)
    patch(pd, '_VecOps', _CpuVecOps)
    patch(qq, 'run_program', emulate_wave)
    patch(kq, 'run_program', emulate_wave)
    patch(kq, 'run_chain', lambda prog, A, B1, B2, C, dtype=0, beta=0, stream=None:
          emulate_chain(prog, A, B1, B2, C, beta))
    patch(qq, 'gather', _gather)
    patch(kq, 'gather', _gather)

    def env_buffer(self, dtype=0):
        # the model computes in FP64 whatever the requested precision
        if isinstance(self.buf, np.ndarray):
            self.buf = upload(self.buf)
        return self.buf
    patch(kq.EnvStore, 'buffer', env_buffer)
    patch(dv, 'sync', lambda: None)
    try:
        yield
    finally:
        for (mod, name), fn in saved.items():
            setattr(mod, name, fn)
