"""Device Davidson eigensolver (counterpart of source/tools/pdvdson.py:22-390).

Basis and sigma vectors live on the device in FP64; every inner product,
linear combination and the |diag-e|+shift preconditioner are libzmpo kernels
with a fixed reduction order.  Only the (<= 2*nfac*neig)-dimensional
Rayleigh-Ritz matrix and the subspace-collapse coefficients are solved on the
host (numpy eigh of a <= 24x24 matrix), as the reference does on rank 0.
"""
import numpy as np

from . import _lib
from . import device as dv


class DavidsonError(RuntimeError):
    pass


class _VecOps:
    MAXD = 64   # rows per dots call (workspace size)

    def __init__(self, n):
        self.n = n
        self.work = dv.empty(_lib.lib().zmpo_dots_workspace(self.MAXD), 0)
        self.out = dv.empty(self.MAXD * self.MAXD, 0)
        self.one = dv.upload(np.ones(1), 0)

    def copy(self, out, src):
        """out = src (FP64 vectors of length n)."""
        _lib.check(_lib.lib().zmpo_lincomb(0, dv.ptr(out), dv.ptr(src), self.n, 1,
                                           dv.ptr(self.one), 1.0, 0.0, self.n, dv.stream()))
        return out

    def scale(self, x, s):
        """x *= s (s = 0 clears x)."""
        _lib.check(_lib.lib().zmpo_lincomb(0, dv.ptr(x), None, self.n, 0, None, 1.0, float(s),
                                           self.n, dv.stream()))
        return x

    def axpy(self, out, a, x):
        """out += a x."""
        _lib.check(_lib.lib().zmpo_lincomb(0, dv.ptr(out), dv.ptr(x), self.n, 1,
                                           dv.ptr(self.one), float(a), 1.0, self.n, dv.stream()))
        return out

    def _dots_into(self, X, nx, y, dst):
        _lib.check(_lib.lib().zmpo_dots(0, dv.ptr(X), X.stride(0), nx, dv.ptr(y), self.n,
                                        dv.ptr(dst), dv.ptr(self.work), dv.stream()))

    def dots(self, X, nx, y):
        """[X_j . y for j < nx] as a host array (X rows of a 2-D device tensor)."""
        if nx == 0:
            return np.zeros(0)
        self._dots_into(X, nx, y, self.out)
        return dv.download(self.out[:nx]).copy()

    def gram(self, X, nx, Y, ny):
        """[[X_j . Y_i for j < nx] for i < ny] with one device-to-host copy."""
        for i in range(ny):
            self._dots_into(X, nx, Y[i], self.out[i * nx:])
        return dv.download(self.out[:nx * ny]).reshape(ny, nx).copy()

    def project_out(self, r, X, nx):
        """r -= sum_j (X_j . r) X_j, the coefficients staying on the device."""
        if nx == 0:
            return r
        self._dots_into(X, nx, r, self.out)
        _lib.check(_lib.lib().zmpo_lincomb(0, dv.ptr(r), dv.ptr(X), X.stride(0), nx,
                                           dv.ptr(self.out), -1.0, 1.0, self.n, dv.stream()))
        return r

    def lincomb(self, out, X, coef, alpha=1.0, beta=0.0):
        nx = len(coef)
        if nx == 0:
            if beta != 1.0:
                self.scale(out, beta)
            return out
        c = dv.upload(np.asarray(coef, dtype=np.float64), 0)
        _lib.check(_lib.lib().zmpo_lincomb(0, dv.ptr(out), dv.ptr(X), X.stride(0), nx, dv.ptr(c),
                                           float(alpha), float(beta), self.n, dv.stream()))
        return out

    def residual(self, r, W, V, nb, coef, theta):
        """r = W c - theta V c over the first nb rows (zmpo_davidson_residual)."""
        c = dv.upload(np.asarray(coef[:nb], dtype=np.float64), 0)
        _lib.check(_lib.lib().zmpo_davidson_residual(self.n, dv.ptr(W), dv.ptr(V), W.stride(0),
                                                     nb, dv.ptr(c), float(theta), None, 0.0,
                                                     dv.ptr(r), dv.stream()))
        return r

    def norm(self, x):
        return float(np.sqrt(max(self.dots(x.view(1, -1), 1, x)[0], 0.0)))

    def precond(self, r, diag, e, shift):
        _lib.check(_lib.lib().zmpo_precond(0, dv.ptr(r), dv.ptr(diag), float(e), float(shift),
                                           self.n, dv.stream()))


def _host_ortho(vbas, rbas, crit_indp):
    """Subspace-coefficient version of dvdson_ortho (pdvdson.py:350-390)."""
    rbas = rbas.copy()
    v2 = np.zeros_like(rbas)
    nindp = 0
    for _ in range(5):
        rbas = rbas - (rbas @ vbas.T) @ vbas
    for i in range(rbas.shape[0]):
        rv = rbas[i].copy()
        rii = np.linalg.norm(rv)
        if rii <= crit_indp:
            continue
        rv = rv / rii
        rv = rv / np.linalg.norm(rv)
        v2[nindp] = rv
        nindp += 1
        for _ in range(5):
            rbas[i:] -= (rbas[i:] @ vbas.T) @ vbas
            rbas[i:] -= (rbas[i:] @ v2[:nindp].T) @ v2[:nindp]
    return nindp, v2[:nindp].copy()


class EigenSolver:
    """pdvdson.eigenSolver restricted to iop=4, real, one process."""

    def __init__(self):
        self.maxcycle = 20
        self.crit_e = 1e-8
        self.crit_vec = 1e-8
        self.crit_indp = 1e-12
        self.lshift = 1e-4
        self.nfac = 10
        self.neig = 1
        self.ndim = 0
        self.diag = None        # device float64
        self.HVec = None        # HVec(x_dev, out_dev)
        self.nmvp = 0
        self.niter = 0
        # symmetry check of the projected H: sigma applications in FP32
        # leave an asymmetry at the FP32 rounding level
        self.herm_tol = 1e-10
        # residual criterion max(crit_vec, tol_scale * max(1, |e|)): the
        # level a low-precision sigma still resolves scales with |H psi|
        self.tol_scale = 0.0

    def _ortho(self, ops, V, nb, R, rows, out_start):
        """Orthonormalise R[rows] against V[:nb] (5 passes) into V[out_start:]."""
        for _ in range(5):
            for i in rows:
                ops.project_out(R[i], V, nb)
        nindp = 0
        for pos, i in enumerate(rows):
            rii = ops.norm(R[i])
            if rii <= self.crit_indp:
                continue
            dst = V[out_start + nindp]
            ops.copy(dst, R[i])
            ops.scale(dst, 1.0 / rii)
            ops.scale(dst, 1.0 / ops.norm(dst))
            nindp += 1
            for _ in range(5):
                for j in rows[pos:]:
                    ops.project_out(R[j], V, nb)
                    ops.project_out(R[j], V[out_start:], nindp)
        return nindp

    def solve_iter(self, v0=None, w0=None):
        """Lowest neig eigenpairs.  w0 (neig = 1 only): the sigma of v0,
        already known (for instance in a higher precision than HVec): the
        first basis vector keeps it instead of calling HVec."""
        n, neig = self.ndim, self.neig
        if neig > n:
            raise DavidsonError('neig > ndim')
        ops = _VecOps(n)
        maxv = self.nfac * neig + 2 * neig + 2
        # every row is written before it is read
        V = dv.empty(maxv * n, 0).view(maxv, n)
        W = dv.empty(maxv * n, 0).view(maxv, n)
        T = dv.empty(maxv * n, 0).view(maxv, n)
        R = dv.empty(neig * n, 0).view(neig, n)
        self.nmvp = 0
        if v0 is None:
            d = dv.download(self.diag) + 1e-12 * np.arange(1, 1 + n) / float(n)
            idx = np.argsort(d, kind='stable')[:neig]
            e = np.zeros((neig, n))
            e[np.arange(neig), idx] = 1.0
            ev = dv.upload(e.reshape(-1), 0).view(neig, n)
            for i in range(neig):
                ops.copy(V[i], ev[i])
        else:
            src = v0 if not isinstance(v0, np.ndarray) else dv.upload(v0, 0)
            src = src.reshape(neig, n).contiguous()
            for i in range(neig):
                ops.copy(V[i], src[i])
        # genOrthoBas (pdvdson.py:335-345)
        if w0 is not None and neig != 1:
            raise DavidsonError('a given first sigma needs neig = 1')
        nrm0 = ops.norm(V[0])
        ops.scale(V[0], 1.0 / nrm0)
        if neig > 1:
            R2 = dv.empty((neig - 1) * n, 0).view(neig - 1, n)
            for i in range(neig - 1):
                ops.copy(R2[i], V[1 + i])
            nindp = self._ortho(ops, V, 1, R2, list(range(neig - 1)), 1)
            if nindp + 1 != neig:
                raise DavidsonError('insufficient orthonormal basis')
        if w0 is not None:
            ops.copy(W[0], w0.reshape(-1).contiguous())
            ops.scale(W[0], 1.0 / nrm0)
        else:
            for i in range(neig):
                self.HVec(V[i], W[i])
            self.nmvp += neig
        nb = neig
        eigs = np.zeros(neig) + 1e3
        vr = None
        for niter in range(1, self.maxcycle):
            self.niter = niter
            S = ops.gram(V, nb, V, nb)
            if np.linalg.norm(S - np.identity(nb)) > 1e-10:
                raise DavidsonError('basis not orthonormal: %g' % np.linalg.norm(S - np.identity(nb)))
            H = ops.gram(W, nb, V, nb)
            if np.linalg.norm(H - H.T) > self.herm_tol * max(1.0, np.abs(H).max()):
                raise DavidsonError('projected H not Hermitian: %g' % np.linalg.norm(H - H.T))
            H = 0.5 * (H + H.T)
            e, u = np.linalg.eigh(H)
            vr = u[:, :neig].T.copy()
            teig = e[:neig]
            econv = [abs(teig[i] - eigs[i]) <= self.crit_e for i in range(neig)]
            eigs = teig.copy()
            rconv, rindx = [], []
            for i in range(neig):
                ops.residual(R[i], W, V, nb, vr[i], eigs[i])
                rn = ops.norm(R[i])
                rconv.append(rn <= max(self.crit_vec, self.tol_scale * max(1.0, abs(eigs[i]))))
                if not rconv[i] and not econv[i]:
                    rindx.append(i)
            if not rindx or niter == self.maxcycle - 1:
                break
            if nb > self.nfac * neig:
                pr = np.identity(nb)[:neig, :] - vr
                nindp, vr2 = _host_ortho(vr, pr[rindx, :], self.crit_indp)
                cm = np.vstack((vr, vr2)) if nindp else vr
                for k in range(cm.shape[0]):
                    ops.lincomb(T[k], V, cm[k], 1.0, 0.0)
                for k in range(cm.shape[0]):
                    ops.copy(V[k], T[k])
                for k in range(cm.shape[0]):
                    ops.lincomb(T[k], W, cm[k], 1.0, 0.0)
                for k in range(cm.shape[0]):
                    ops.copy(W[k], T[k])
                nb = cm.shape[0]
                vr = np.identity(nb)[:neig]
            for i in range(neig):
                if rconv[i]:
                    continue
                ops.precond(R[i], self.diag, eigs[i], self.lshift)
            nindp = self._ortho(ops, V, nb, R, rindx, nb)
            if nindp == 0:
                raise DavidsonError('unable to generate new direction')
            for k in range(nindp):
                self.HVec(V[nb + k], W[nb + k])
            self.nmvp += nindp
            nb += nindp
        vec = dv.empty(neig * n, 0).view(neig, n)
        for i in range(neig):
            ops.lincomb(vec[i], V, vr[i], 1.0, 0.0)
        return eigs, vec, self.nmvp
