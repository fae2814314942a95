"""Two-site sweep driver over the device kernels.

Counterpart of source/mpo_dmrg_class.py (mpo_dmrg fields :47-192, build
:279-297, default :329-434) and source/mpo_dmrg_opt.py (sweep :31-95,
checkOpers :153-193, sweep_updateSite :313-344, sweep_updateBoundary
:347-377), with the 'Fixed2' schedule (mpo_dmrg_schedule.py:100-122) and one
process.  ``dmrg.noise`` (default 0) adds the perturbative noise term
(pRDMQt, mpo_dmrg_kernelQt.py:453; used at mpo_dmrg_dotutil.py:313-317) to the
decimation density matrix.  Environments and the MPO live in device memory
(EnvStore / MPOStore) instead of HDF5 files; the Qt MPS starts from an occupation
product state (mpo_dmrg_util.py:118-155) followed by initOps (:193-214).
"""
import time

import numpy as np

from . import device as dv
from . import kernelQt
from .dot import ci_solver
from .mpo import MPOStore, qphys
from .qtensor import QTensor


def default_occupation(K, nelec, sz=0.0):
    """Aufbau occupation, alpha on even / beta on odd spin orbitals."""
    na = int(round(nelec / 2 + sz))
    nb = int(round(nelec / 2 - sz))
    occ = np.zeros(2 * K)
    occ[0:2 * na:2] = 1.0
    occ[1:2 * nb:2] = 1.0
    return occ


def mps0occun(K, occun):
    """Product-state MPS from an occupation (mpo_dmrg_util.py:118-155)."""
    qnuml = [np.array([[0., 0.]])]
    sites = []
    ne, ms = 0.0, 0.0
    for i in range(K):
        na, nb = occun[2 * i], occun[2 * i + 1]
        t = np.zeros(4)
        if na < 0.5 and nb < 0.5:
            t[0] = 1.0
        elif na < 0.5:
            ne, ms = ne + 1.0, ms - 0.5
            t[1] = 1.0
        elif nb < 0.5:
            ne, ms = ne + 1.0, ms + 0.5
            t[2] = 1.0
        else:
            ne += 2.0
            t[3] = 1.0
        sites.append(t.reshape(1, 4, 1))
        qnuml.append(np.array([[ne, ms]]))
    qnumr = [np.array([[ne - q[0][0], ms - q[0][1]]]) for q in qnuml]
    return qnuml, qnumr, sites


class MPODMRG:
    """The subset of the reference's mpo_dmrg object used on this path."""

    def __init__(self, int1e, int2e, nelec, sz=0.0, Dmax=100, tol=1e-8, crit_vec=1e-8,
                 occun=None, dtype=0, mpo=None, budget=kernelQt.DEFAULT_BUDGET,
                 keep_envs=False, env_budget=None, host_budget=None):
        self.fhop = mpo if mpo is not None else MPOStore(int1e, int2e)
        self.nsite = self.fhop.nsite
        self.sbas = self.fhop.sbas
        self.nops = self.fhop.nops
        self.opers = self.fhop.opers
        self.hpwts = self.fhop.hpwts
        self.qphys = [qphys() for _ in range(self.nsite)]
        self.qsectors = {(float(nelec), float(sz)): 1}
        self.Dmax = Dmax
        self.crit_e = tol
        self.crit_vec = crit_vec
        self.thresh = -1e-14
        self.noise = 0.0     # perturbative noise of the decimation RDM (pRDMQt)
        self.dtype = dtype
        self.budget = budget
        self.psi0 = None
        self.nsigma = 0
        self.sigma_flops = 0
        self.niters = []
        # wall seconds: 'solve' (the two-site step: 'diag', 'sigma_plan',
        # 'davidson', 'decimation', 'guess') and 'renorm'
        self.timing = {'solve': 0.0, 'renorm': 0.0}
        # keep_envs=False: a half-sweep releases the environments it will
        # not read again (the next half-sweep rebuilds them before use), so
        # about one environment per bond is resident instead of two
        self.keep_envs = keep_envs
        # env_budget (bytes): device memory the environments may hold; the
        # ones read farthest ahead move to host memory (at most host_budget
        # bytes) until the rest fit.  None keeps every environment resident.
        self.env_budget, self.host_budget = env_budget, host_budget
        K = self.nsite
        occun = default_occupation(K, nelec, sz) if occun is None else occun
        self.qnuml, self.qnumr, s0 = mps0occun(K, occun)
        self.lmps, self.rmps = {}, {}
        for i in range(K):
            q = QTensor([False, False, True])
            q.fromDenseTensor(s0[i], [self.qnuml[i], self.qphys[i], self.qnuml[i + 1]])
            self.lmps[i] = q
        # initOps: right boundary + every left environment (util.py:193-214)
        self.envR = {K: kernelQt.boundary_store(self.nops)}
        self.envL, _ = kernelQt.genHops(self.fhop, self.lmps, 'L', fit=self._fit_built)

    def _fit_built(self, envs, i):
        # left environments built for the 'R' half-sweep that follows
        fit_envs(self, i + 1, 'R', envL=envs)


def env_bytes(dmrg):
    """(device bytes, host bytes) of the FP64 environment values."""
    dev = host = 0
    for env in (dmrg.envL, dmrg.envR):
        for e in env.values():
            if e.resident():
                dev += e.nbytes()
            else:
                host += e.nbytes()
    return dev, host


def fit_envs(dmrg, isite, status, envL=None):
    """Hold the resident environments within dmrg.env_budget bytes before
    the two-site step at isite (half-sweep ``status``), counting the two it
    reads (envL[isite-1], envR[isite+2]) as resident.  The environments
    whose next read lies farthest ahead move to host memory first: in an
    'L' half-sweep the left ones (read in the next 'R' half-sweep, smallest
    bond last) and then the right ones from the far end; an 'R' half-sweep
    mirrors this.  The files of mpo_dmrg_opt.py:196-238 play this part in
    the reference, where every environment lives on disk."""
    budget = getattr(dmrg, 'env_budget', None)
    if budget is None:
        return
    envL = dmrg.envL if envL is None else envL
    envR = dmrg.envR
    need = {('L', isite - 1), ('R', isite + 2)}
    left = [('L', j) for j in sorted(envL)]
    right = [('R', j) for j in sorted(envR, reverse=True)]
    order = left + right if status == 'L' else right + left
    store = {'L': envL, 'R': envR}
    dev = host = 0
    for side, j in left + right:
        e = store[side][j]
        if e.resident() or (side, j) in need:
            dev += e.nbytes()
        else:
            host += e.nbytes()
    hbudget = getattr(dmrg, 'host_budget', None)
    for side, j in order:
        if dev <= budget:
            break
        e = store[side][j]
        if (side, j) in need or not e.resident():
            continue
        n = e.nbytes()
        if hbudget is not None and host + n > hbudget:
            raise MemoryError('environments need %.1f GiB of device memory and %.1f GiB of host '
                              'memory; budgets %.1f / %.1f GiB' % (
                                  dev / 2 ** 30, (host + n) / 2 ** 30, budget / 2 ** 30,
                                  hbudget / 2 ** 30))
        t0 = time.time()
        e.spill()
        dmrg.timing['spill'] = dmrg.timing.get('spill', 0.0) + time.time() - t0
        dev -= n
        host += n


def restart_from_left_mps(dmrg, sites, qnuml):
    """Continue from a left-canonical MPS (the state after an 'L' half-sweep;
    the last site holds the normalised coefficients): its sites and bond
    QNs replace the product state and every left environment is rebuilt
    (initOps, mpo_dmrg_util.py:193-214)."""
    K = dmrg.nsite
    for i in range(K):
        dmrg.lmps[i] = sites[i]
    for i in range(K + 1):
        dmrg.qnuml[i] = np.asarray(qnuml[i])
    dmrg.envL, dmrg.envR = {}, {}      # released before the new ones are built
    dmrg.envL, _ = kernelQt.genHops(dmrg.fhop, dmrg.lmps, 'L', fit=dmrg._fit_built)
    dmrg.envR = {K: kernelQt.boundary_store(dmrg.nops)}
    # the first step of the next 'R' half-sweep (sites K-2, K-1) starts from
    # the MPS's own two-site wavefunction
    a = dmrg.lmps[K - 2].toDenseTensor()
    b = dmrg.lmps[K - 1].toDenseTensor()
    dmrg.psi0 = np.tensordot(a, b, axes=([2], [0])).reshape(1, -1)


def sweep_updateSite(dmrg, isite, qred, site, status):
    """opt.py:313-344 (ncsite=2)."""
    # site: dense (host decimation) or the QTensor the device decimation built
    if status == 'L':
        dmrg.qnuml[isite + 1] = qred.copy()
        if isinstance(site, QTensor):
            q = site
        else:
            q = QTensor([False, False, True])
            q.fromDenseTensor(site, [dmrg.qlst[0], dmrg.qlst[1], qred])
        dmrg.lmps[isite] = q
    else:
        dmrg.qnumr[isite + 1] = qred.copy()
        if isinstance(site, QTensor):
            q = site
        else:
            q = QTensor([True, False, False])
            q.fromDenseTensor(site, [qred, dmrg.qlst[-2], dmrg.qlst[-1]])
        dmrg.rmps[isite + 1] = q
    return q


def sweep_updateBoundary(dmrg, qred, srotR, status):
    """opt.py:347-377 (ncsite=2)."""
    K = dmrg.nsite
    s0 = srotR[0] if isinstance(srotR, np.ndarray) else dv.download(srotR[0])
    smat = s0 / np.linalg.norm(s0)
    key = list(dmrg.qsectors.keys())[0]
    qsym = np.array([key])
    if status == 'L':
        q = QTensor([False, False, True])
        q.fromDenseTensor(smat, [qred, dmrg.qphys[K - 1], qsym])
        dmrg.lmps[K - 1] = q
        dmrg.qnuml[K] = qsym
    else:
        q = QTensor([True, False, False])
        q.fromDenseTensor(smat, [qsym, dmrg.qphys[0], qred])
        dmrg.rmps[0] = q
        dmrg.qnumr[0] = qsym


def release_dead_envs(dmrg, isite, status, after):
    """Drop the environments a half-sweep will not read again.

    'R' (right to left): step isite reads envL[isite-1] and envR[isite+2]
    and writes envR[isite+1]; envL[j >= isite] are not read again before
    the next 'L' half-sweep rebuilds them, nor is envL[isite-1] after the
    step.  'L' mirrors this.  Boundaries (envL[-1], envR[K]) stay."""
    K = dmrg.nsite
    if status == 'R':
        env, first = dmrg.envL, (isite - 1 if after else isite)
        dead = [j for j in env if j >= max(first, 0)]
        if not after:
            dead_r = [isite + 1] if isite + 1 < K and (isite + 1) in dmrg.envR else []
            for j in dead_r:
                del dmrg.envR[j]
    else:
        env, last = dmrg.envR, (isite + 2 if after else isite + 1)
        dead = [j for j in env if j <= min(last, K - 1)]
        if not after and isite in dmrg.envL and isite >= 0:
            del dmrg.envL[isite]
    for j in dead:
        del env[j]


def sweep(dmrg, status, max_steps=None):
    """One half sweep (opt.py:31-95); status 'R' runs right-to-left.  With
    max_steps the half-sweep stops early and keeps every environment."""
    K = dmrg.nsite
    actlst = list(range(K - 2, -1, -1)) if status == 'R' else list(range(0, K - 1))
    elst, dlst, nmvp = [], [], 0
    qred = srotR = None
    release = not getattr(dmrg, 'keep_envs', True) and max_steps is None
    for n, isite in enumerate(actlst):
        if max_steps is not None and n >= max_steps:
            break
        if release:
            release_dead_envs(dmrg, isite, status, after=False)
        fit_envs(dmrg, isite, status)
        flst = [[dmrg.envL[isite - 1], dmrg.envR[isite + 2]]]
        t0 = time.time()
        nm, eigs, sigs, dwts, qred, site, srotR = ci_solver(dmrg, isite, 2, actlst, flst, status)
        t1 = time.time()
        nmvp += nm
        elst.append(eigs)
        dlst.append(dwts)
        qsite = sweep_updateSite(dmrg, isite, qred, site, status)
        flstN = [[None]]
        kernelQt.renormQt_H00(dmrg, isite, 2, flst, flstN, qsite, status)
        if status == 'L':
            dmrg.envL[isite] = flstN[0][0]
        else:
            dmrg.envR[isite + 1] = flstN[0][0]
        del flst, flstN
        if release:
            release_dead_envs(dmrg, isite, status, after=True)
        t2 = time.time()
        dmrg.timing['solve'] += t1 - t0
        dmrg.timing['renorm'] += t2 - t1
        log = getattr(dmrg, 'step_log', None)
        if log is not None:
            log(status, isite, float(eigs[0]), t2 - t0, dmrg.nsigma)
    if max_steps is None:
        sweep_updateBoundary(dmrg, qred, srotR, status)
    eavs = np.array([e[0] for e in elst])
    indx = int(np.argmin(eavs))
    return nmvp, indx, float(eavs[indx]), elst, dlst


def run(dmrg, nsweeps, tol=None, schedule=None, log=None):
    """Sweeps 'R' then 'L'; energy = min of both half-sweep minima
    (class.py:385-430).  Without a schedule: Fixed2 at dmrg.Dmax, stopping
    when the energy changes by less than tol; with a schedule.Schedule
    (Fixed2 or Normal2) its parameters are set before every sweep and its
    convergence test is used (schedule.py:37-93).  log(isweep, eav, dt) is
    called after every sweep.  Returns the per-sweep energies."""
    tol = dmrg.crit_e if tol is None else tol
    eold = 1e3
    out = []
    for isweep in range(nsweeps):
        t0 = time.time()
        if schedule is not None:
            schedule.apply(dmrg, isweep)
        r1 = sweep(dmrg, 'R')
        r2 = sweep(dmrg, 'L')
        eav = min(r1[2], r2[2])
        out.append(eav)
        if log is not None:
            log(isweep, eav, time.time() - t0)
        conv = schedule.converged(isweep, eav - eold) if schedule is not None \
            else abs(eav - eold) < tol
        if conv:
            break
        eold = eav
    return out
