"""Sweep schedules of the two-site driver (source/mpo_dmrg_schedule.py).

``Schedule.fixed()`` is the 'Fixed2' schedule (schedule.py:105-122): every
sweep at maxM with one Davidson tolerance and noise.  ``Schedule.normal()``
is 'Normal2' (schedule.py:126-168, the two-site form of the BLOCK code's
schedule): the bond dimension starts at min(max(int(1.5 sqrt(maxM)), 25),
maxM) and doubles every sweep until it reaches maxM; the Davidson
tolerance starts at 1e-4 and is divided by 10 after every sweep at maxM
until it falls below ``tol``; the perturbative noise (pRDMQt) is 1e-4 until
the first sweep at maxM and 0 afterwards.  Convergence (schedule.py:82-93):
'Normal2' needs isweep > change = (number of scheduled sweeps) + 2 and
|dE| < tol; 'Fixed2' only |dE| < tol.  The one-site schedules (Normal0/1,
Fixed1) are not provided: the device path implements the two-site step.
"""
import math


class Schedule:
    def __init__(self, maxM=1000, tol=1e-7, maxiter=40):
        self.maxM = maxM
        self.tol = tol
        self.maxiter = maxiter
        self.startM = None
        self.startTol = 1e-4
        self.startNoise = 1e-4
        self.coff = 2
        self.change = None
        self.tag = None
        self.sweeps, self.maxMs, self.tols, self.noises = [], [], [], []

    def fixed(self, maxM=None, tol=None, maxiter=None, noise=1e-4):
        """'Fixed2': one entry, used for every sweep."""
        self.tag = 'Fixed2'
        if maxM is not None:
            self.maxM = maxM
        if tol is not None:
            self.tol = tol
        if maxiter is not None:
            self.maxiter = maxiter
        self.sweeps, self.maxMs, self.tols, self.noises = [0], [self.maxM], [self.tol], [noise]
        self.change = 0
        return self

    def normal(self):
        """'Normal2': bond-dimension ramp, then tightening tolerance."""
        self.tag = 'Normal2'
        self.sweeps, self.maxMs, self.tols, self.noises = [], [], [], []
        if self.startM is None:
            self.startM = min(max(int(1.5 * math.sqrt(self.maxM)), 25), self.maxM)
        m = min(self.startM, self.maxM)
        n = 0
        tol, noise = self.startTol, self.startNoise
        while m < self.maxM:
            self._add(n, m, tol, noise)
            n += 1
            m *= 2
        inoise = 0
        while tol >= self.tol:
            self._add(n, self.maxM, tol, noise)
            n += 1
            tol /= 10.0
            inoise += 1
            if inoise == 1:
                noise = 0.0
        self.change = n + self.coff
        return self

    def _add(self, n, m, tol, noise):
        self.sweeps.append(n)
        self.maxMs.append(m)
        self.tols.append(tol)
        self.noises.append(noise)

    def parameters(self, isweep):
        """(Dmax, crit_e, noise) of sweep isweep (schedule.py:37-80)."""
        if self.tag == 'Normal2':
            i = min(len(self.maxMs) - 1, isweep)
        elif self.tag == 'Fixed2':
            i = -1
        else:
            raise ValueError('schedule not set: call fixed() or normal()')
        return self.maxMs[i], self.tols[i], self.noises[i]

    def apply(self, dmrg, isweep):
        dmrg.Dmax, dmrg.crit_e, dmrg.noise = self.parameters(isweep)

    def converged(self, isweep, delta_e):
        if self.tag == 'Normal2':
            return isweep > self.change and abs(delta_e) < self.tol
        return abs(delta_e) < self.tol
