"""MPS interchange: the reference's on-disk qtensor layout and a checkpoint.

``qtensor_from_dump`` rebuilds a QTensor from the fields qtensor.dump writes
per tensor (qtensor/qtensor.py:188-261): attributes rank, nsyms, status and
the datasets qsyms (the sectors of all legs stacked), ndims, idlst (all
legs' index lists flattened) and value (allowed blocks in ascending block
order, row-major) -- read with h5py where available, or from arrays already
extracted.  The block tables are recomputed and must reproduce the stored
idx_allowed.  ``save`` / ``load`` checkpoint the left-canonical MPS of a
sweep (sites and bond quantum numbers) to one .npz file.
"""
import numpy as np

from .qtensor import QTensor


def _split(flat, counts):
    out, k = [], 0
    for c in counts:
        out.append(flat[k:k + c])
        k += c
    return out


def qtensor_from_dump(rank, nsyms, status, qsyms, ndims, idlst, value, idx_allowed=None,
                      nqnum=2):
    nsyms = [int(x) for x in np.ravel(nsyms)]
    q = np.asarray(qsyms, dtype=np.float64).reshape(-1, nqnum)
    nd = np.asarray(ndims, dtype=np.int64).ravel()
    qt = QTensor([bool(x) for x in np.ravel(status)])
    qt.fromQsyms(int(rank), _split(q, nsyms), _split(nd, nsyms), ifallocate=False)
    if idx_allowed is not None and not np.array_equal(qt.idx_allowed,
                                                      np.asarray(idx_allowed).astype(np.int64)):
        raise ValueError('stored block table differs from the block rule')
    ids = np.asarray(idlst, dtype=np.int64).ravel()
    legs, k = [], 0
    for leg_nd in _split(nd, nsyms):
        secs = []
        for n in leg_nd:
            secs.append([int(x) for x in ids[k:k + n]])
            k += n
        legs.append(secs)
    qt.idlst = legs
    v = np.asarray(value, dtype=np.float64).ravel()
    if len(v) != qt.size_allowed:
        raise ValueError('stored value length %d, block tables give %d' % (len(v), qt.size_allowed))
    qt.value = v
    return qt


def save(path, sites, qnums):
    """Checkpoint an MPS: ``sites`` {i: QTensor (host or device values)},
    ``qnums`` the bond quantum numbers [K + 1 arrays]."""
    arrs = {'nsite': len(sites), 'nbond': len(qnums)}
    for i, qt in sites.items():
        v = qt._host_value()
        arrs['s%d_status' % i] = np.asarray(qt.status)
        arrs['s%d_nsyms' % i] = np.asarray(qt.nsyms)
        arrs['s%d_qsyms' % i] = np.concatenate(qt.qsyms)
        arrs['s%d_ndims' % i] = np.concatenate(qt.ndims)
        arrs['s%d_idlst' % i] = np.array([x for leg in qt.idlst for sec in leg for x in sec],
                                         dtype=np.int64)
        arrs['s%d_value' % i] = v
    for b, q in enumerate(qnums):
        arrs['q%d' % b] = np.asarray(q)
    np.savez_compressed(path, **arrs)


def load(path):
    """(sites, qnums) written by ``save`` (values on the host)."""
    d = np.load(path)
    sites = {}
    for i in range(int(d['nsite'])):
        st = d['s%d_status' % i]
        sites[i] = qtensor_from_dump(len(st), d['s%d_nsyms' % i], st, d['s%d_qsyms' % i],
                                     d['s%d_ndims' % i], d['s%d_idlst' % i], d['s%d_value' % i],
                                     nqnum=d['s%d_qsyms' % i].shape[1])
    qnums = [d['q%d' % b] for b in range(int(d['nbond']))]
    return sites, qnums
