"""Extract the H6/STO-3G integrals of the reference's worked example into a fixture.

Source: samps/example_conversion/0.h6_nqt/mole.h5 in the reference (its FCI
energies are recorded in that directory's README:2-5).  h5py is not
installed, so this script asks ``h5ls`` for the single chunk's file address
of each dataset and decompresses the LZF-filtered chunk itself (LZF format:
control byte < 32 = literal run of ctrl+1 bytes, otherwise a back reference
of length (ctrl>>5)+2 (+ next byte when that field is 7) at distance
((ctrl&31)<<8) + next byte + 1).  Run once in a container that has the
reference mounted; writes tests/golden/h6_sto3g.npz.
"""
import os
import re
import subprocess
import sys

import numpy as np

REF = '/root/reference/source/samps/example_conversion/0.h6_nqt/mole.h5'
H5LS = '/opt/conda/bin/h5ls'
H5DUMP = '/opt/conda/bin/h5dump'


def lzf_decompress(data, out_len):
    out = bytearray()
    ip = 0
    n = len(data)
    while ip < n:
        ctrl = data[ip]
        ip += 1
        if ctrl < 32:
            out += data[ip:ip + ctrl + 1]
            ip += ctrl + 1
        else:
            ln = ctrl >> 5
            ref = len(out) - ((ctrl & 0x1f) << 8) - 1
            if ln == 7:
                ln += data[ip]
                ip += 1
            ref -= data[ip]
            ip += 1
            ln += 2
            for k in range(ln):
                out.append(out[ref + k])
    assert len(out) == out_len, (len(out), out_len)
    return bytes(out)


def read_dataset(name, shape):
    txt = subprocess.run([H5LS, '-v', '--address', REF + '/' + name], capture_output=True,
                         text=True, check=True).stdout
    rows = re.findall(r'0x[0-9a-f]+\s+(\d+)\s+(\d+)\s+\[', txt)
    assert len(rows) == 1, txt
    nbytes, addr = int(rows[0][0]), int(rows[0][1])
    with open(REF, 'rb') as f:
        f.seek(addr)
        raw = f.read(nbytes)
    out_len = int(np.prod(shape)) * 8
    buf = raw if nbytes == out_len else lzf_decompress(raw, out_len)
    return np.frombuffer(buf, dtype='<f8').reshape(shape).copy()


def read_attr(name):
    txt = subprocess.run([H5DUMP, '-m', '%.17g', '-a', name, REF], capture_output=True,
                         text=True, check=True).stdout
    return float(re.search(r'\(0\):\s*([-0-9.eE+]+)', txt).group(1))


def main(out=None):
    out = out or os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h6_sto3g.npz')
    h = read_dataset('int1e_spatial', (6, 6))
    eri = read_dataset('int2e_spatial', (6, 6, 6, 6))
    int1e = read_dataset('int1e', (12, 12))
    enuc = read_attr('cal/enuc')
    nelec = read_attr('cal/nelec')
    # README:2-5 of the same directory (pyscf FCI, total energies)
    efci = np.array([-5.1491130015, -5.1372384525, -5.1233644821, -5.1180021882])
    np.savez(out, h=h, eri=eri, int1e=int1e, enuc=enuc, nelec=nelec, efci_readme=efci)
    print('wrote', out)


if __name__ == '__main__':
    main(*sys.argv[1:])
