"""Build libzmpo.so in-tree for gfx950 (``python -m zmpo_dmrg_amd.build``)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'csrc', 'zmpo.hip')
OUT = os.path.join(HERE, 'libzmpo.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I' + os.path.join(ROOT, 'include'),
         # rocSOLVER: the per-sector eigensolver of the truncation
         '-L/opt/rocm/lib', '-Wl,-rpath,/opt/rocm/lib', '-lrocsolver', '-lrocblas']


def build(verbose=False):
    tmp = OUT + '.tmp'
    cmd = [HIPCC] + FLAGS + ['-o', tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed for %s' % SRC)
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT)
    return OUT


if __name__ == '__main__':
    build(verbose=True)
