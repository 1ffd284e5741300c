"""Loader of oracle/_ref/_gpawref -- the reference's own serial C subset
compiled from /root/reference by oracle/ref/Makefile.  TEST ONLY: used in
the build container to write golden fixtures and to check the oracle; the
reference does not exist on the GPU box, so nothing there may need this.
"""
import importlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, '_ref')
REFERENCE = '/root/reference'


def available():
    return os.path.isdir(os.path.join(REFERENCE, 'c', 'bmgs'))


def build():
    subprocess.run(['make', '-s', '-C', os.path.join(HERE, 'ref'),
                    'REF=' + REFERENCE], check=True)


def load():
    """Import _gpawref.  Symbols resolve lazily: the reference's lfc.c
    references zgemm_ (used only by lcao_to_grid_k, never called here) and
    no BLAS is linked."""
    if '_gpawref' in sys.modules:
        return sys.modules['_gpawref']
    flags = sys.getdlopenflags()
    sys.path.insert(0, REF_DIR)
    try:
        sys.setdlopenflags(os.RTLD_LAZY)
        return importlib.import_module('_gpawref')
    finally:
        sys.setdlopenflags(flags)
        sys.path.remove(REF_DIR)
