# Kernel meta data: literals only, so that a plugin loader can read them without importing
# the module (the reference's loader parses exactly these names by AST,
# src/gravitation/lib/load.py:77-99 and :125-144).

__longname__ = 'hip-backend (1)'
__version__ = '0.1.0'
__description__ = 'MI355X (gfx950) HIP kernel: SGPR-broadcast all-pairs stage 1, compensated FP32 or FP64, fused stage 2'
__requirements__ = ['numpy', 'torch']
__externalrequirements__ = ['rocm']
__interpreters__ = ['python3']
__parallel__ = False
__license__ = 'GPLv2'
__authors__ = [
    'gravitation-mi355x contributors',
]

from .wrapper import universe  # noqa: E402,F401
