"""gravitation_amd: MI355X-native (gfx950, HIP) direct-summation N-body kernel with the
plugin contract of pleiszenburg/gravitation.

* ``gravitation_amd.kernel.hip1``: the drop-in kernel (``universe`` class).
* ``gravitation_amd.stages``: torch-tensor wrappers of the stage kernels.
* ``gravitation_amd._native``: ctypes binding of libgravhip.so (include/gravitation_hip.h).
"""

__version__ = '0.1.0'
