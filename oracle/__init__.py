"""CPU oracle of the gravitation hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only to check results against (or as the timed CPU baseline), never
as a code path of the product.  The product (``gravitation_amd``) has no CPU fallback.

Contents
--------
``oracle.oracle``   NumPy restatements (FP64 all-pairs accelerations, semi-implicit Euler
                    stage 2, a pure-Python pair loop for tiny N, the np1/np4 CPU algorithms,
                    the adjoint of a trajectory) each citing the reference lines it follows.
``oracle.c``        the same FP64 all-pairs sum in plain C with Neumaier-compensated
                    accumulation (``liboracle.so``, built by ``oracle/Makefile``).
``oracle.ref``      ctypes bindings of the reference's own C stage 1 (``_lib1_/lib.c``,
                    ``_lib4_/lib.c``) compiled from /root/reference into ``oracle/_ref/`` by
                    ``oracle/Makefile`` (not committed; the built files travel with the tree).

How each function is pinned (checks live in tests/test_oracle.py):
* ``accel_f64``: brute-force pure-Python py1 loop on N <= 8; two-body and regular-polygon
  closed forms; the reference's solar-system known answer |a_earth| = G M_sun / AU^2; the
  softened pair closed form; invariants (momentum, translation, scaling, permutation); the
  compiled reference ``_lib1_`` (FP32) and the C oracle (FP64, compensated).
* ``stage2``: closed forms of semi-implicit Euler (a = 0: straight line; constant a over k
  steps: exact polynomial), update order (v first) by a one-step known answer.
* ``np1_accel`` / ``np4_accel``: the FP64 oracle (same inputs, FP32 tolerance).
* ``trajectory`` / ``adjoint``: FP64 central finite differences and torch autograd (CPU FP64).
"""
