"""Plugin base class: the front-end contract every gravitation kernel implements.

This restates the reference's ``universe_base`` (src/gravitation/kernel/_base_.py:64-177) and
``_point_mass`` (:39-62) so that a kernel written against either runs unchanged:

* lifecycle PREINIT -> STARTED -> STOPPED with the same ``SyntaxError`` messages on misuse
  (:110-113, :124-127, :139-142, :167-170) and ``NotImplementedError`` for a missing stage 1
  (:147-150);
* ``G <- G * scale_r**3 / scale_m`` unless ``scale_off`` is passed (:85-88); leftover keyword
  arguments land in ``self._meta`` (:93);
* ``add_object`` scales r, v and m in place in the caller's objects (:114-117);
* ``step()`` = stage 1, stage 2, stage 3 (:136-145); the default stage 2 is the semi-implicit
  Euler ``_point_mass.move`` (:58-62); stage 3 advances t by T (:158-161).

Additions that do not change the contract: ``add_objects`` (bulk insertion of N bodies from
arrays, same scaling rules) and ``add_mass`` (an alias of ``add_object``).
"""

import numpy as np

STATE_PREINIT = 0
STATE_STARTED = 1
STATE_STOPPED = 2


class _point_mass:
    """One body: name, position r, velocity v, acceleration a, mass m.

    r, v and a are length-3 sequences (lists, or NumPy row views once a kernel binds them to
    its own arrays, as src/gravitation/kernel/np2.py:70-75 does).
    """

    def __init__(self, name, r, v, m):
        self._name = name
        self._r = r
        self._v = v
        self._a = [0.0] * len(r)
        self._m = m

    def __str__(self):
        x, y, z = (float(c) for c in self._r)
        vx, vy, vz = (float(c) for c in self._v)
        return '%s | %.4e, %.4e, %.4e | %.4e, %.4e, %.4e' % (self._name, x, y, z, vx, vy, vz)

    def move(self, T):
        """Default stage 2 for one body: v <- a*T + v, then r <- v*T + r, then a <- 0."""
        self._v[:] = [a * T + v for v, a in zip(self._v, self._a)]
        self._r[:] = [v * T + r for r, v in zip(self._r, self._v)]
        self._a[:] = [0.0] * len(self._a)


class universe_base:
    """Kernel base class.  Derive and implement at least ``step_stage1``."""

    def __init__(self, t=0.0, T=1.0e3, G=6.6740831e-11, scale_m=1.0, scale_r=1.0,
                 dtype='float32', threads=1, **kwargs):
        """Must not be overridden."""
        self._scale_m = scale_m
        self._scale_r = scale_r
        self._t = t
        self._T = T
        if kwargs.pop('scale_off', False):
            self._G = G
        else:
            self._G = G * (self._scale_r ** 3) / self._scale_m
        self._mass_list = []
        self._state = STATE_PREINIT
        self._dtype = dtype
        self._threads = threads
        self._meta = kwargs

    def __iter__(self):
        return iter(self._mass_list)

    def __len__(self):
        return len(self._mass_list)

    def __str__(self):
        return '\n'.join(str(pm) for pm in self._mass_list)

    def _check_preinit(self):
        if self._state == STATE_STARTED:
            raise SyntaxError('simulation was started')
        if self._state == STATE_STOPPED:
            raise SyntaxError('simulation was stopped')

    def add_object(self, **kwargs):
        """Adds one point mass (keywords name, r, v, m); only before ``start``."""
        self._check_preinit()
        if not kwargs.pop('scale_off', False):
            kwargs['r'][:] = [c * self._scale_r for c in kwargs['r']]
            kwargs['v'][:] = [c * self._scale_r for c in kwargs['v']]
            kwargs['m'] *= self._scale_m
        self._mass_list.append(_point_mass(**kwargs))

    add_mass = add_object

    def add_objects(self, r, v, m, names=None, scale_off=False):
        """Adds N point masses from arrays r (N, 3), v (N, 3), m (N,); only before ``start``.

        The same scaling as ``add_object`` is applied to the caller's arrays in place (when
        they are writable float arrays) and every body's r/v is a row view of them.
        """
        self._check_preinit()
        r = np.asarray(r)
        v = np.asarray(v)
        m = np.asarray(m)
        if r.ndim != 2 or r.shape != v.shape or m.shape != (r.shape[0],):
            raise ValueError('add_objects: r and v must be (N, D) and m (N,)')
        if not scale_off:
            if self._scale_r != 1.0:
                r *= self._scale_r
                v *= self._scale_r
            if self._scale_m != 1.0:
                m *= self._scale_m
        if names is None:
            names = ['obj%d' % (len(self._mass_list) + k) for k in range(r.shape[0])]
        self._mass_list.extend(
            _point_mass(name, r[k], v[k], m[k].item()) for k, name in enumerate(names))

    def start(self):
        """Moves PREINIT -> STARTED and calls ``start_kernel``."""
        if self._state == STATE_STARTED:
            raise SyntaxError('simulation is running')
        if self._state == STATE_STOPPED:
            raise SyntaxError('simulation was stopped')
        self._state = STATE_STARTED
        self.start_kernel()

    def start_kernel(self):
        """Kernel-specific set-up; override if needed."""

    def step(self):
        """Runs stage 1, stage 2 and stage 3 of one time step.  Must not be overridden."""
        if self._state == STATE_PREINIT:
            raise SyntaxError('simulation was not started')
        if self._state == STATE_STOPPED:
            raise SyntaxError('simulation was stopped')
        self.step_stage1()
        self.step_stage2()
        self.step_stage3()

    def step_stage1(self):
        """Computes accelerations; must be overridden."""
        raise NotImplementedError()

    def step_stage2(self):
        """Advances velocities and positions (semi-implicit Euler); may be overridden."""
        for pm in self._mass_list:
            pm.move(self._T)

    def step_stage3(self):
        """Advances the simulation time.  Must not be overridden."""
        self._t += self._T

    def stop(self):
        """Moves STARTED -> STOPPED and calls ``stop_kernel``."""
        if self._state == STATE_PREINIT:
            raise SyntaxError('simulation was not started')
        if self._state == STATE_STOPPED:
            raise SyntaxError('simulation was stopped before')
        self._state = STATE_STOPPED
        self.stop_kernel()

    def stop_kernel(self):
        """Kernel-specific tear-down; override if needed."""
