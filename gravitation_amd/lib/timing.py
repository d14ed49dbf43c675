"""Step timing for the benchmark worker.

What is measured is what the reference's worker reports (src/gravitation/cli/worker.py:119-136,
its timers in src/gravitation/lib/timing.py:41-78): the wall-clock duration of every timed run
in integer nanoseconds, and the shortest of them as the best time.  Here the clock is the
monotonic ``perf_counter_ns``, which no system clock adjustment can move during a run.  A
``hip1`` step with ``sync_host=True`` returns only after its device work and copies have
finished, so the wall-clock duration of ``step()`` covers the whole step.
"""

from time import perf_counter_ns


class RunClock:
    """Durations [ns] of the runs bracketed by ``start()`` / ``stop()``."""

    __slots__ = ('runs', '_t0')

    def __init__(self):
        self.runs = []
        self._t0 = None

    def start(self):
        if self._t0 is not None:
            raise RuntimeError('RunClock.start(): a run is already being timed')
        self._t0 = perf_counter_ns()

    def stop(self):
        """Ends the current run; returns and records its duration in ns."""
        t1 = perf_counter_ns()
        if self._t0 is None:
            raise RuntimeError('RunClock.stop(): no run is being timed')
        dt = t1 - self._t0
        self._t0 = None
        self.runs.append(dt)
        return dt

    def best(self):
        """The shortest recorded run [ns]."""
        if not self.runs:
            raise RuntimeError('RunClock.best(): no run recorded')
        return min(self.runs)


def stopwatch():
    """A callable returning the nanoseconds elapsed since stopwatch() was called."""
    t0 = perf_counter_ns()
    return lambda: perf_counter_ns() - t0
