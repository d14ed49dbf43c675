"""Wall-clock timers with the reference's semantics (src/gravitation/lib/timing.py:41-99): every
run is recorded in integer nanoseconds and the benchmark reports the minimum.  The device work
of a ``hip1`` step is complete when ``step()`` returns with ``sync_host=True`` (the D2H copy
synchronises), so wall-clock timing of ``step()`` measures the whole step."""

from time import time_ns


class best_run_timer:
    """Records run times [ns]; ``min``/``avg``/``sum`` over all runs."""

    def __init__(self):
        self._state = []
        self._running = False
        self._lower = None

    def _need(self):
        if not self._state:
            raise SyntaxError('Nothing has been recorded.')

    def avg(self):
        self._need()
        return sum(self._state) // len(self._state)

    def min(self):
        self._need()
        return min(self._state)

    def sum(self):
        self._need()
        return sum(self._state)

    def start(self):
        if self._running:
            raise SyntaxError('Timer is running!')
        self._running = True
        self._lower = time_ns()

    def stop(self):
        upper = time_ns()
        if not self._running:
            raise SyntaxError('Timer is NOT running!')
        runtime = upper - self._lower
        self._lower = None
        self._running = False
        self._state.append(runtime)
        return runtime


class average_timer(best_run_timer):
    """As best_run_timer, keeping only the last ``maxlen`` runs."""

    def __init__(self, maxlen):
        super().__init__()
        self._maxlen = maxlen

    def stop(self):
        runtime = super().stop()
        if len(self._state) > self._maxlen:
            self._state.pop(0)
        return runtime


class elapsed_timer:
    """Nanoseconds since construction."""

    def __init__(self):
        self._state = time_ns()

    def __call__(self):
        return time_ns() - self._state
