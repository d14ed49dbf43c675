"""Single-kernel benchmark worker writing the reference's JSON-lines log.

    python -m gravitation_amd.lib.worker -k hip1 --scenario galaxy \
        --scenario_param '{"stars_len": 16384}' -i 10 -t 10

Restates the reference's worker (src/gravitation/cli/worker.py:109-244): it emits
``START``, ``INPUT`` (simulation / python / platform), ``PROCEDURE``, ``SIZE``, then per step
``STEP`` {runtime, gctime, counter} and ``BEST_TIME`` {value}, and finally ``EXIT`` ``OK``; on
an exception ``ERROR`` {msg: traceback} and ``EXIT`` ``BAD``.  Timing follows it too: GC
disabled, ``gc.collect()`` around each step, wall clock in ns, best = minimum
(src/gravitation/lib/timing.py:41-78; ``timing.RunClock`` here).  It runs ``min_iterations``
steps and then extra steps until ``min_total_runtime`` seconds (:226-242).  The log is accepted by the reference's
``analyze`` (src/gravitation/cli/analyze.py:47-116).

Additions: ``--seed`` (the reference's scenarios are unseeded), the ``plummer`` and ``cube``
scenarios, and ``--kernel_param`` (JSON keywords for the kernel, e.g. ``{"eps": 1e-3}``).
"""

import argparse
import gc
import importlib
import json
import os
import platform
import sys
import traceback

from .simulation import create_simulation, store_simulation
from .timing import RunClock, stopwatch

KERNEL_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'kernel')


def inventory():
    """Kernel names: entries of gravitation_amd/kernel not starting with '_' (the rule of
    src/gravitation/lib/load.py:45-48)."""
    out = []
    for item in sorted(os.listdir(KERNEL_DIR)):
        if item.startswith('_') or item.startswith('.'):
            continue
        if item.endswith('.py'):
            out.append(item[:-3])
        elif os.path.isdir(os.path.join(KERNEL_DIR, item)):
            out.append(item)
    return out


def load_kernel(name):
    return importlib.import_module('gravitation_amd.kernel.%s' % name).universe


class _Exit(Exception):
    pass


def run_worker(kernel, scenario='galaxy', scenario_param=None, data_out_file='data.npz',
               save_after_iteration=(), min_iterations=10, min_total_runtime=10, threads=1,
               seed=None, kernel_param=None, universe_class=None, out=None):
    """Runs the worker and writes its log lines to ``out`` (default stdout).  Returns 0 on
    ``EXIT OK`` and 1 on ``EXIT BAD``.  ``universe_class`` overrides the kernel lookup."""
    out = out or sys.stdout
    scenario_param = dict(scenario_param or {})
    save_after_iteration = tuple(save_after_iteration)
    counter = [0]

    def _msg(**d):
        out.write(json.dumps(d) + '\n')
        out.flush()

    def _bad():
        _msg(log='ERROR', msg=traceback.format_exc())
        _msg(log='EXIT', msg='BAD')
        raise _Exit()

    def _store(s):
        _msg(log='PROCEDURE', msg='Saving data after step %d ...' % counter[0])
        try:
            store_simulation(s, data_out_file, 'kernel=%s;len=%d;step=%d'
                             % (kernel, len(s), counter[0]))
        except Exception:
            _bad()
        _msg(log='PROCEDURE', msg='Data saved after step %d.' % counter[0])

    def _step(s, rt, gt):
        try:
            gc.collect()
            rt.start()
            s.step()
            rt_ = rt.stop()
            gt.start()
            gc.collect()
            gt_ = gt.stop()
        except Exception:
            _bad()
        counter[0] += 1
        if counter[0] in save_after_iteration:
            _store(s)
        _msg(log='STEP', runtime=rt_, gctime=gt_, counter=counter[0])
        _msg(log='BEST_TIME', value=rt.best())

    try:
        _msg(log='START')
        try:
            import psutil
            cores, threads_max = psutil.cpu_count(logical=False), psutil.cpu_count(logical=True)
        except ImportError:
            cores = threads_max = os.cpu_count()
        _msg(log='INPUT',
             simulation=dict(kernel=kernel, scenario=scenario, scenario_param=scenario_param,
                             min_iterations=min_iterations, min_total_runtime=min_total_runtime,
                             threads=threads),
             python=dict(build=list(platform.python_build()),
                         compiler=platform.python_compiler(),
                         implementation=platform.python_implementation(),
                         version=list(sys.version_info)),
             platform=dict(system=platform.system(), release=platform.release(),
                           version=platform.version(), machine=platform.machine(),
                           processor=platform.processor(), cores=cores, threads=threads_max,
                           _cpu={}, _gpu=_gpu_info()))
        min_total_ns = min_total_runtime * 10 ** 9
        _msg(log='PROCEDURE', msg='Creating simulation ...')
        try:
            cls = universe_class or load_kernel(kernel)
            if kernel_param:
                base = cls

                def cls(*a, **kw):       # noqa: E306  (kernel options through **kwargs)
                    kw.update(kernel_param)
                    return base(*a, **kw)
            s = create_simulation(scenario, cls, scenario_param, threads, seed)
        except Exception:
            _bad()
        _msg(log='PROCEDURE', msg='Simulation created.')
        _msg(log='SIZE', value=len(s))
        rt, gt, et = RunClock(), RunClock(), stopwatch()
        gc.disable()
        try:
            if 0 in save_after_iteration:
                _store(s)
            for _ in range(min_iterations):
                _step(s, rt, gt)
            et_ = et()
            if et_ >= min_total_ns:
                _msg(log='PROCEDURE', msg='Minimum steps sufficient.')
                _msg(log='EXIT', msg='OK')
                return 0
            _msg(log='PROCEDURE', msg='Extra steps required.')
            remaining = (min_total_ns - et_) // max(et_, 1) * min_iterations
            for _ in range(remaining):
                _step(s, rt, gt)
            _msg(log='EXIT', msg='OK')
            return 0
        finally:
            gc.enable()
    except _Exit:
        return 1


def _gpu_info():
    try:
        import torch
        if torch.cuda.is_available():
            p = torch.cuda.get_device_properties(0)
            return [{'name': p.name, 'gcnArchName': getattr(p, 'gcnArchName', ''),
                     'multi_processor_count': p.multi_processor_count,
                     'total_memory': p.total_memory}]
    except Exception:
        pass
    return {}


def main(argv=None):
    p = argparse.ArgumentParser(description='isolated single-kernel benchmark worker')
    p.add_argument('--kernel', '-k', required=True, choices=inventory())
    p.add_argument('--scenario', default='galaxy')
    p.add_argument('--scenario_param', default='{}')
    p.add_argument('--kernel_param', default='{}')
    p.add_argument('--data_out_file', '-o', default='data.npz')
    p.add_argument('--save_after_iteration', '-s', type=int, action='append', default=[])
    p.add_argument('--min_iterations', '-i', type=int, default=10)
    p.add_argument('--min_total_runtime', '-t', type=int, default=10)
    p.add_argument('--threads', '-p', type=int, default=1)
    p.add_argument('--seed', type=int, default=None)
    a = p.parse_args(argv)
    return run_worker(a.kernel, a.scenario, json.loads(a.scenario_param), a.data_out_file,
                      a.save_after_iteration, a.min_iterations, a.min_total_runtime, a.threads,
                      a.seed, json.loads(a.kernel_param))


if __name__ == '__main__':
    sys.exit(main())
