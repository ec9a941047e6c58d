"""Progress messages around sub-tasks (reference: helpers/status.py).

``Status(message)`` prints "message ..." when created and "message ... done
(elapsed)" when it is updated or deleted, "FAILED" instead of "done" after
``fail()``; ``StatusDecorator(message)`` does the same around a function.
Verbosity is per instance, or the class attribute ``verbose`` when None.
"""
import functools
import time
from datetime import timedelta


def _elapsed(t0):
    return str(timedelta(seconds=time.perf_counter() - t0))


class Status:
    verbose = True

    def __init__(self, message, verbose=None):
        self._verbose = verbose
        self._message = message
        self._failed = False
        self._open = True
        if self._loud():
            print(f"{message} ...")
        self._t0 = time.perf_counter()

    def _loud(self):
        return Status.verbose if self._verbose is None else self._verbose

    def fail(self):
        self._failed = True

    def _close(self):
        if self._open and self._loud():
            print(f"{self._message} ... {'FAILED' if self._failed else 'done'} "
                  f"({_elapsed(self._t0)})")
        self._open = False

    def update(self, message=None, verbose=None):
        """Close the current sub-task and open the next (same message and
        verbosity unless given)."""
        self._close()
        self.__init__(self._message if message is None else message,
                      self._verbose if verbose is None else verbose)

    def __del__(self):
        try:
            self._close()
        except Exception:
            pass


class StatusDecorator:
    verbose = True

    def __init__(self, message, verbose=None):
        self._message = message
        self._verbose = verbose

    def _loud(self):
        return StatusDecorator.verbose if self._verbose is None else self._verbose

    def __call__(self, f):
        @functools.wraps(f)
        def run(*args, **kwargs):
            if self._loud():
                print(f"{self._message} ...")
            t0 = time.perf_counter()
            result = f(*args, **kwargs)
            if self._loud():
                print(f"{self._message} ... done ({_elapsed(t0)})")
            return result
        return run
