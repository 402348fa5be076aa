"""Name-based backend dispatch, the counterpart of shenfun's
shenfun/optimization/__init__.py:37-80.

``optimizer(func)`` looks at ``config['optimization']['mode']`` and returns the
function of the same name -- ``func.__name__``, else
``func.__qualname__.replace('.', '_')`` (``TDMA.Solve`` -> ``TDMA_Solve``) --
from the backend module of that mode, or ``func`` itself when the mode names
no backend or the backend lacks the name (``verbose`` reports it).  The only
backend here is 'hip' (:mod:`shenfun_amd.optimization.hip`), which exports
shenfun's Cython names (la.pyx, transforms.pyx, Matvec.pyx) with their
signatures.  ``runtimeoptimizer`` resolves at every call, so a change of the
mode takes effect at once.
"""
import importlib
from functools import wraps

from ..config import config

BACKENDS = {"hip": "shenfun_amd.optimization.hip"}


def backend(mode=None):
    """The backend module of ``mode`` (default: the configured one), or None."""
    mode = (config["optimization"]["mode"] if mode is None else mode).lower()
    name = BACKENDS.get(mode)
    return importlib.import_module(name) if name else None


def optimizer(func, wrap=True):
    """Replace ``func`` by the configured backend's function of the same name
    (shenfun/optimization/__init__.py:47-80)."""
    verbose = config["optimization"].get("verbose", False)
    mod = backend()
    if mod is None:
        if verbose:
            print(func.__qualname__ + " not optimized")
        return func
    fun = getattr(mod, func.__name__, func)
    if fun is func:
        fun = getattr(mod, func.__qualname__.replace(".", "_"), func)
    if verbose and fun is func:
        print(func.__qualname__ + " not optimized")
    if wrap is False:
        return fun

    @wraps(func)
    def wrapped_function(*args, **kwargs):
        return fun(*args, **kwargs)
    return wrapped_function


class runtimeoptimizer:
    """Resolve ``func`` through :func:`optimizer` at every call
    (shenfun/optimization/__init__.py:37-45)."""

    def __init__(self, func):
        self.func = func
        self.__doc__ = func.__doc__
        self.__name__ = func.__name__
        self.__qualname__ = func.__qualname__

    def __call__(self, *args, **kwargs):
        return optimizer(self.func)(*args, **kwargs)
