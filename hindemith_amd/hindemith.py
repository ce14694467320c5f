"""Drop-in for the reference's ``hindemith`` module (hindemith.py:1-30).

``from hindemith_amd.hindemith import *`` gives ``hm``, ``defines``, the
operand types (``Stencil``, ``SpMat2D``, ``sum2d``) and numpy, like the
reference's star imports (hindemith.py:2-4).  Decorating a function with
``@hm`` compiles it on first call per kwargs signature and runs it on the GPU
through libhmx; without the decorator the same body runs on NumPy.

Calls take keyword arguments only; positional arguments are ignored, as in
the reference (hindemith.py:22-27).
"""
import inspect
import warnings

from numpy import *  # noqa: F401,F403  (the reference re-exports numpy)
import scipy.sparse  # noqa: F401

from .hlib import *  # noqa: F401,F403
from .hlib import SpMat2D, Stencil, sum2d  # noqa: F401
from . import frontend as _fe
from ._lib import HindemithError  # noqa: F401

state = {}
defines = {}


class hm_fn:
    """The decorated function (hindemith.py:17-27)."""

    def __init__(self, fn):
        self.fn = fn
        self.source = inspect.getsource(fn)
        self.__name__ = getattr(fn, "__name__", "hm_fn")
        self.__doc__ = getattr(fn, "__doc__", None)
        self.last_record = None

    def __call__(self, *args, **kwargs):
        if args:
            warnings.warn("@hm functions take keyword arguments only; positional arguments "
                          "are ignored (as in the reference)", stacklevel=2)
        rec = _fe.mycompile(self.source, kwargs, defines, fn_key=(id(self.fn), self.source))
        self.last_record = rec
        return _fe.run_backend(rec, kwargs)

    def compile(self, **kwargs):
        """Compile for this kwargs signature without running; returns the plan."""
        rec = _fe.mycompile(self.source, kwargs, defines, fn_key=(id(self.fn), self.source))
        self.last_record = rec
        return rec.plan()

    def plan(self):
        """Launch plan of the most recent call (one line per kernel group)."""
        return self.last_record.plan() if self.last_record else ""

    def program_source(self, compile=False):
        """HIP source of the kernels generated for the most recent call's fused
        groups (compile=True also builds them with hipRTC)."""
        return self.last_record.program_source(compile) if self.last_record else ""


def hm(fn):
    return hm_fn(fn)
