"""minim_amd -- MI355X-native energy/gradient and minimiser hot path of minim.

The compute lives in libminim_amd.so (hand-written HIP kernels for gfx950 and
a C++ host library behind the C ABI of include/minim_hip.h).  This package is
the Python mirror of the reference's State / Potential / Minimiser API used by
the tests and the benchmark.  The library is loaded on first use (so that
`python -m minim_amd.build` can rebuild it); there is no CPU fallback: a
missing library raises ImportError.
"""
from . import _capi
from .api import (BarAndHinge, Fire, GradDescent, Lbfgs, Lj2d, Lj3d, LjNd, Minimiser, PhaseField,
                  PhaseFieldUnstructured, Potential, State, dot, potential_constrain, potential_eval, potential_eval_async, potential_eval_mixed,
                  potential_eval_step, potential_time_kernel, potential_time_steps, autotune_clear, site_types, triangulation_lists)


def __getattr__(name):
    if name == "lib":
        return _capi.load()
    raise AttributeError(name)

__all__ = ["BarAndHinge", "Fire", "GradDescent", "Lbfgs", "Lj2d", "Lj3d", "LjNd", "Minimiser", "PhaseField",
           "PhaseFieldUnstructured", "Potential", "State", "dot", "potential_constrain", "potential_eval", "potential_eval_async", "potential_eval_mixed",
           "potential_eval_step", "potential_time_kernel", "potential_time_steps", "autotune_clear", "site_types", "triangulation_lists", "lib"]
