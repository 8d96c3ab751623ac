"""msplat on MI355X: a HIP (gfx950) backend for msplat's 3D Gaussian Splatting training
and render path, behind msplat's backend interface (core/metal/bindings.h) and Python
surface (python/bindings.cpp, python/msplat/__init__.py). The names the reference's
package exports come from msplat._core, loaded on first use."""

__version__ = "1.1.4"

_CORE = ("TrainingConfig", "TrainingStats", "Dataset", "GaussianTrainer", "sync", "cleanup")
__all__ = list(_CORE)


def __getattr__(name):
    if name in _CORE:
        from . import _core
        return getattr(_core, name)
    raise AttributeError(f"module 'msplat' has no attribute {name!r}")
