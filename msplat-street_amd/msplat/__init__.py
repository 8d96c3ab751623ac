"""msplat on MI355X: a HIP (gfx950) backend for msplat's 3D Gaussian Splatting training
and render path, behind msplat's backend interface (core/metal/bindings.h) and Python
surface (python/bindings.cpp)."""

__version__ = "1.1.4"
