"""cardiax_amd: MI355X-native (gfx950) drop-in for the forward-Euler path of
cardiax.solve, computed by hand-written HIP kernels behind a C-ABI."""
from . import convert, params, stimulus  # noqa: F401
from . import solve  # noqa: F401

__all__ = ["convert", "params", "stimulus", "solve"]
