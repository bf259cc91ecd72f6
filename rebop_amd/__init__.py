"""rebop_amd -- batched Gillespie SSA for MI355X (gfx950).

``Gillespie`` mirrors rebop's Python API; ``GillespieNet``/``Rate``/``Batch``
mirror its Rust run-time API (src/gillespie.rs), batched over independent
trajectories; ``define_system`` is the analogue of its macro DSL
(src/gillespie_macro.rs).  The compute path is the HIP library ``librebop_gpu.so``
(C ABI: include/rebop_gpu.h); there is no CPU fallback.
"""

from .engine import Batch, GillespieNet, Rate
from .gillespie import Gillespie
from .system import define_system
from ._native import RebopError, RebopValueError, Stats

__version__ = "0.1.0"

__all__ = ("Gillespie", "GillespieNet", "Rate", "Batch", "define_system", "RebopError", "RebopValueError", "Stats", "__version__")
