"""Nodal DC resistivity on the GPU (``simpeg.electromagnetics.static.resistivity`` surface)."""

from . import receivers, sources  # noqa: F401
from .fields import Fields3DNodal  # noqa: F401
from .simulation import Simulation3DNodal  # noqa: F401
from .survey import Survey  # noqa: F401
from . import survey  # noqa: F401
from . import simulation  # noqa: F401
