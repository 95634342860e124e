"""The reference's import name for this package: ``import mach`` gives ``mach_amd``.

Code written for the reference (``src/mach/__init__.py:17-31``) keeps working unchanged:
``mach.beamform``, ``mach.kernel`` (``beamform``, ``nb_beamform``, ``InterpolationType``),
``mach.experimental``, ``mach.wavefront``, ``mach.geometry``, ``mach.io.must`` and ``mach.io.uff``, including imports of
the submodules (``import mach.kernel``, ``from mach.experimental import beamform``). Each name is the
``mach_amd`` module itself, registered under the reference's dotted name, not a copy. Install only one
of the two packages named ``mach`` on a path.
"""

import sys

from mach_amd import InterpolationType, __version__, beamform, experimental, geometry, io, kernel, wavefront
from mach_amd.io import must as _must
from mach_amd.io import uff as _uff

# submodules under the reference's dotted names: ``import mach.kernel`` finds them here
for _name, _mod in (("kernel", kernel), ("experimental", experimental), ("wavefront", wavefront),
                    ("geometry", geometry), ("io", io), ("io.must", _must), ("io.uff", _uff)):
    sys.modules[f"{__name__}.{_name}"] = _mod
del _name, _mod

Array = object  # the reference's array-API annotation type (``src/mach/_array_api.py``); any DLPack array

__all__ = ["Array", "InterpolationType", "__version__", "beamform", "experimental", "geometry", "io", "kernel",
           "wavefront"]
