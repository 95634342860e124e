"""mach_amd: MI355X-native (gfx950) delay-and-sum ultrasound beamforming.

Drop-in for the reference package ``mach`` (``src/mach/__init__.py:10-31``): ``beamform``,
``kernel``, ``experimental``, ``wavefront`` and ``geometry`` keep the reference's names and
behaviour; the compute runs in hand-written HIP kernels behind the C ABI of ``include/mach_hip.h``.
``import mach_amd as mach`` is the intended use. Beyond the reference: ``preprocess`` (RF -> IQ
analytic signal on the device), ``postprocess`` (power
Doppler and B-mode images computed in the kernel's epilogue), ``transmit`` (transmit arrivals evaluated
on the device from wavefront descriptors), ``batch`` (many requests per persistent launch), ``serving``
(continuous batching), ``autotune`` and ``phantoms`` (synthetic workloads).
"""

__version__ = "0.1.3"

from . import geometry, probes, wavefront  # noqa: E402  (host-only helpers, no native code)
from . import experimental, kernel, postprocess, preprocess, transmit  # noqa: E402  (native library; fails loudly)
from .kernel import InterpolationType, beamform  # noqa: E402

__all__ = [
    "InterpolationType",
    "__version__",
    "beamform",
    "experimental",
    "geometry",
    "kernel",
    "postprocess",
    "preprocess",
    "probes",
    "transmit",
    "wavefront",
]
