"""fastoverlap_amd -- MI355X-native (gfx950) overlap-maximisation alignment.

Drop-in replacements for fastoverlap's ``SphericalAlign``, ``PeriodicAlign``
and ``SOFT`` whose numerics run in hand-written HIP kernels
(``libfastoverlap_hip.so``), plus batched and served interfaces.
"""

from ._lib import FastOverlapError, Context, default_context
from .batch import (SphericalBatch, PeriodicBatch, align_spherical_batch,
                    align_periodic_batch, compact_to_dense, dense_to_compact, kcount,
                    next_fast_len, periodic_defaults)
from .soft import SOFT
from .spherical import (SphericalAlign, SphericalHarmonicAlign, SphericalHarmonicAlignFortran,
                        BaseSphericalAlignment)
from .periodic import PeriodicAlign, PeriodicAlignFortran, BasePeriodicAlignment

__all__ = ["SphericalAlign", "SphericalHarmonicAlign", "SphericalHarmonicAlignFortran",
           "PeriodicAlign", "PeriodicAlignFortran",
           "SOFT",
           "SphericalBatch", "PeriodicBatch", "align_spherical_batch", "align_periodic_batch", "FastOverlapError", "Context",
           "default_context", "compact_to_dense", "dense_to_compact", "kcount",
           "next_fast_len", "periodic_defaults", "BaseSphericalAlignment",
           "BasePeriodicAlignment"]
