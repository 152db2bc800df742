"""ctlab_amd: MI355X-native (gfx950) branchless distance-driven cone-beam
projector / backprojector with CTLAB's C entry points, device-resident
reconstruction updates and a learned proximal step.

The compute path is libctlab_mi355x.so (hand-written HIP); this package is
the host-side mirror of the reference interface.  Importing it does not load
the library; the first call does, and raises if it is not built.
"""
from .geometry import ConeGeometry, arc_detector, circular_cone, fan_multislice, flat_detector  # noqa: F401

__all__ = ["ConeGeometry", "arc_detector", "circular_cone", "fan_multislice", "flat_detector"]
