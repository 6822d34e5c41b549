"""Host-side mirror of the MI355X hot path of ppf-contact-solver.

The compute lives in libppf_mi355x.so (hand-written HIP for gfx950, C ABI in
include/ppf_mi355x.h); this package only moves torch device buffers across
that ABI and generates synthetic scenes.
"""
from ._lib import PPFError, load  # noqa: F401
