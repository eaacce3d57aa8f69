"""MI355X (gfx950) collocation-BEM hot path of cc_abstemious.

Host mirror of the reference's influence_matrix / solve API over the C ABI of
libccab.so (include/ccab.h).  The HIP library is required: importing the
solver modules raises if it has not been built.
"""
from . import capi  # noqa: F401

__all__ = ["capi", "mesh", "preprocess", "influence_matrix", "incident_wave", "solve"]
