"""MI355X (gfx950) implementation of Pylians' density-assignment -> P(k) / B(k) path.

The drop-in modules are ``pylians_amd.MAS_library``, ``pylians_amd.Pk_library`` (Pk, XPk, Bk, ...),
``pylians_amd.redshift_space_library``, ``pylians_amd.smoothing_library``, ``pylians_amd.readgadget``,
``pylians_amd.MAS_gadget`` and ``pylians_amd.Pk_snapshot``.  ``use_dropin()`` puts the directory
``pylians_amd/dropin`` first on ``sys.path``, so that code written for the reference --
``import MAS_library as MASL``, ``import Pk_library as PKL`` -- runs unchanged.
"""
import os as _os

DROPIN = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "dropin")


def use_dropin():
    """Serve the reference's import names (MAS_library, Pk_library, redshift_space_library,
    smoothing_library, readgadget) from this package."""
    import sys
    if DROPIN not in sys.path:
        sys.path.insert(0, DROPIN)
