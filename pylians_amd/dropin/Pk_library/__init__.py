"""`import Pk_library as PKL` unchanged: the names of the reference package (Pk_library/__init__.py:1-4:
Pk_library, Pk_snapshot and bispectrum_library), served by pylians_amd."""
from pylians_amd.Pk_library import *  # noqa: F401,F403
from pylians_amd.Pk_snapshot import *  # noqa: F401,F403
