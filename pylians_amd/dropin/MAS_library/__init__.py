"""`import MAS_library as MASL` unchanged: the names of the reference package (MAS_library/__init__.py:1-3:
the MAS_library module and MAS_gadget), served by pylians_amd."""
from pylians_amd.MAS_gadget import *  # noqa: F401,F403
from pylians_amd.MAS_library import *  # noqa: F401,F403
