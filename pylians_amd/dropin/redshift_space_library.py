"""`import redshift_space_library as RSL` unchanged (redshift_space_library.pyx), served by pylians_amd."""
from pylians_amd.redshift_space_library import *  # noqa: F401,F403
