"""`import smoothing_library as SL` unchanged (smoothing_library.pyx), served by pylians_amd."""
from pylians_amd.smoothing_library import *  # noqa: F401,F403
