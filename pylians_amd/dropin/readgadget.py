"""`import readgadget` unchanged (readgadget.py), served by pylians_amd."""
from pylians_amd.readgadget import *  # noqa: F401,F403
