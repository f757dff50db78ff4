"""The reference's import names (MAS_library/__init__.py:1-3, Pk_library/__init__.py:1-4) resolve to
this package's functions after use_dropin(), in a fresh interpreter (no GPU calls)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_import_names():
    code = """
import pylians_amd
pylians_amd.use_dropin()
import MAS_library as MASL, Pk_library as PKL, redshift_space_library as RSL, smoothing_library as SL, readgadget
import pylians_amd.MAS_library as M, pylians_amd.Pk_library as P, pylians_amd.Pk_snapshot as S
assert MASL.MA is M.MA and MASL.CICWc3D is M.CICWc3D and MASL.FLOAT_type is M.FLOAT_type
assert MASL.density_field_gadget.__module__ == 'pylians_amd.MAS_gadget'
assert PKL.Pk is P.Pk and PKL.XPk is P.XPk and PKL.Bk is P.Bk and PKL.FFT3Dr_f is P.FFT3Dr_f
assert PKL.Pk_comp is S.Pk_comp and PKL.Pk_Gadget is S.Pk_Gadget
assert RSL.pos_redshift_space.__module__ == 'pylians_amd.redshift_space_library'
assert SL.FT_filter.__module__ == 'pylians_amd.smoothing_library'
assert readgadget.header.__module__ == 'pylians_amd.readgadget'
assert not hasattr(PKL, 'np') and not hasattr(MASL, 'torch')
print('ok')
"""
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, cwd=ROOT, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-3000:]


def test_dropin_modules_do_not_import_torch():
    """A numpy caller of the drop-in modules never imports torch: device buffers then come from the
    library (pyl_dev_alloc); without a GPU the first compute call fails loudly, still without torch."""
    code = """
import sys
import numpy as np
import pylians_amd
pylians_amd.use_dropin()
import MAS_library as MASL, Pk_library as PKL, redshift_space_library as RSL, smoothing_library as SL, readgadget
import pylians_amd.pipeline, pylians_amd.staging, pylians_amd.Pk_snapshot, pylians_amd.MAS_gadget
assert 'torch' not in sys.modules, 'import pulled in torch'
from pylians_amd import _lib
n = _lib.lib()
c = __import__('ctypes').c_int(-1)
_lib.check(n.pyl_device_count(__import__('ctypes').byref(c)))
if c.value == 0:
    try:
        MASL.MA(np.zeros((4, 3), np.float32), np.zeros((8, 8, 8), np.float32), 1.0, 'CIC')
        raise AssertionError('no error without a GPU')
    except RuntimeError as e:
        assert 'no GPU' in str(e)
assert 'torch' not in sys.modules, 'a call pulled in torch'
print('ok')
"""
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, cwd=ROOT, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-3000:]
