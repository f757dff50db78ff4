"""The numpy drop-in path without torch: a child process that never imports torch runs MA, Pk, XPk,
FFT3Dr/IFFT3Dr, pos_redshift_space and field_smoothing on numpy arrays (device buffers from the
library, pyl_dev_alloc and pyl_memcpy_*), and the results equal those of the same calls made here
with torch imported (torch device buffers).  The child loads one HIP runtime, also after a later
`import torch` (MAS_library.pyx:57-112, Pk_library.pyx:266-425, 534-798)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASE = """
import numpy as np
def inputs():
    rng = np.random.default_rng(11)
    pos = (rng.random((40000, 3)) * 200.0).astype(np.float32)
    vel = rng.normal(0, 300, (40000, 3)).astype(np.float32)
    W = rng.random(40000).astype(np.float32)
    return pos, vel, W
def run(MASL, PKL, RSL, SL):
    pos, vel, W = inputs()
    out = {}
    g = np.zeros((32, 32, 32), np.float32)
    MASL.MA(pos, g, 200.0, 'CIC', W=W)
    out['grid'] = g
    d = (g / np.float32(g.mean()) - np.float32(1.0)).astype(np.float32)
    p = PKL.Pk(d, 200.0, axis=0, MAS='CIC')
    out['pk'] = np.concatenate([p.k3D, p.Pk.ravel(), p.Pkphase, p.Nmodes3D, p.Pk1D, p.Pk2D, p.Nmodes2D])
    x = PKL.XPk([d, d * np.float32(0.5)], 200.0, axis=2, MAS=['CIC', 'CIC'])
    out['xpk'] = np.concatenate([x.Pk.ravel(), x.XPk.ravel(), x.Nmodes3D])
    dk = PKL.FFT3Dr_f(d)
    out['fft'] = np.abs(dk).astype(np.float64)
    out['ifft'] = PKL.IFFT3Dr_f(dk)
    q = pos.copy()
    RSL.pos_redshift_space(q, vel, 200.0, 100.0, 0.5, 1)
    out['rsd'] = q
    fk = SL.FT_filter(200.0, 10.0, 32, 'Top-Hat')
    out['smooth'] = SL.field_smoothing(d, fk)
    return out
"""

CHILD = CASE + """
import sys, ctypes
import pylians_amd
pylians_amd.use_dropin()
import MAS_library as MASL, Pk_library as PKL, redshift_space_library as RSL, smoothing_library as SL
out = run(MASL, PKL, RSL, SL)
assert 'torch' not in sys.modules
np.savez(sys.argv[1], **out)
def runtimes():
    maps = open('/proc/self/maps').read().splitlines()
    return sorted(set(l.split()[-1] for l in maps if 'libamdhip64' in l))
before = runtimes()
import torch
t = torch.ones(1000, device='cuda')
g = np.zeros((16, 16, 16), np.float32)
MASL.MA(np.random.default_rng(1).random((500, 3)).astype(np.float32), g, 1.0, 'CIC')
after = runtimes()
print(__import__('json').dumps({'before': before, 'after': after, 'sum': float(g.sum()), 't': float(t.sum())}))
"""


def test_numpy_path_without_torch_matches_torch_path(gpu, tmp_path):
    env = dict(os.environ, PYTHONPATH=ROOT)
    f = str(tmp_path / "child.npz")
    print("\ntorch-free child: MA, Pk, XPk, FFT, RSD, smoothing on numpy arrays", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, f], capture_output=True, text=True, env=env, cwd=ROOT,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    info = json.loads(r.stdout.strip().splitlines()[-1])
    print("HIP runtimes loaded:", info["before"], "->", info["after"])
    # one runtime before and after `import torch`, and the same one
    assert len(info["before"]) == 1 and info["after"] == info["before"]
    assert info["sum"] == pytest.approx(500.0, rel=1e-5) and info["t"] == 1000.0
    child = np.load(f)
    ns = {}
    exec(CASE, ns)
    import pylians_amd
    from pylians_amd import MAS_library, Pk_library, redshift_space_library, smoothing_library
    here = ns["run"](MAS_library, Pk_library, redshift_space_library, smoothing_library)
    for k, v in here.items():
        c = child[k]
        assert c.shape == v.shape and c.dtype == v.dtype, k
        scale = np.max(np.abs(v)) if v.size else 1.0
        err = np.max(np.abs(c.astype(np.float64) - v.astype(np.float64))) / scale
        print("%-6s max |torch-free - torch| / max = %.2e" % (k, err))
        # the same kernels on the same inputs; float64 atomic order may flip a float32 rounding
        assert err <= 1e-6, k
