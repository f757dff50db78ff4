"""Persistence of the deposit autotuner's choices (see pyl_tune in include/pylians_hip.h).

    from pylians_amd import tuning
    tuning.load()        # $PYLIANS_AMD_TUNE_CACHE or ~/.cache/pylians_amd/tune.json, if present
    ...                  # first calls per workload shape time the binned, atomic and presorted deposits
    tuning.save()
"""
import ctypes
import json
import os

from . import _lib


_KINDS = ["binned", "atomic", "presorted"]   # choice codes 0, 1, 2 of pyl_tune_export


def default_path():
    return os.environ.get("PYLIANS_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "pylians_amd", "tune.json"))


def export_text():
    n = _lib.lib().pyl_tune_export(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    _lib.lib().pyl_tune_export(buf, n + 1)
    return buf.value.decode()


def entries():
    out = []
    for line in export_text().splitlines():
        d, s, p, w, nl, c = (int(v) for v in line.split())
        out.append({"dims": d, "scheme": s, "prec": p, "weights": w, "log2_n": nl,
                    "deposit": _KINDS[c]})
    return out


def save(path=None):
    path = path or default_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(entries(), f, indent=1)
    return path


def load(path=None):
    path = path or default_path()
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        ent = json.load(f)
    text = "".join("%d %d %d %d %d %d\n" % (e["dims"], e["scheme"], e["prec"], e["weights"], e["log2_n"],
                                            _KINDS.index(e["deposit"])) for e in ent)
    _lib.check(_lib.lib().pyl_tune_import(text.encode()))
    return len(ent)


def enable(on=True):
    _lib.check(_lib.lib().pyl_tune(1 if on else 0))
