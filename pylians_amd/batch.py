"""Batched particles -> P(k) for many independent realizations of uneven sizes.

``PkBatch(jobs)`` takes a list of dicts ``{pos, BoxSize, dims, MAS='CIC', axis=0, W=None,
precision='fp32'}`` with positions already on the GPU (or numpy arrays, uploaded once), runs the
whole batch once inside workspaces the graph owns, captures it into a single hipGraph and replays
it with ``run()`` on the batch's own stream; ``results()`` returns one ``Pk``-like object per job,
after the mode-count check of every job (check_number_modes, Pk_library.pyx:90-102, on the counts
of the binning kernel).  Eager calls between replays do not disturb the graph (its workspaces are
its own).  Without a graph, ``run()`` queues the jobs on the current torch stream, in the shared
workspaces like any other call.  Builder-defined: the reference computes one field per call
(Pk/Pk.py:17-27).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .Pk_library import check_number_modes, pk_layout, unpack_pk

_PREC = {"fp32": _lib.F32, "fp64": _lib.F64}


class _Res:
    pass


class PkBatch:
    def __init__(self, jobs, flags=_lib.PK_ALL, graph=True):
        _lib.require_gpu()
        self._keep = []
        self._outs = []
        arr = (_lib.PkJob * len(jobs))()
        for i, j in enumerate(jobs):
            p = _lib.to_dev(j["pos"], torch.float32)
            w = None if j.get("W") is None else _lib.to_dev(j["W"], torch.float32)
            dims = int(j["dims"])
            out = torch.empty(pk_layout(dims)[3], dtype=torch.float64, device="cuda")
            self._keep += [p, w]
            self._outs.append((out, dims))
            arr[i] = _lib.PkJob(p.data_ptr(), 0 if w is None else w.data_ptr(), int(p.shape[0]), dims,
                                float(j["BoxSize"]), _lib.SCHEMES[j.get("MAS", "CIC")], int(j.get("axis", 0)),
                                _PREC[j.get("precision", "fp32")], int(flags), out.data_ptr())
        self._jobs = arr
        self.n_particles = sum(int(j["pos"].shape[0]) for j in jobs)
        self.stream = torch.cuda.Stream()
        self._graph = None
        if graph:
            h = ctypes.c_void_p()
            torch.cuda.synchronize()
            _lib.check(_lib.lib().pyl_graph_create(arr, len(jobs), ctypes.c_void_p(self.stream.cuda_stream),
                                                   ctypes.byref(h)))
            self._graph = h

    def run(self):
        """Queue the whole batch: one graph launch on this batch's stream, or (without a graph) the
        jobs one by one on the current stream."""
        if self._graph is not None:
            s = ctypes.c_void_p(self.stream.cuda_stream)
            _lib.check(_lib.lib().pyl_graph_launch(self._graph, s))
        else:
            self.stream.wait_stream(torch.cuda.current_stream())
            _lib.check(_lib.lib().pyl_pk_batch(self._jobs, len(self._jobs), _lib.stream()))
            self.stream.wait_stream(torch.cuda.current_stream())

    def results(self):
        self.stream.synchronize()
        res = []
        for out, dims in self._outs:
            r = _Res()
            unpack_pk(r, _lib.to_host(out), dims)
            # the DC mode is counted by the kernel and dropped from the output (Pk_library.pyx:411-412)
            check_number_modes(np.concatenate([[1.0], r.Nmodes3D]), dims)
            res.append(r)
        return res

    def close(self):
        if self._graph is not None:
            self.stream.synchronize()
            _lib.check(_lib.lib().pyl_graph_destroy(self._graph))
            self._graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
