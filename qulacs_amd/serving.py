"""Request server over the qm_server_* entry points: circuits of mixed size run
on one device with continuous batching and admission by HBM footprint.

    srv = Server(hbm_budget_bytes=64 << 30)
    rid = srv.submit(n, gates, terms, uniforms)    # from any thread
    energy, samples, latency = srv.wait(rid)
    energy, samples, flags, latency = srv.wait_checked(rid)   # with the boundary check
    srv.close()

`gates` and `terms` use the batch forms of qulacs_amd.batch; every request
starts from |0...0>."""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from .batch import _marshal


class Server:
    def __init__(self, hbm_budget_bytes: int, large_slots: int = 4, small_max_qubits: int = 16, device: int = 0):
        self.h = _lib.call("qm_server_create", int(hbm_budget_bytes), int(large_slots), int(small_max_qubits), device)
        if not self.h:
            raise _lib.QulacsMI355Error("qm_server_create failed")
        self._counts: dict[int, int] = {}
        self._lock = threading.Lock()

    def submit(self, n: int, gates, terms, uniforms) -> int:
        gates = list(gates)
        arr, keep = _marshal(gates)
        terms = list(terms)
        t = max(1, len(terms))
        bf = (C.c_ulonglong * t)(*[int(x[0]) for x in terms])
        pf = (C.c_ulonglong * t)(*[int(x[1]) for x in terms])
        rot = (C.c_uint * t)(*[int(x[2]) for x in terms])
        coef = (C.c_double * t)(*[float(x[3]) for x in terms])
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
        rid = _lib.call("qm_server_submit", self.h, int(n), arr, len(gates), bf, pf, rot, coef, len(terms),
                        u.ctypes.data_as(_lib.PD), u.size)
        if rid < 0:
            raise _lib.QulacsMI355Error("qm_server_submit rejected the request")
        with self._lock:
            self._counts[rid] = u.size
        del keep
        return rid

    def wait(self, rid: int):
        """(energy, samples, latency_seconds) of request rid."""
        e, out, _, lat = self.wait_checked(rid)
        return e, out, lat

    def wait_checked(self, rid: int):
        """(energy, samples, flags, latency_seconds) of request rid.  flags[c] = 1
        marks a sample whose uniform lies within the rounding margin of a CDF
        step (qm_sampling_checked_host); every other sample is the same for any
        evaluation of the circuit that differs from this one by rounding."""
        with self._lock:
            count = self._counts.pop(rid)
        e = C.c_double()
        lat = C.c_double()
        flagged = C.c_uint()
        out = np.zeros(max(1, count), dtype=np.uint64)
        flags = np.zeros(max(1, count), dtype=np.uint8)
        _lib.call("qm_server_wait_checked", self.h, rid, C.byref(e), out.ctypes.data_as(_lib.PI),
                  flags.ctypes.data_as(_lib.PB), C.byref(flagged), C.byref(lat))
        return e.value, out[:count], flags[:count], lat.value

    def stats(self) -> dict:
        v = [C.c_ulonglong() for _ in range(6)]
        _lib.call("qm_server_stats", self.h, *[C.byref(x) for x in v])
        return dict(zip(("admitted", "completed", "failed", "batches", "max_batch", "peak_bytes"),
                        (int(x.value) for x in v)))

    def close(self) -> None:
        if self.h:
            _lib.call("qm_server_destroy", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
