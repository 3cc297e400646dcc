"""Continuous-batching serving loop over the C ABI (msdp_server_*).

Requests are prepared `MadSDPSolver` objects (problem data staged on the host,
nothing on the device yet) with an arrival time in seconds after `start()`.
The native loop admits them in submission order as worker slots and device
memory free up, solves each on its slot's HIP stream concurrently with the
others, and releases its device state; the solution (X, S, y, x, z), the
status and the iteration trace stay readable on the solver object.

MadSDP itself solves one problem per call (src/solver.jl:385-396); the loop
has no counterpart in the reference and adds no arithmetic of its own.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, List

import numpy as np

from . import lib as L
from .solver import MadSDPSolver


@dataclass
class RequestResult:
    arrival: float
    admitted: float
    uploaded: float
    done: float
    footprint: int
    slot: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class Server:
    """`Server(slots, mem_budget, device)`: `slots` concurrent solves at most,
    admitted while their summed footprints stay within `mem_budget` bytes."""

    def __init__(self, slots: int = 8, mem_budget: int = 64 << 30, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.msdp_server_create(C.byref(h), device, int(slots), int(mem_budget)),
                "msdp_server_create")
        self._h = h
        self._solvers: Dict[int, MadSDPSolver] = {}
        self.started = False

    def start(self):
        L.check(self.lib.msdp_server_start(self._h), "msdp_server_start")
        self.started = True

    def submit(self, solver: MadSDPSolver, arrival: float = 0.0) -> int:
        rid = C.c_int64()
        L.check(self.lib.msdp_server_submit(self._h, solver._h, float(arrival), C.byref(rid)),
                "msdp_server_submit")
        self._solvers[rid.value] = solver  # keeps the handle alive while in flight
        return rid.value

    def wait_all(self):
        L.check(self.lib.msdp_server_wait_all(self._h), "msdp_server_wait_all")

    def request(self, rid: int) -> RequestResult:
        info = L.RequestInfo()
        L.check(self.lib.msdp_server_request(self._h, rid, C.byref(info)), "msdp_server_request")
        L.check(info.rc, f"request {rid}")
        return RequestResult(info.arrival, info.admitted, info.uploaded, info.done, info.footprint,
                             info.slot)

    def stats(self) -> L.ServerStats:
        st = L.ServerStats()
        L.check(self.lib.msdp_server_get_stats(self._h, C.byref(st)), "msdp_server_get_stats")
        return st

    def close(self):
        if getattr(self, "_h", None):
            self.lib.msdp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve(solvers: List[MadSDPSolver], arrivals, slots: int = 8, mem_budget: int = 64 << 30,
          device: int = 0):
    """Runs a request stream to completion; returns (results, stats) with
    solves/s = completed / (last completion - first arrival) and latency
    percentiles in seconds."""
    srv = Server(slots, mem_budget, device)
    ids = [srv.submit(s, a) for s, a in zip(solvers, arrivals)]
    srv.start()
    srv.wait_all()
    res = [srv.request(i) for i in ids]
    st = srv.stats()
    srv.close()
    lat = np.array([r.latency for r in res])
    service = np.array([r.done - r.admitted for r in res])
    span = max(r.done for r in res) - min(r.arrival for r in res)
    summary = dict(completed=len(res), span_s=span, solves_per_s=len(res) / span,
                   p50_latency_s=float(np.percentile(lat, 50)),
                   p99_latency_s=float(np.percentile(lat, 99)),
                   mean_service_s=float(service.mean()), max_service_s=float(service.max()),
                   mean_upload_s=float(np.mean([r.uploaded - r.admitted for r in res])),
                   peak_active=int(st.peak_active), peak_bytes=int(st.peak_bytes))
    return res, summary
