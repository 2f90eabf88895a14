"""Serving: continuous-batching deep scan (gf2t_server_*) and a request
loop that takes (r, s-range, bounds) jobs of mixed sizes.

A job's s range is sieved on the device (jobs of one r that are waiting
together share one sieve pass), its survivors go to that r's server, and
the job completes when every s has its line.  Jobs are admitted to a
server while its slots (sized from an HBM budget) are free; the rest
queue.  Each s gets the same line alone or in any mix.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from ._capi import ScanParams, check, lib
from .certs import sieve_lines
from .device import init
from .engine import run_gpu


class Server:
    """Continuous-batching deep-scan server for one r."""

    def __init__(self, r: int, mem_bytes: int = 0, block_m: int = 0, L0: int = 64,
                 Lmax: int = 4096, gcd_threads: int = 0):
        init()
        self.r = r
        P = ScanParams(k0=1, maxd=0, L0=L0, Lmax=Lmax, gcd_threads=gcd_threads, batch=0,
                       verbose=0, block_m=block_m)
        h = C.c_void_p()
        check(lib.gf2t_server_create(r, C.byref(P), mem_bytes, C.byref(h)), "gf2t_server_create")
        self._h = h
        self.last_live = self.last_queued = 0
        self.fresh = False
        slots, per = C.c_size_t(), C.c_size_t()
        check(lib.gf2t_server_capacity(h, C.byref(slots), C.byref(per)), "gf2t_server_capacity")
        self.slots, self.bytes_per_slot = slots.value, per.value

    def submit(self, job: int, svals, k0: int, maxd: int = 0) -> None:
        sv = np.ascontiguousarray(np.asarray(svals, dtype=np.uint64))
        check(lib.gf2t_server_submit(self._h, job, sv.ctypes.data_as(C.POINTER(C.c_uint64)),
                                     sv.size, k0, maxd), "gf2t_server_submit")
        self.fresh = True

    def step(self, rounds: int = 1) -> tuple[int, int]:
        live, queued = C.c_size_t(), C.c_size_t()
        check(lib.gf2t_server_step(self._h, rounds, C.byref(live), C.byref(queued)),
              "gf2t_server_step")
        self.last_live, self.last_queued = live.value, queued.value
        self.fresh = False
        return live.value, queued.value

    def poll(self) -> list[tuple[int, str]]:
        out = []
        n = C.c_size_t()
        jobs = (C.c_uint64 * 4096)()
        buf = C.create_string_buffer(1 << 20)
        while True:
            check(lib.gf2t_server_poll(self._h, 4096, jobs, buf, len(buf), C.byref(n)),
                  "gf2t_server_poll")
            if n.value == 0:
                return out
            lines = buf.value.decode().splitlines()
            out.extend((int(jobs[i]), lines[i]) for i in range(n.value))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.gf2t_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


@dataclass
class Job:
    job_id: int
    r: int
    s_lo: int
    s_hi: int
    depth: int = 32          # sieve bound
    maxd: int = 2048         # deep bound (degrees)
    arrival: float = 0.0
    lines: dict = field(default_factory=dict)
    pending: int = 0
    done_at: float = 0.0

    def certificate(self) -> list[str]:
        return [self.lines[s] for s in range(self.s_lo, self.s_hi + 1)]


class RequestLoop:
    """Admits jobs as they arrive, sieves each new batch of one r in one
    pass, feeds survivors to the per-r servers and steps every server with
    work, so the work of jobs that overlap in time overlaps on the device."""

    def __init__(self, mem_fraction: float = 0.6, block_m: int = 0, gcd_threads: int = 0,
                 L0: int = 64, Lmax: int = 4096):
        from .device import mem_info
        init()
        self.free0 = mem_info()[0]
        self.mem_fraction = mem_fraction
        self.block_m = block_m
        self.gcd_threads = gcd_threads
        self.L0, self.Lmax = L0, Lmax
        self.servers: dict[int, Server] = {}
        self.jobs: dict[int, Job] = {}
        self.waiting: list[Job] = []

    def server(self, r: int) -> Server:
        if r not in self.servers:
            budget = int(self.free0 * self.mem_fraction / 4)   # up to four r at once
            self.servers[r] = Server(r, mem_bytes=budget, block_m=self.block_m,
                                     gcd_threads=self.gcd_threads, L0=self.L0, Lmax=self.Lmax)
        return self.servers[r]

    def submit(self, job: Job) -> None:
        self.jobs[job.job_id] = job
        self.waiting.append(job)

    def _sieve_waiting(self) -> None:
        by_r: dict[int, list[Job]] = {}
        for j in self.waiting:
            by_r.setdefault(j.r, []).append(j)
        self.waiting = []
        for r, js in by_r.items():
            smax = max(j.s_hi for j in js)
            depth = max(j.depth for j in js)
            best, _ = run_gpu(r, depth, smax)
            found = {}
            for ln in sieve_lines(best[: smax + 1]):
                s = int(ln.split()[0], 10)
                found[s] = ln
            for j in js:
                surv = []
                for s in range(j.s_lo, j.s_hi + 1):
                    if s in found and int(found[s].split()[1]) <= j.depth:
                        j.lines[s] = found[s]
                    else:
                        surv.append(s)
                j.pending = len(surv)
                if surv:
                    self.server(r).submit(j.job_id, surv, k0=j.depth, maxd=j.maxd)
                else:
                    j.done_at = time.perf_counter()

    def busy(self) -> bool:
        return bool(self.waiting) or any(j.pending for j in self.jobs.values())

    def pump(self, rounds: int = 1) -> list[Job]:
        """One scheduling pass: sieve the newly arrived jobs, then step every
        server with work on its own host thread (servers of different r
        share the device; each one's gcds run on its own thread pool).
        Returns the jobs that completed in it."""
        if self.waiting:
            self._sieve_waiting()
        finished = []
        busy = [sv for sv in self.servers.values() if sv.last_live or sv.last_queued or sv.fresh]
        if len(busy) > 1:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(max_workers=len(busy)) as ex:
                list(ex.map(lambda sv: sv.step(rounds), busy))
        elif busy:
            busy[0].step(rounds)
        for sv in self.servers.values():
            for job_id, line in sv.poll():
                j = self.jobs[job_id]
                j.lines[int(line.split()[0])] = line
                j.pending -= 1
                if j.pending == 0:
                    j.done_at = time.perf_counter()
                    finished.append(j)
        return finished

    def close(self) -> None:
        for sv in self.servers.values():
            sv.close()
        self.servers = {}
