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


class _Worker:
    """Runs one server on its own host thread: drains its submissions,
    steps it while it has work and hands finished lines back, so servers
    of different r progress independently (a long gcd phase of one does
    not hold up the others)."""

    def __init__(self, server: Server, rounds: int):
        import queue
        import threading
        self.server = server
        self.rounds = rounds
        self.inbox: "queue.Queue" = queue.Queue()
        self.outbox: "queue.Queue" = queue.Queue()
        self.error = None
        self._stop = threading.Event()
        self._wake = threading.Event()
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()

    def submit(self, job: int, svals, k0: int, maxd: int) -> None:
        self.inbox.put((job, svals, k0, maxd))
        self._wake.set()

    def _run(self) -> None:
        import queue
        try:
            while not self._stop.is_set():
                while True:
                    try:
                        job, svals, k0, maxd = self.inbox.get_nowait()
                    except queue.Empty:
                        break
                    self.server.submit(job, svals, k0, maxd)
                    self.server.fresh = True
                if self.server.last_live or self.server.last_queued or self.server.fresh:
                    self.server.step(self.rounds)
                    for item in self.server.poll():
                        self.outbox.put(item)
                else:
                    self._wake.wait(0.01)
                    self._wake.clear()
        except Exception as e:  # reported to the caller by RequestLoop.pump
            self.error = e

    def close(self) -> None:
        self._stop.set()
        self._wake.set()
        self._thread.join()
        self.server.close()


class RequestLoop:
    """Admits jobs as they arrive, sieves each new batch of one r in one
    pass and feeds the survivors to that r's server; every server runs on
    its own worker thread, so the work of jobs that overlap in time
    overlaps on the device."""

    def __init__(self, mem_fraction: float = 0.6, block_m: int = 0, gcd_threads: int = 0,
                 L0: int = 64, Lmax: int = 4096, rounds: int = 1):
        from .device import mem_info
        init()
        self.free0 = mem_info()[0]
        self.mem_fraction = mem_fraction
        self.block_m = block_m
        self.gcd_threads = gcd_threads
        self.L0, self.Lmax = L0, Lmax
        self.rounds = rounds
        self.workers: dict[int, _Worker] = {}
        self.jobs: dict[int, Job] = {}
        self.waiting: list[Job] = []

    @property
    def servers(self) -> dict[int, Server]:
        return {r: w.server for r, w in self.workers.items()}

    def server(self, r: int) -> Server:
        if r not in self.workers:
            budget = int(self.free0 * self.mem_fraction / 4)   # up to four r at once
            sv = Server(r, mem_bytes=budget, block_m=self.block_m,
                        gcd_threads=self.gcd_threads, L0=self.L0, Lmax=self.Lmax)
            self.workers[r] = _Worker(sv, self.rounds)
        return self.workers[r].server

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
            self.server(r)
            for j in js:
                seg = best[j.s_lo: j.s_hi + 1]
                surv = []
                for k, s in enumerate(range(j.s_lo, j.s_hi + 1)):
                    key = int(seg[k])
                    if key != 0xFFFFFFFFFFFFFFFF and (key >> 48) <= j.depth:
                        j.lines[s] = f"{s} {key >> 48} p{key & 0xFFFFFFFFFFFF:x}"
                    else:
                        surv.append(s)
                j.pending = len(surv)
                if surv:
                    self.workers[r].submit(j.job_id, surv, j.depth, j.maxd)
                else:
                    j.done_at = time.perf_counter()

    def busy(self) -> bool:
        return bool(self.waiting) or any(j.pending for j in self.jobs.values())

    def pump(self, rounds: int = 1) -> list[Job]:
        """Sieve the newly arrived jobs and collect finished lines; returns
        the jobs that completed.  (`rounds` is kept for callers; each
        worker steps on its own.)"""
        import queue
        if self.waiting:
            self._sieve_waiting()
        finished = []
        got_any = False
        for w in self.workers.values():
            if w.error is not None:
                raise w.error
            while True:
                try:
                    job_id, line = w.outbox.get_nowait()
                except queue.Empty:
                    break
                got_any = True
                j = self.jobs[job_id]
                j.lines[int(line.split()[0])] = line
                j.pending -= 1
                if j.pending == 0:
                    j.done_at = time.perf_counter()
                    finished.append(j)
        if not got_any and not finished:
            time.sleep(0.002)
        return finished

    def close(self) -> None:
        for w in self.workers.values():
            w.close()
        self.workers = {}
