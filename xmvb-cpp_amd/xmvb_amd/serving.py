"""Serving: continuous batching of E+grad and full-convergence requests over resident problems.

No reference counterpart: the reference evaluates one problem per process (`cli/run.cpp`), so this
module is the engine's own request loop. Requests name a registered problem. Problems that share
a structure topology (same active space and structure set, any n_bf / n_aux) live in one pool.

E+grad requests go to the engine's native scheduler (`xmvb_amd.Scheduler`, vbe_server_*): its
dispatcher thread batches every queued request of a pool whose slot is not in flight into one
launch as soon as one of the pool's launch lanes is free, on the context's shared streams (at most
four), and stamps each request with its queue, device and host times. This module only keeps the
residency: a problem becomes resident (its S, h, B and workspace uploaded into a pool slot) when a
request for it arrives, provided the pool has a free slot and the context's HBM budget has room
(`Context(hbm_budget_bytes=...)`; the engine refuses an allocation over it). Otherwise the least
recently used resident problem with no queued or running request is unloaded to make room; if there
is none, the request waits here until completions free slots or memory.
"""
from __future__ import annotations

import collections
import threading
import time

import numpy as np

from . import Pool, Scheduler, VbeError

VBE_ERR_MEMORY = 6


class Request:
    """One request; `energy`, `grad` (E+grad) or `result` (convergence) are set on completion, with
    `status` (0 on success) and, for E+grad requests, `queue_ms` (arrival to launch), `device_ms`
    (its launch on the device) and `batch` (requests in that launch)."""

    __slots__ = ("id", "key", "C", "gradient", "arrival", "done", "energy", "grad", "result", "options", "status",
                 "queue_ms", "device_ms", "batch")

    def __init__(self, rid, key, C, gradient, arrival, options=None):
        self.id, self.key, self.C, self.gradient, self.arrival = rid, key, C, gradient, arrival
        self.done = None
        self.energy = None
        self.grad = None
        self.result = None
        self.options = options or {}
        self.status = None
        self.queue_ms = self.device_ms = None
        self.batch = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _Problem:
    __slots__ = ("key", "topo", "loader", "e_nuc", "home", "slot", "pending", "last_used")

    def __init__(self, key, topo, loader, e_nuc):
        self.key, self.topo, self.loader, self.e_nuc = key, topo, loader, e_nuc
        self.home = None      # _PoolState of the problem's topology
        self.slot = None      # resident slot, or None
        self.pending = 0      # queued + running requests
        self.last_used = 0.0


class _PoolState:
    def __init__(self, ctx, topo, capacity):
        self.pool = Pool(ctx, topo, capacity)
        self.free_slots = list(range(capacity - 1, -1, -1))


class _Residency:
    """Problems, their pools and which of them are resident (one instance per host thread)."""

    def __init__(self, ctx, capacity: int):
        self.ctx, self.capacity = ctx, capacity
        self.pools = {}       # id(topology) -> _PoolState
        self.problems = {}
        self.loads = 0
        self.evictions = 0

    def register(self, key, topology, loader, e_nuc=0.0):
        """loader() -> (n_bf, n_aux, S, h, B): the problem's static data, called at admission."""
        if key in self.problems:
            raise ValueError(f"problem {key!r} is already registered")
        p = _Problem(key, topology, loader, e_nuc)
        ps = self.pools.get(id(topology))
        if ps is None:
            ps = self.pools[id(topology)] = _PoolState(self.ctx, topology, self.capacity)
        p.home = ps
        self.problems[key] = p

    def _victim(self, home=None, exclude=None):
        best = None
        for p in self.problems.values():
            if p.slot is None or p.pending or p is exclude or (home is not None and p.home is not home):
                continue
            if best is None or p.last_used < best.last_used:
                best = p
        return best

    def _evict(self, p):
        p.home.pool.unload(p.slot)
        p.home.free_slots.append(p.slot)
        p.slot = None
        self.evictions += 1

    def admit(self, p) -> bool:
        """Make p resident; False if neither a slot nor the memory can be freed now."""
        if p.slot is not None:
            return True
        ps = p.home
        if not ps.free_slots:
            v = self._victim(home=ps, exclude=p)
            if v is None:
                return False
            self._evict(v)
        slot = ps.free_slots.pop()
        n_bf, n_aux, S, h, B = p.loader()
        while True:
            try:
                ps.pool.load(slot, n_bf, n_aux, S, h, B, p.e_nuc)
                break
            except VbeError as exc:
                if exc.status != VBE_ERR_MEMORY:
                    ps.free_slots.append(slot)
                    raise
                v = self._victim(exclude=p)
                if v is None:
                    ps.free_slots.append(slot)
                    return False
                self._evict(v)
        p.slot = slot
        p.last_used = time.perf_counter()
        self.loads += 1
        return True

    def resident(self) -> int:
        return sum(1 for p in self.problems.values() if p.slot is not None)

    def close(self):
        for ps in self.pools.values():
            ps.pool.close()
        self.pools.clear()


class Server:
    """E+grad requests over registered problems, batched by the native scheduler.

    Call `submit` and `pump` from one thread. `submit` admits the problem if needed and hands the
    request to the scheduler at once (or holds it until admission is possible); `pump` returns the
    requests completed since the previous call (waiting up to wait_us for one) and retries held
    admissions."""

    def __init__(self, ctx, capacity: int = 256, max_batch: int = 64, max_inflight_per_pool: int = 4):
        self.res = _Residency(ctx, capacity)
        self.sched = Scheduler(ctx, max_batch=max_batch, max_inflight_per_pool=max_inflight_per_pool)
        self.held = collections.deque()  # waiting for admission
        self.inflight = {}               # scheduler id -> request
        self._next_id = 0

    def register(self, key, topology, loader, e_nuc=0.0):
        self.res.register(key, topology, loader, e_nuc)

    def _send(self, r) -> bool:
        p = self.res.problems[r.key]
        if p.slot is None and not self.res.admit(p):
            return False
        ps = p.home
        grad = np.empty((ps.pool.n_act, ps.pool.n_bf[p.slot])) if r.gradient else None
        sid = self.sched.submit(ps.pool, p.slot, r.C, grad, arrival=r.arrival)
        self.inflight[sid] = r
        return True

    def submit(self, key, C, gradient: bool = True, arrival: float | None = None) -> Request:
        p = self.res.problems[key]
        r = Request(self._next_id, key, C, gradient, time.perf_counter() if arrival is None else arrival)
        self._next_id += 1
        p.pending += 1
        if self.held or not self._send(r):
            self.held.append(r)
        return r

    def busy(self) -> bool:
        return bool(self.held) or bool(self.inflight)

    def pump(self, wait_us: int = 0) -> list:
        done = []
        for info in self.sched.collect(wait_us if self.inflight else 0):
            r = self.inflight.pop(info["id"])
            p = self.res.problems[r.key]
            r.status = info["status"]
            r.energy = info["energy"]
            r.grad = info["grad"]
            r.done = info["t_complete"]
            r.queue_ms = (info["t_dispatch"] - r.arrival) * 1e3
            r.device_ms = info["device_ms"]
            r.batch = info["batch_size"]
            p.pending -= 1
            p.last_used = r.done
            done.append(r)
        while self.held and self._send(self.held[0]):
            self.held.popleft()
        return done

    def drain(self) -> list:
        out = []
        while self.busy():
            out += self.pump(wait_us=2000)
        return out

    @property
    def launches(self) -> int:
        return self.sched.counters()["launches"]

    @property
    def loads(self) -> int:
        return self.res.loads

    @property
    def evictions(self) -> int:
        return self.res.evictions

    def close(self):
        self.sched.close()
        self.res.close()


class ConvergenceWorker:
    """Full-convergence requests (device P-LBFGS) on a host thread of their own: the requests that
    have arrived are converged together in lockstep, one launch per optimizer phase for all of them
    (one pool per topology, one request per problem per round). Problems are admitted into this
    worker's own pools under the same context budget."""

    def __init__(self, ctx, capacity: int = 64, max_iterations: int = 150, **options):
        self.res = _Residency(ctx, capacity)
        self.options = dict(options, max_iterations=max_iterations)
        self._lock = threading.Lock()
        self._pending = collections.deque()
        self._wake = threading.Event()
        self._stop = False
        self._thread = None
        self.completed = []
        self.rounds = []
        self.error = None
        self._next_id = 0

    def register(self, key, topology, loader, e_nuc=0.0):
        self.res.register(key, topology, loader, e_nuc)

    def submit(self, key, C, arrival: float | None = None, **options) -> Request:
        with self._lock:
            r = Request(self._next_id, key, C, True, time.perf_counter() if arrival is None else arrival, options)
            self._next_id += 1
            self.res.problems[key].pending += 1
            self._pending.append(r)
        self._wake.set()
        return r

    def preload(self, keys=None):
        """Admit problems ahead of their requests (all registered ones by default), so that pool
        creation and factor uploads do not fall inside a serving window."""
        with self._lock:
            for key in (list(self.res.problems) if keys is None else keys):
                self.res.admit(self.res.problems[key])

    def start(self):
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()

    def finish(self):
        """Wait until every submitted request has completed, then stop the thread."""
        self._stop = True
        self._wake.set()
        if self._thread is not None:
            self._thread.join()
        if self.error is not None:
            raise self.error

    def _run(self):
        try:
            while True:
                with self._lock:
                    work = list(self._pending)
                    self._pending.clear()
                if not work:
                    if self._stop:
                        return
                    self._wake.wait(0.005)
                    self._wake.clear()
                    continue
                later = self._round(work)
                with self._lock:
                    self._pending.extendleft(reversed(later))
        except BaseException as exc:  # re-raised by finish()
            self.error = exc

    def _round(self, work):
        later, groups, seen = [], {}, set()
        for r in work:
            p = self.res.problems[r.key]
            if r.key in seen:
                later.append(r)
                continue
            with self._lock:  # pending counts change under the lock, so eviction sees them settled
                admitted = self.res.admit(p)
            if not admitted:
                later.append(r)
                continue
            seen.add(r.key)
            groups.setdefault(id(p.home), (p.home, []))[1].append(r)
        if not groups and later:
            time.sleep(0.002)  # nothing admissible now: memory is held elsewhere
        for ps, reqs in groups.values():
            opts = dict(self.options)
            opts.update(reqs[0].options)
            res = ps.pool.optimize([self.res.problems[r.key].slot for r in reqs], [r.C for r in reqs], **opts)
            now = time.perf_counter()
            with self._lock:
                for r, out in zip(reqs, res):
                    p = self.res.problems[r.key]
                    r.result, r.energy, r.done = out, out["energy"], now
                    p.pending -= 1
                    p.last_used = now
                    self.completed.append(r)
            self.rounds.append(len(reqs))
        return later

    def close(self):
        self.res.close()
