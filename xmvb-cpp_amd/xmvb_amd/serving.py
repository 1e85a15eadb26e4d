"""Serving: continuous batching of E+grad and full-convergence requests over resident problems.

No reference counterpart: the reference evaluates one problem per process (`cli/run.cpp`), so this
module is the engine's own request loop. Requests name a registered problem. Problems that share
a structure topology (same active space and structure set, any n_bf / n_aux) live in one pool;
each pool takes up to two submissions at once (its two streams), and pools of different
topologies run concurrently on the device.

Admission: a problem becomes resident (its S, h, B and workspace uploaded into a pool slot) when a
request for it is first dispatched, provided the pool has a free slot and the context's HBM budget
has room (`Context(hbm_budget_bytes=...)`; the engine refuses an allocation over it). Otherwise the
least recently used resident problem with no queued or running request is unloaded to make room;
if there is none, the request waits until completions free slots or memory. Requests that arrive
while others run join the next launch of their pool (continuous batching), so a result does not
depend on which requests it was batched with beyond rounding.
"""
from __future__ import annotations

import collections
import threading
import time

from . import Pool, VbeError

VBE_ERR_MEMORY = 6


class Request:
    """One request; `energy`, `grad` (E+grad) or `result` (convergence) are set on completion."""

    __slots__ = ("id", "key", "C", "gradient", "arrival", "done", "energy", "grad", "result", "options")

    def __init__(self, rid, key, C, gradient, arrival, options=None):
        self.id, self.key, self.C, self.gradient, self.arrival = rid, key, C, gradient, arrival
        self.done = None
        self.energy = None
        self.grad = None
        self.result = None
        self.options = options or {}

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
        self.inflight = {}    # ticket -> list of requests
        self.running = set()  # slots held by in-flight submissions


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
    """E+grad requests, continuously batched per pool (at most two submissions per pool in flight).

    Single-threaded: call `submit` and `pump` from one thread. `pump` collects finished launches,
    admits queued requests' problems and starts new launches; it returns the requests completed
    since the previous call."""

    def __init__(self, ctx, capacity: int = 256, max_batch: int = 64):
        self.res = _Residency(ctx, capacity)
        self.max_batch = max_batch
        self.queue = collections.deque()
        self.launches = 0
        self.batch_sizes = []
        self._next_id = 0

    def register(self, key, topology, loader, e_nuc=0.0):
        self.res.register(key, topology, loader, e_nuc)

    def submit(self, key, C, gradient: bool = True, arrival: float | None = None) -> Request:
        p = self.res.problems[key]
        r = Request(self._next_id, key, C, gradient, time.perf_counter() if arrival is None else arrival)
        self._next_id += 1
        p.pending += 1
        self.queue.append(r)
        return r

    def busy(self) -> bool:
        return bool(self.queue) or any(ps.inflight for ps in self.res.pools.values())

    def pump(self) -> list:
        done = []
        for ps in self.res.pools.values():
            for ticket in list(ps.inflight):
                if not ps.pool.poll(ticket):
                    continue
                e, grads = ps.pool.wait(ticket)
                now = time.perf_counter()
                for i, r in enumerate(ps.inflight.pop(ticket)):
                    p = self.res.problems[r.key]
                    r.energy = float(e[i])
                    r.grad = grads[i] if grads is not None else None
                    r.done = now
                    p.pending -= 1
                    p.last_used = now
                    ps.running.discard(p.slot)
                    done.append(r)
        if self.queue:
            self._dispatch()
        return done

    def _dispatch(self):
        batches = {}   # id(pool state) -> (pool state, requests, gradient flag)
        rest = collections.deque()
        for r in self.queue:
            p = self.res.problems[r.key]
            ps = p.home
            b = batches.get(id(ps))
            if len(ps.inflight) >= 2 or (b and len(b[1]) >= self.max_batch):
                rest.append(r)
                continue
            if b and b[2] != r.gradient:  # one gradient flag per launch; the other kind waits
                rest.append(r)
                continue
            if p.slot is None and not self.res.admit(p):
                rest.append(r)
                continue
            if p.slot in ps.running or (b and any(q.key == r.key for q in b[1])):
                rest.append(r)  # the problem's slot is busy in this or a running launch
                continue
            if b is None:
                b = batches[id(ps)] = (ps, [], r.gradient)
            b[1].append(r)
        for ps, reqs, gradient in batches.values():
            slots = [self.res.problems[r.key].slot for r in reqs]
            ticket = ps.pool.submit(slots, [r.C for r in reqs], gradient=gradient)
            ps.inflight[ticket] = reqs
            ps.running.update(slots)
            self.launches += 1
            self.batch_sizes.append(len(reqs))
        self.queue = rest

    def drain(self) -> list:
        out = []
        while self.busy():
            out += self.pump()
        return out

    @property
    def loads(self) -> int:
        return self.res.loads

    @property
    def evictions(self) -> int:
        return self.res.evictions

    def close(self):
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
