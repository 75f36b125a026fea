"""Continuous-batching request loop for HHG scan requests on one GPU.

The reference runs its scans one after another inside one process
(main.cu:460-481).  Here requests (atom, intensity, wavelength, pulse length,
grid, optional second colour) arrive over time and are served concurrently:

* A *lane* is one propagator context (one HIP stream) with ``slots`` instance
  slots sharing the static tables of one (atom, N, r_max, Lmax, dt) key.  Its
  field is a ring buffer, so a request can be seated in any free slot at any
  chunk boundary while the other slots keep running (their orbitals and
  potentials are untouched by the seating, bit for bit).
* Every lane runs its own loop in its own host thread (ctypes releases the
  GIL): it seats waiting requests of its key into free slots, runs one chunk
  of steps, hands finished requests on, and starts again -- no lane waits for
  another lane's chunk, and a request is admitted at the first chunk boundary
  of its lane after it arrives.  ``max_concurrent`` caps how many lanes run a
  chunk at the same moment (their kernels share the GPU).
* Lanes are created on demand while their device footprint (ggb_footprint +
  field + observable buffers) fits the memory budget; when it does not, idle
  lanes are closed, least recently used first.  Static tables (ground state,
  eigenpairs) are computed once per key and kept for ``max_systems`` keys.
* Faults stay with their request: a field that is not finite (or above
  10 a.u.) is refused at admission; a request whose observables turn
  non-finite or whose orbital norms grow (the absorber only removes norm) is
  stopped and fails alone; ``cancel(rid)`` and per-request deadlines free the
  slot at the next chunk boundary; a lane whose chunk raises fails only the
  requests seated in it.
* A request finishes when its own number of steps is done; its dipole
  acceleration is turned into a spectrum on the GPU by a spectrum thread while
  its lane runs on (``ggb_spectrum_noalloc``, which allocates nothing and so
  runs beside the chunks; ``ggb_spectrum`` with the device alone for series
  beyond its scratch), and the request's arrival -> spectrum time is recorded.

Calls that allocate or free (lane creation and destruction, a chunk that
captures a graph, the allocating spectrum) take the device alone; chunk runs
and the allocation-free spectrum share it (``_Gate``).

Every instance's arithmetic is column-local in every kernel, so a request's
result does not depend on what else shares its lane (tests check this).
"""
import contextlib
import math
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import spectrum as spec
from .pulse import Pulse
from .system import AtomSystem

FIELD_LIMIT = 10.0   # a.u. (3.5e18 W/cm^2): larger fields are refused at admission


@dataclass
class Request:
    rid: int
    atom: str
    intensity: float          # W/cm^2
    wavelength: float         # nm
    n_cycles: float
    N: int
    r_max: float
    lmax: int
    dt: float
    arrival: float = 0.0      # seconds after the server's start
    second: Optional[tuple] = None   # (harmonic, intensity ratio, phase)
    deadline: Optional[float] = None  # seconds after arrival; unfinished by then -> fails

    def key(self):
        return (self.atom, self.N, float(self.r_max), self.lmax, float(self.dt))

    def pulse(self):
        return Pulse(self.intensity, self.wavelength, self.n_cycles, second=self.second)


@dataclass
class Result:
    request: Request
    dip: np.ndarray = None        # [n_steps, n_orb]
    acc: np.ndarray = None
    norm: np.ndarray = None
    a_total: np.ndarray = None    # Ehrenfest acceleration of all electrons
    omega: np.ndarray = None
    spectrum: np.ndarray = None
    t_admit: float = math.nan
    t_finish: float = math.nan
    error: Optional[str] = None

    @property
    def time_to_spectrum(self):
        return self.t_finish - self.request.arrival


@dataclass
class _Active:
    req: Request
    slot: int
    start: int                    # lane step at which it was seated
    n_steps: int
    E_end: np.ndarray
    buf: list = field(default_factory=list)
    done: int = 0
    error: Optional[str] = None


class _Gate:
    """Device access shared by chunk runs, exclusive for calls that allocate,
    free, capture a graph or use the legacy stream."""

    def __init__(self):
        self._cv = threading.Condition()
        self._readers = 0
        self._writer = False
        self._writers_waiting = 0

    @contextlib.contextmanager
    def shared(self):
        with self._cv:
            while self._writer or self._writers_waiting:
                self._cv.wait()
            self._readers += 1
        try:
            yield
        finally:
            with self._cv:
                self._readers -= 1
                if not self._readers:
                    self._cv.notify_all()

    @contextlib.contextmanager
    def exclusive(self):
        with self._cv:
            self._writers_waiting += 1
            while self._writer or self._readers:
                self._cv.wait()
            self._writers_waiting -= 1
            self._writer = True
        try:
            yield
        finally:
            with self._cv:
                self._writer = False
                self._cv.notify_all()


class Lane:
    """One context, `slots` instances of one static key."""

    def __init__(self, system, slots, chunk, ring, factory, device=0, norm_tol=1e-6):
        self.system, self.slots, self.chunk, self.ring = system, slots, chunk, ring
        self.norm_tol = norm_tol
        self.prop = factory(system, n_inst=slots, device=device)
        self.step = 0
        self.active = [None] * slots
        self.orbital_steps = 0
        self.last_used = time.perf_counter()
        self.graph_dirty = False  # the next chunk captures its graph
        # idle slots skip the angular and potential kernels (and kinetic tiles
        # whose instances are all idle)
        self._masked = hasattr(self.prop, "set_active")
        self._variant = None
        self._variants = None  # {variant: (ms with one slot in use, ms with all)}
        try:
            if hasattr(self.prop, "autotune"):  # kinetic tile shapes for this lane's shape
                if self._masked:
                    self.prop.set_active([True] * slots)
                self._variant, t_all = self.prop.autotune(reps=3)
                if self._masked and slots > 1 and hasattr(self.prop, "set_kinetic_variant"):
                    # with few slots in use a narrower variant can win (it skips idle
                    # columns at a finer grain); time each with one slot in use too
                    self.prop.set_active([True] + [False] * (slots - 1))
                    _, t_one = self.prop.autotune(reps=3, cache=False)
                    self.prop.set_kinetic_variant(self._variant)
                    keys = set(t_all) & set(t_one)
                    if len(keys) > 1:
                        self._variants = {k: (t_one[k], t_all[k]) for k in keys}
            self.prop.field_ring(ring)
            self._mark()
            if hasattr(self.prop, "set_graph"):  # each chunk replays a captured run of steps
                self.prop.set_graph(min(64, chunk))
                # One idle chunk here, while the creator holds the device alone,
                # captures the graph and sizes the observable buffer, so that the
                # lane's own chunks neither capture nor allocate.
                self.prop.run(0, chunk)
                self.step = chunk
            self.graph_dirty = False
        except Exception:
            if hasattr(self.prop, "close"):  # no half-made lane keeps device memory
                self.prop.close()
            raise

    def _mark(self):
        """Hand the slots in use to the device and, when the variants were timed
        at one slot and at all, take the one whose linear interpolation in the
        number of slots in use is fastest."""
        if not self._masked:
            return
        mask = [a is not None for a in self.active]
        self.prop.set_active(mask)
        n = sum(mask)
        if self._variants and n > 0:
            f = (n - 1) / (self.slots - 1)
            best = min(sorted(self._variants),
                       key=lambda k: self._variants[k][0] + f * (self._variants[k][1] - self._variants[k][0]))
            if best != self._variant:
                self.prop.set_kinetic_variant(best)  # the next chunk captures its graph again
                self._variant = best
                self.graph_dirty = True

    def free_slots(self):
        return [s for s, a in enumerate(self.active) if a is None]

    def seat(self, req, slot, now):
        p = req.pulse()
        n = p.n_steps(req.dt)
        if n < 1:
            raise ValueError("request of no steps")
        if n + self.chunk > self.ring:
            raise ValueError("request of %d steps does not fit the field ring (%d)" % (n, self.ring))
        E = p.midpoints(req.dt, n)
        E_end = p((np.arange(n) + 1.0) * req.dt)
        if not (np.all(np.isfinite(E)) and np.all(np.isfinite(E_end))):
            raise ValueError("field is not finite")
        if max(np.abs(E).max(), np.abs(E_end).max()) > FIELD_LIMIT:
            raise ValueError("field above %g a.u." % FIELD_LIMIT)
        self.active[slot] = _Active(req=req, slot=slot, start=self.step, n_steps=n, E_end=E_end)
        try:
            self._mark()  # in use before the potential of the new orbitals is built
            self.prop.upload_instance(slot)
            # field for the request's own steps, zero for the chunk tail after it
            self.prop.set_field_slice(slot, self.step, np.concatenate([E, np.zeros(self.chunk)]))
        except Exception:
            self.active[slot] = None  # the slot stays free
            self._mark()
            raise
        return n

    def drop(self, slot):
        """Free a slot at a chunk boundary (cancelled or past its deadline)."""
        self.active[slot] = None
        self._mark()

    def advance(self):
        """Run one chunk; return the requests that ended in it (finished, or
        stopped with ``error`` set: observables not finite or norm growth)."""
        return self.consume(self.prop.run(self.step, self.chunk))

    def consume(self, obs):
        """Take the observables of the chunk just run (from self.step): hand
        each request its rows, and return the requests that ended in it."""
        k = self.chunk
        self.step += k
        self.last_used = time.perf_counter()
        ended = []
        for s, a in enumerate(self.active):
            if a is None:
                continue
            take = min(k, a.n_steps - a.done)
            d, ac, nr = obs["dip"][:take, s], obs["acc"][:take, s], obs["norm"][:take, s]
            self.orbital_steps += take * self.system.n_orb
            if not (np.all(np.isfinite(d)) and np.all(np.isfinite(ac)) and np.all(np.isfinite(nr))):
                a.error = "observables not finite at step %d" % (a.done + take)
            elif nr.max() > 1.0 + self.norm_tol:
                a.error = "orbital norm grew to %.3g at step %d" % (nr.max(), a.done + take)
            a.buf.append((d, ac, nr))
            a.done += take
            if a.error is not None or a.done >= a.n_steps:
                ended.append(a)
                self.active[s] = None
        if ended:
            self._mark()
        return ended

    def busy(self):
        return any(a is not None for a in self.active)


class Server:
    """Continuous batching over lanes; see the module docstring."""

    def __init__(self, memory_budget=64 << 30, slots=8, chunk=256, ring=1 << 16,
                 device=0, factory=None, system_factory=None, spectrum_beta=24.0,
                 precision=0, max_concurrent=None, norm_tol=1e-6, max_systems=16,
                 device_ground_state=None):
        import functools
        from . import engine
        self.memory_budget = int(memory_budget)
        self.slots, self.chunk, self.ring, self.device = slots, chunk, ring, device
        self.precision = int(precision)
        self.norm_tol = norm_tol
        self.max_systems = max_systems
        self.factory = factory or functools.partial(engine.Propagator, precision=self.precision)
        self.footprint = (lambda sysm: engine.footprint(sysm, slots, ring, chunk, self.precision)) \
            if factory is None else (lambda sysm: 0)
        # with the GPU propagator the ground state of a new configuration is
        # computed on the device too (by default; a stand-in backend uses the
        # host iteration)
        self._device_gs = (factory is None) if device_ground_state is None else bool(device_ground_state)
        self.system_factory = system_factory or self._make_system
        self.systems = {}
        self.lanes = {}
        self._need = {}
        self.used = 0
        self.beta = spectrum_beta
        # spectra on the GPU with the real propagator; the host definition
        # (same numbers, numpy FFT) with a stand-in backend
        self.device_spectrum = factory is None
        self._gate = _Gate()
        self._sem = threading.BoundedSemaphore(max_concurrent) if max_concurrent else None
        self._cv = threading.Condition()
        self._cancelled = set()
        self._closed_orbital_steps = 0
        self.lanes_created = 0
        self.lanes_closed = 0

    def _make_system(self, key):
        """Static data of a key: the ground state and the propagator tables'
        eigenpairs on the device (both allocate, so each holds the device
        alone; lanes wait meanwhile), the angular tables on the host between
        them while the lanes run."""
        if not self._device_gs:
            return AtomSystem(key[0], key[1], key[2], key[3], key[4])
        from .atoms import Atom
        from .grid import RadialGrid
        from .groundstate import DeviceGroundState
        grid = RadialGrid(key[1], key[2])
        with self._gate.exclusive():
            gs = DeviceGroundState(grid, Atom(key[0]), device=self.device)
        sysm = AtomSystem(key[0], key[1], key[2], key[3], key[4], ground_state=gs, tables=None)
        with self._gate.exclusive():  # the batched eigendecomposition allocates
            sysm.build_tables("device", self.device)
        return sysm

    def _system(self, key):
        if key not in self.systems:
            self.systems[key] = self.system_factory(key)
        return self.systems[key]

    def cancel(self, rid):
        """Cancel request `rid`: a waiting request is dropped, a running one
        frees its slot at its lane's next chunk boundary."""
        with self._cv:
            self._cancelled.add(rid)
            self._cv.notify_all()

    def _result(self, lane, a):
        """Observables of a finished request and its total dipole acceleration."""
        d = np.concatenate([b[0] for b in a.buf])
        ac = np.concatenate([b[1] for b in a.buf])
        nr = np.concatenate([b[2] for b in a.buf])
        res = Result(request=a.req, dip=d, acc=ac, norm=nr)
        res.a_total = lane.system.total_acceleration(ac, nr, a.E_end)
        return res

    def _spectrum(self, res):
        dt = res.request.dt
        if not self.device_spectrum:
            res.omega, res.spectrum = spec.acceleration_spectrum(res.a_total, dt, beta=self.beta)
        elif spec.noalloc_fits(res.a_total.size):
            # no allocation: runs beside the lanes' chunks (their non-blocking
            # streams do not wait on the legacy stream it uses)
            with self._gate.shared():
                res.omega, res.spectrum, _ = spec.device_spectrum(res.a_total, dt, beta=self.beta,
                                                                  device=self.device, noalloc=True)
        else:
            with self._gate.exclusive():  # ggb_spectrum allocates
                res.omega, res.spectrum, _ = spec.device_spectrum(res.a_total, dt, beta=self.beta,
                                                                  device=self.device)

    def _close_lane(self, key):
        lane = self.lanes.pop(key)
        with self._gate.exclusive():
            lane.prop.close()
        self.used -= self._need.pop(key, 0)
        self._closed_orbital_steps += lane.orbital_steps
        self.lanes_closed += 1
        if len(self.systems) > self.max_systems:
            self.systems.pop(key, None)

    def serve(self, requests, realtime=True):
        """Serve `requests` (arrival times in seconds from now).  Returns
        (results in request order, stats)."""
        t0 = time.perf_counter()
        cv = self._cv
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting = {}          # key -> deque of arrived requests without a slot
        results, errors, admit_time = {}, {}, {}
        threads = {}          # key -> (thread, stop flag)
        finished = deque()    # (result without spectrum) for the spectrum thread
        spec_stop = [False]
        total = len(requests)
        ids = {r.rid for r in requests}

        def now():
            return time.perf_counter() - t0

        def overdue(req):
            return req.deadline is not None and now() > req.arrival + req.deadline

        def fail(rid, msg):
            with cv:
                errors[rid] = msg
                cv.notify_all()

        def worker(key, lane, stop):
            while True:
                with cv:
                    while True:
                        if stop[0]:
                            return
                        drops = [a for a in lane.active if a is not None and
                                 (a.req.rid in self._cancelled or overdue(a.req))]
                        seats = []
                        free = [s for s in lane.free_slots()]
                        q = waiting.get(key)
                        while q and free:
                            req = q.popleft()
                            if req.rid in self._cancelled:
                                errors[req.rid] = "cancelled"
                            elif overdue(req):
                                errors[req.rid] = "deadline passed before admission"
                            else:
                                seats.append((req, free.pop(0)))
                        if seats or drops or lane.busy():
                            break
                        cv.wait(0.05)
                with self._gate.shared():
                    for a in drops:
                        lane.drop(a.slot)
                        fail(a.req.rid, "cancelled" if a.req.rid in self._cancelled else
                             "deadline passed after %d of %d steps" % (a.done, a.n_steps))
                    for req, slot in seats:
                        try:
                            lane.seat(req, slot, now())
                            with cv:
                                admit_time[req.rid] = now()
                        except Exception as exc:  # a bad request fails alone
                            fail(req.rid, repr(exc))
                if not lane.busy():
                    continue
                try:
                    with (self._sem or contextlib.nullcontext()):
                        with (self._gate.exclusive() if lane.graph_dirty else self._gate.shared()):
                            dirty = lane.graph_dirty
                            ended = lane.advance()
                            if dirty:
                                lane.graph_dirty = False
                except Exception as exc:  # the lane's requests fail together
                    for a in lane.active:
                        if a is not None:
                            fail(a.req.rid, repr(exc))
                    lane.active = [None] * lane.slots
                    try:
                        with self._gate.shared():
                            lane._mark()
                    except Exception:
                        pass
                    continue
                for a in ended:
                    if a.error is not None:
                        fail(a.req.rid, a.error)
                        continue
                    try:
                        res = self._result(lane, a)
                    except Exception as exc:  # this request fails alone
                        fail(a.req.rid, repr(exc))
                        continue
                    with cv:  # its spectrum is the spectrum thread's; the lane runs on
                        finished.append(res)
                        cv.notify_all()

        def spectra():
            """Spectra of finished requests, one after another, beside the lanes."""
            while True:
                with cv:
                    while not finished and not spec_stop[0]:
                        cv.wait(0.05)
                    if not finished:
                        return
                    res = finished.popleft()
                rid = res.request.rid
                try:
                    self._spectrum(res)
                except Exception as exc:  # this request fails alone
                    fail(rid, repr(exc))
                    continue
                res.t_finish = now()
                with cv:
                    res.t_admit = admit_time.get(rid, math.nan)
                    results[rid] = res
                    cv.notify_all()

        def start_lane(key):
            """Create the lane of `key` if its footprint fits (closing idle
            lanes, least recently used first); returns False when it cannot."""
            sysm = self._system(key)
            need = self.footprint(sysm)
            while self.used + need > self.memory_budget:
                with cv:
                    idle = [(l.last_used, k) for k, l in self.lanes.items()
                            if not l.busy() and not waiting.get(k)]
                if not idle:
                    return False
                k = min(idle)[1]
                th, stop = threads.pop(k)
                with cv:
                    stop[0] = True
                    cv.notify_all()
                th.join()
                self._close_lane(k)
            with self._gate.exclusive():
                lane = Lane(sysm, self.slots, self.chunk, self.ring, self.factory, self.device,
                            self.norm_tol)
            self.lanes[key] = lane
            self._need[key] = need
            self.used += need
            self.lanes_created += 1
            stop = [False]
            th = threading.Thread(target=worker, args=(key, lane, stop), daemon=True)
            threads[key] = (th, stop)
            th.start()
            return True

        spec_thread = threading.Thread(target=spectra, daemon=True)
        spec_thread.start()
        try:
            while True:
                with cv:
                    while pending and (not realtime or pending[0].arrival <= now()):
                        req = pending.popleft()
                        if req.rid in self._cancelled:
                            errors[req.rid] = "cancelled"
                        else:
                            waiting.setdefault(req.key(), deque()).append(req)
                    for k, q in waiting.items():
                        for req in [r for r in q if r.rid in self._cancelled]:
                            q.remove(req)
                            errors[req.rid] = "cancelled"
                    need_lane = [k for k, q in waiting.items() if q and k not in self.lanes]
                    if len(results) + len(errors) >= total:
                        break
                    cv.notify_all()
                for key in need_lane:
                    try:
                        ok = start_lane(key)
                    except Exception as exc:  # the key's requests fail, others go on
                        with cv:
                            for req in waiting.pop(key, ()):
                                errors[req.rid] = repr(exc)
                        continue
                    if not ok:
                        with cv:
                            running = any(l.busy() for l in self.lanes.values()) or \
                                any(q for k, q in waiting.items() if k in self.lanes)
                            if not running:
                                # nothing running and this key cannot get a lane
                                for req in waiting.pop(key, ()):
                                    errors[req.rid] = "no lane fits the memory budget"
                with cv:
                    if len(results) + len(errors) >= total:
                        break
                    wait = 0.05
                    if pending and realtime:
                        wait = max(0.0, min(wait, pending[0].arrival - now()))
                    cv.wait(wait)
        finally:
            with cv:
                for _, stop in threads.values():
                    stop[0] = True
                spec_stop[0] = True
                cv.notify_all()
            for th, _ in threads.values():
                th.join()
            spec_thread.join()
        out = []
        for req in requests:
            if req.rid in results:
                out.append(results[req.rid])
            else:
                out.append(Result(request=req, error=errors.get(req.rid, "not served")))
        wall = now()
        tts = np.array([r.time_to_spectrum for r in out if r.error is None])
        ots = sum(l.orbital_steps for l in self.lanes.values()) + self._closed_orbital_steps
        stats = dict(wall_s=wall, served=int(tts.size), failed=total - int(tts.size),
                     p50_time_to_spectrum_s=float(np.percentile(tts, 50)) if tts.size else math.nan,
                     p99_time_to_spectrum_s=float(np.percentile(tts, 99)) if tts.size else math.nan,
                     orbital_timesteps_per_s=ots / wall if wall > 0 else math.nan,
                     lanes=len(self.lanes), lanes_created=self.lanes_created,
                     lanes_closed=self.lanes_closed, device_bytes=self.used,
                     cancelled=sum(1 for r in out if r.error == "cancelled"))
        with cv:
            self._cancelled -= ids
        return out, stats

    def close(self):
        for key in list(self.lanes):
            self._close_lane(key)


def poisson_arrivals(rate, n, seed=0):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
