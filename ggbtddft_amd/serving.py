"""Continuous-batching request loop for HHG scan requests on one GPU.

The reference runs its scans one after another inside one process
(main.cu:460-481).  Here requests (atom, intensity, wavelength, pulse length,
grid, optional second colour) arrive over time and are served concurrently:

* A *lane* is one propagator context (one HIP stream) with ``slots`` instance
  slots sharing the static tables of one (atom, N, r_max, Lmax, dt) key.  Its
  field is a ring buffer, so a request can be seated in any free slot at any
  chunk boundary while the other slots keep running (their orbitals and
  potentials are untouched by the seating, bit for bit).
* Lanes are created on demand while their device footprint (ggb_footprint +
  field + observable buffers) fits the memory budget; a request whose key has
  no lane and no room waits.  Static tables (ground state, eigenpairs) are
  computed once per key and kept.
* Each lane runs in its own host thread (ctypes releases the GIL), so work of
  different lanes overlaps on the device through their streams.
* A request finishes when its own number of steps is done; its dipole
  acceleration is turned into a spectrum on the GPU (``ggb_spectrum``) and the
  request's arrival -> spectrum time is recorded.

Every instance's arithmetic is column-local in every kernel, so a request's
result does not depend on what else shares its lane (tests check this).
"""
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


class Lane:
    """One context, `slots` instances of one static key."""

    def __init__(self, system, slots, chunk, ring, factory, device=0):
        self.system, self.slots, self.chunk, self.ring = system, slots, chunk, ring
        self.prop = factory(system, n_inst=slots, device=device)
        self.step = 0
        self.active = [None] * slots
        self.orbital_steps = 0
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
                # One idle chunk here, in the creating thread while no lane runs,
                # captures the graph and sizes the observable buffer, so the
                # chunks run from the lane threads neither capture nor allocate.
                self.prop.run(0, chunk)
                self.step = chunk
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

    def free_slots(self):
        return [s for s, a in enumerate(self.active) if a is None]

    def seat(self, req, slot, now):
        p = req.pulse()
        n = p.n_steps(req.dt)
        if n + self.chunk > self.ring:
            raise ValueError("request of %d steps does not fit the field ring (%d)" % (n, self.ring))
        E = p.midpoints(req.dt, n)
        E_end = p((np.arange(n) + 1.0) * req.dt)
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

    def advance(self):
        """Run one chunk; return the finished (active, obs) items."""
        k = self.chunk
        obs = self.prop.run(self.step, k)
        self.step += k
        finished = []
        for s, a in enumerate(self.active):
            if a is None:
                continue
            take = min(k, a.n_steps - a.done)
            a.buf.append((obs["dip"][:take, s], obs["acc"][:take, s], obs["norm"][:take, s]))
            a.done += take
            self.orbital_steps += take * self.system.n_orb
            if a.done >= a.n_steps:
                finished.append(a)
                self.active[s] = None
        if finished:
            self._mark()
        return finished

    def busy(self):
        return any(a is not None for a in self.active)


class Server:
    """Continuous batching over lanes; see the module docstring."""

    def __init__(self, memory_budget=64 << 30, slots=8, chunk=256, ring=1 << 16,
                 device=0, factory=None, system_factory=None, spectrum_beta=24.0,
                 precision=0):
        import functools
        from . import engine
        self.memory_budget = int(memory_budget)
        self.slots, self.chunk, self.ring, self.device = slots, chunk, ring, device
        self.precision = int(precision)
        self.factory = factory or functools.partial(engine.Propagator, precision=self.precision)
        self.footprint = (lambda sysm: engine.footprint(sysm, slots, ring, chunk, self.precision)) \
            if factory is None else (lambda sysm: 0)
        # with the GPU propagator the ground state of a new configuration is
        # computed on the device too
        gs = "device" if factory is None else None
        self.system_factory = system_factory or (
            lambda k: AtomSystem(k[0], k[1], k[2], k[3], k[4], ground_state=gs, gs_device=device))
        self.systems = {}
        self.lanes = {}
        self.used = 0
        self.beta = spectrum_beta
        # spectra on the GPU with the real propagator; the host definition
        # (same numbers, numpy FFT) with a stand-in backend
        self.device_spectrum = factory is None

    def _system(self, key):
        if key not in self.systems:
            self.systems[key] = self.system_factory(key)
        return self.systems[key]

    def _finish(self, lane, a, t0):
        d = np.concatenate([b[0] for b in a.buf])
        ac = np.concatenate([b[1] for b in a.buf])
        nr = np.concatenate([b[2] for b in a.buf])
        res = Result(request=a.req, dip=d, acc=ac, norm=nr)
        res.a_total = lane.system.total_acceleration(ac, nr, a.E_end)
        if self.device_spectrum:
            res.omega, res.spectrum, _ = spec.device_spectrum(res.a_total, a.req.dt, beta=self.beta,
                                                              device=self.device)
        else:
            res.omega, res.spectrum = spec.acceleration_spectrum(res.a_total, a.req.dt,
                                                                 beta=self.beta)
        res.t_finish = time.perf_counter() - t0
        return res

    def serve(self, requests, realtime=True):
        """Serve `requests` (arrival times in seconds from now).  Returns
        (results in request order, stats)."""
        t0 = time.perf_counter()
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting = deque()
        results = {}
        admit_time = {}
        errors = {}

        def now():
            return time.perf_counter() - t0

        def lane_for(key):
            lane = self.lanes.get(key)
            if lane is not None:
                return lane
            sysm = self._system(key)
            need = self.footprint(sysm)
            if self.used + need > self.memory_budget:
                return None
            lane = Lane(sysm, self.slots, self.chunk, self.ring, self.factory, self.device)
            self.lanes[key] = lane
            self.used += need
            return lane

        def admit():
            while pending and (not realtime or pending[0].arrival <= now()):
                waiting.append(pending.popleft())
            still = deque()
            while waiting:
                req = waiting.popleft()
                try:
                    lane = lane_for(req.key())
                    free = lane.free_slots() if lane is not None else []
                    if not free:
                        still.append(req)
                        continue
                    lane.seat(req, free[0], now())
                    admit_time[req.rid] = now()
                except Exception as exc:  # a bad request fails alone
                    errors[req.rid] = repr(exc)
            waiting.extend(still)

        def run_lane(lane, out):
            try:
                out.extend(lane.advance())
            except Exception as exc:
                out.append(exc)

        total = len(requests)
        while len(results) + len(errors) < total:
            admit()
            busy = [l for l in self.lanes.values() if l.busy()]
            if not busy:
                if pending and realtime:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
                    continue
                if not pending and waiting:
                    # nothing running and nothing admissible: budget too small
                    for req in waiting:
                        errors[req.rid] = "no lane fits the memory budget"
                    waiting.clear()
                continue
            outs = [[] for _ in busy]
            threads = [threading.Thread(target=run_lane, args=(l, o)) for l, o in zip(busy, outs)]
            for th in threads:
                th.start()
            for th in threads:
                th.join()
            for lane, out in zip(busy, outs):
                for item in out:
                    if isinstance(item, Exception):  # the lane's requests fail together
                        for a in lane.active:
                            if a is not None:
                                errors[a.req.rid] = repr(item)
                        lane.active = [None] * lane.slots
                        try:
                            lane._mark()
                        except Exception:
                            pass
                        continue
                    try:
                        res = self._finish(lane, item, t0)
                    except Exception as exc:  # e.g. the spectrum: this request fails alone
                        errors[item.req.rid] = repr(exc)
                        continue
                    res.t_admit = admit_time.get(item.req.rid, math.nan)
                    results[item.req.rid] = res
        out = []
        for req in requests:
            if req.rid in results:
                out.append(results[req.rid])
            else:
                out.append(Result(request=req, error=errors.get(req.rid, "not served")))
        wall = now()
        tts = np.array([r.time_to_spectrum for r in out if r.error is None])
        ots = sum(l.orbital_steps for l in self.lanes.values())
        stats = dict(wall_s=wall, served=int(tts.size), failed=total - int(tts.size),
                     p50_time_to_spectrum_s=float(np.percentile(tts, 50)) if tts.size else math.nan,
                     p99_time_to_spectrum_s=float(np.percentile(tts, 99)) if tts.size else math.nan,
                     orbital_timesteps_per_s=ots / wall if wall > 0 else math.nan,
                     lanes=len(self.lanes), device_bytes=self.used)
        return out, stats

    def close(self):
        for lane in self.lanes.values():
            lane.prop.close()
        self.lanes.clear()


def poisson_arrivals(rate, n, seed=0):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
