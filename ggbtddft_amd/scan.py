"""Batched scans of uneven instances (different grids, intensities,
wavelengths and pulse lengths).

The reference runs the scans of one grid one after another
(main.cu:460-481).  Here a scan is a list of requests; instances of one static
key (atom, N, r_max, Lmax, dt) share a context's slots and tables, and the
instances of different keys run side by side (see ``serving`` and
``ScanGroup``).

``scan_requests`` draws the batched-scan workload: n instances from a seed
over He and Ne (uniform), intensity log-uniform in [5e13, 5e14] W/cm^2,
wavelength uniform in [600, 2000] nm, sin^2 length uniform integer in
[4, 16] cycles (so step counts are uneven), and the grid from the quiver
radius alpha = E0 / omega0^2: N = 256, Lmax = 32, r_max = 100 below 25 a.u.,
N = 384, Lmax = 48, r_max = 150 below 60 a.u., else N = 512, Lmax = 64,
r_max = 200; dt = 0.05.
"""
import math
import time

import numpy as np

from . import spectrum as spec
from .pulse import Pulse
from .serving import Lane, Request, Result

GRIDS = ((25.0, 256, 32, 100.0), (60.0, 384, 48, 150.0), (np.inf, 512, 64, 200.0))


def grid_for(pulse):
    """(N, Lmax, r_max) of the quiver radius alpha = E0 / omega0^2."""
    alpha = pulse.E0 / pulse.omega0 ** 2
    for lim, nr, lmax, rmax in GRIDS:
        if alpha < lim:
            return nr, lmax, rmax
    raise AssertionError


def scan_requests(n, seed=0, dt=0.05, cycle_scale=1.0):
    """The draw above; cycle_scale < 1 shortens every pulse by that factor
    (same grids and fields' shapes, proportionally uneven step counts) for a
    bounded timing run."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        atom = "He" if rng.uniform() < 0.5 else "Ne"
        I = float(np.exp(rng.uniform(np.log(5e13), np.log(5e14))))
        lam = float(rng.uniform(600.0, 2000.0))
        cyc = int(rng.integers(4, 17))
        nr, lmax, rmax = grid_for(Pulse(I, lam, cyc))
        reqs.append(Request(rid=k, atom=atom, intensity=I, wavelength=lam,
                            n_cycles=cyc * float(cycle_scale), N=nr, r_max=rmax, lmax=lmax, dt=dt))
    return reqs


class ScanGroup:
    """A batched scan served as one batch on one GPU: one lane (context) per
    static key with a slot per request of that key, the chunks of all busy
    lanes advanced by one hipGraph launch (engine.Group), requests seated in
    free slots at chunk boundaries (longest first) until all have run their
    own number of steps.  No host threads: one loop launches every lane's
    chunk at once.
    Results are each request's observables, Ehrenfest acceleration and
    spectrum (as the serving loop's)."""

    def __init__(self, systems, slots, chunk=256, ring=1 << 17, precision=0, device=0,
                 graph_steps=64, spectrum_beta=24.0):
        import functools
        from . import engine
        self.systems, self.slots, self.chunk, self.ring = systems, slots, chunk, ring
        self.device, self.graph_steps, self.beta = device, graph_steps, spectrum_beta
        self.factory = functools.partial(engine.Propagator, precision=int(precision))
        self.engine = engine

    def run(self, requests):
        """Serve every request (arrival times ignored); returns (results in
        request order, stats)."""
        t0 = time.perf_counter()
        waiting = {}
        # longest first within each lane (all requests are there at the start):
        # the short ones fill the slots the long ones leave, so a lane's last
        # chunks are not one long request running alone
        for r in sorted(requests, key=lambda r: -r.pulse().n_steps(r.dt)):
            waiting.setdefault(r.key(), []).append(r)
        lanes = {k: Lane(self.systems[k], min(self.slots, len(v)), self.chunk, self.ring,
                         self.factory, self.device) for k, v in waiting.items()}
        results, errors = {}, {}
        group, members = None, None
        try:
            while len(results) + len(errors) < len(requests):
                for k, lane in lanes.items():
                    q = waiting[k]
                    for slot in lane.free_slots():
                        if not q:
                            break
                        req = q.pop(0)
                        try:
                            lane.seat(req, slot, time.perf_counter() - t0)
                        except Exception as exc:  # a bad request fails alone
                            errors[req.rid] = repr(exc)
                busy = [k for k, lane in lanes.items() if lane.busy()]
                if not busy:
                    if not any(waiting.values()):
                        break  # every request has failed or finished
                    continue
                if busy != members:  # the group follows the lanes with work
                    if group is not None:
                        group.close()
                    group = self.engine.Group([lanes[k].prop for k in busy], self.graph_steps)
                    members = busy
                obs = group.run([lanes[k].step for k in busy], self.chunk)
                for k, o in zip(busy, obs):
                    lane = lanes[k]
                    for a in lane.consume(o):
                        if a.error is not None:
                            errors[a.req.rid] = a.error
                            continue
                        results[a.req.rid] = self._finish(lane, a, t0)
        finally:
            if group is not None:
                group.close()
            for lane in lanes.values():
                lane.prop.close()
        out = [results.get(r.rid) or Result(request=r, error=errors.get(r.rid, "not served"))
               for r in requests]
        wall = time.perf_counter() - t0
        ots = sum(l.orbital_steps for l in lanes.values())
        tts = np.array([r.time_to_spectrum for r in out if r.error is None])
        return out, dict(wall_s=wall, served=int(tts.size), failed=len(requests) - int(tts.size),
                         orbital_timesteps_per_s=ots / wall if wall > 0 else math.nan,
                         lanes=len(lanes))

    def _finish(self, lane, a, t0):
        d = np.concatenate([b[0] for b in a.buf])
        ac = np.concatenate([b[1] for b in a.buf])
        nr = np.concatenate([b[2] for b in a.buf])
        res = Result(request=a.req, dip=d, acc=ac, norm=nr)
        res.a_total = lane.system.total_acceleration(ac, nr, a.E_end)
        res.omega, res.spectrum, _ = spec.device_spectrum(res.a_total, a.req.dt, beta=self.beta,
                                                          device=self.device)
        res.t_finish = time.perf_counter() - t0
        return res
