"""Serving loop (continuous batching) on the CPU with the oracle-backed
stand-in propagator: slot reuse, ring-buffer field offsets, per-request
truncation and spectra match a direct oracle run of each request; a bad
request or a failing lane fails only its own requests."""
import numpy as np

from fake_backend import OraclePropagator
from ggbtddft_amd.serving import Request, Server, poisson_arrivals
from oracle import propagate as OP


def _requests():
    arr = poisson_arrivals(50.0, 5, seed=0)
    reqs = []
    for k, (I, lam, cyc) in enumerate([(1e14, 800, 1), (3e14, 700, 1.5), (2e14, 900, 1),
                                        (5e13, 600, 2), (4e14, 800, 1)]):
        reqs.append(Request(rid=k, atom="He", intensity=I, wavelength=lam, n_cycles=cyc, N=24,
                            r_max=20.0, lmax=3, dt=0.2, arrival=float(arr[k])))
    return reqs


def test_serving_matches_direct_runs():
    reqs = _requests()
    srv = Server(slots=2, chunk=40, ring=2048, factory=OraclePropagator)
    results, stats = srv.serve(reqs, realtime=False)
    assert stats["served"] == len(reqs) and stats["failed"] == 0
    O = OP.System("He", 24, 20.0, 3, 0.2)
    for res in results:
        assert res.error is None
        p = res.request.pulse()
        n = p.n_steps(0.2)
        ref = O.run(p.midpoints(0.2, n))
        assert res.acc.shape == (n, O.norb)
        np.testing.assert_allclose(res.acc, ref["acc"], rtol=0, atol=1e-13)
        np.testing.assert_allclose(res.norm, ref["norm"], rtol=0, atol=1e-13)
        assert np.all(np.isfinite(res.spectrum)) and res.spectrum.size == 17 * n
    assert np.isfinite(stats["p99_time_to_spectrum_s"])


def test_budget_too_small_fails_cleanly():
    reqs = _requests()[:1]
    srv = Server(slots=1, chunk=40, ring=2048, factory=OraclePropagator, memory_budget=0)
    srv.footprint = lambda sysm: 1
    results, stats = srv.serve(reqs, realtime=False)
    assert stats["failed"] == 1 and "budget" in results[0].error


def test_request_too_long_for_the_ring_fails_alone():
    reqs = _requests()[:3]
    long_req = Request(rid=99, atom="He", intensity=1e14, wavelength=800, n_cycles=40, N=24,
                       r_max=20.0, lmax=3, dt=0.2)
    srv = Server(slots=2, chunk=40, ring=2048, factory=OraclePropagator)
    results, stats = srv.serve(reqs + [long_req], realtime=False)
    assert stats["served"] == 3 and stats["failed"] == 1
    assert "field ring" in results[-1].error
    assert all(r.error is None for r in results[:3])


class _FailingPropagator(OraclePropagator):
    """Stand-in whose chunk fails for one atom's lane (a device error)."""

    def run(self, step0, nsteps):
        if self.sys.atom.symbol == "Ne":
            raise RuntimeError("device error in this lane")
        return super().run(step0, nsteps)


def test_failing_lane_fails_only_its_requests():
    reqs = _requests()[:2]
    reqs.append(Request(rid=7, atom="Ne", intensity=1e14, wavelength=800, n_cycles=1, N=24,
                        r_max=20.0, lmax=3, dt=0.2))
    srv = Server(slots=2, chunk=40, ring=2048, factory=_FailingPropagator)
    results, stats = srv.serve(reqs, realtime=False)
    assert stats["served"] == 2 and stats["failed"] == 1
    assert "device error" in results[-1].error
    assert all(r.error is None for r in results[:2])
