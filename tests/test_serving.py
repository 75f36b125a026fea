"""Serving loop (continuous batching) on the CPU with the oracle-backed
stand-in propagator: slot reuse, ring-buffer field offsets, per-request
truncation and spectra match a direct oracle run of each request; a bad
request or a failing lane fails only its own requests."""
import numpy as np
import pytest

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


def _he(rid, arrival=0.0, N=24, lmax=3, **kw):
    return Request(rid=rid, atom="He", intensity=1e14, wavelength=800, n_cycles=1, N=N,
                   r_max=20.0, lmax=lmax, dt=0.2, arrival=arrival, **kw)


def test_idle_lanes_are_evicted_for_a_new_key():
    # the budget holds one lane: the second key's request closes the idle first lane
    srv = Server(slots=1, chunk=40, ring=2048, factory=OraclePropagator, memory_budget=10)
    srv.footprint = lambda sysm: 10
    reqs = [_he(0), _he(1, arrival=0.5, N=28)]
    results, stats = srv.serve(reqs, realtime=True)
    assert stats["served"] == 2 and stats["failed"] == 0
    assert stats["lanes_created"] == 2 and stats["lanes_closed"] == 1 and srv.used == 10
    O = OP.System("He", 28, 20.0, 3, 0.2)
    p = reqs[1].pulse()
    ref = O.run(p.midpoints(0.2, p.n_steps(0.2)))
    np.testing.assert_allclose(results[1].acc, ref["acc"], rtol=0, atol=1e-13)
    srv.close()
    assert srv.used == 0 and not srv.lanes


def test_cancel_and_deadline_fail_only_their_requests():
    srv = Server(slots=1, chunk=40, ring=4096, factory=OraclePropagator)
    long_req = Request(rid=5, atom="He", intensity=1e14, wavelength=800, n_cycles=6, N=24,
                       r_max=20.0, lmax=3, dt=0.2, deadline=0.0)
    srv.cancel(1)
    results, stats = srv.serve([_he(0), _he(1), long_req, _he(2)], realtime=False)
    errs = {r.request.rid: r.error for r in results}
    assert errs[0] is None and errs[2] is None
    assert errs[1] == "cancelled" and "deadline" in errs[5]
    assert stats["served"] == 2 and stats["cancelled"] == 1


class _BlowUpPropagator(OraclePropagator):
    """Stand-in whose instance 1 reports a growing norm (mode "norm") or a
    non-finite acceleration (mode "nan")."""
    mode = "norm"

    def run(self, step0, nsteps, channels=False):
        out = super().run(step0, nsteps)
        if self.n_inst > 1:
            if self.mode == "norm":
                out["norm"][:, 1, :] *= 1.01
            else:
                out["acc"][nsteps // 2:, 1, 0] = np.nan
        return out


@pytest.mark.parametrize("mode,msg", [("norm", "norm grew"), ("nan", "not finite")])
def test_unhealthy_request_fails_alone(mode, msg):
    factory = type("P", (_BlowUpPropagator,), {"mode": mode})
    srv = Server(slots=2, chunk=40, ring=2048, factory=factory)
    results, stats = srv.serve([_he(0), _he(1)], realtime=False)
    errs = {r.request.rid: r.error for r in results}
    assert errs[0] is None and msg in errs[1]
    assert stats["served"] == 1 and stats["failed"] == 1


def test_field_guard_refuses_absurd_intensity():
    srv = Server(slots=1, chunk=40, ring=2048, factory=OraclePropagator)
    bad = Request(rid=3, atom="He", intensity=1e20, wavelength=800, n_cycles=1, N=24, r_max=20.0,
                  lmax=3, dt=0.2)
    results, stats = srv.serve([_he(0), bad], realtime=False)
    assert results[0].error is None and "field above" in results[1].error
