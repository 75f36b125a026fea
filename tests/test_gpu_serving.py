"""Serving on the GPU: requests seated into running lanes (continuous
batching, two lanes overlapping on the device) give the same results as each
request propagated alone, within 1e-12, in FP64 and in the FP32 mixed form."""

import numpy as np
import pytest

from ggbtddft_amd.serving import Request, Server, poisson_arrivals

pytestmark = pytest.mark.gpu


def test_mixed_equals_alone():
    print("case: serving mixed vs alone (He and Ne lanes)", flush=True)
    arr = poisson_arrivals(200.0, 6, seed=1)
    specs = [("He", 2e14, 800, 1), ("Ne", 3e14, 700, 1), ("He", 1e14, 900, 1.5),
             ("He", 4e14, 600, 1), ("Ne", 1e14, 800, 1.5), ("He", 3e14, 1000, 1)]
    reqs = [Request(rid=k, atom=a, intensity=I, wavelength=w, n_cycles=c, N=48, r_max=30.0,
                    lmax=6, dt=0.1, arrival=float(arr[k])) for k, (a, I, w, c) in enumerate(specs)]
    srv = Server(slots=2, chunk=64, ring=4096)
    mixed, stats = srv.serve(reqs, realtime=False)
    srv.close()
    print(stats, flush=True)
    assert stats["served"] == len(reqs)
    for r in reqs:
        alone_srv = Server(slots=1, chunk=64, ring=4096)
        alone, _ = alone_srv.serve([Request(**{**r.__dict__, "arrival": 0.0})], realtime=False)
        alone_srv.close()
        m = mixed[r.rid]
        a = alone[0]
        assert m.acc.shape == a.acc.shape
        d = np.abs(m.acc - a.acc).max()
        print("rid", r.rid, "max |acc mix - alone|", d, flush=True)
        assert d <= 1e-12 * max(1.0, np.abs(a.acc).max())
        assert np.abs(m.norm - a.norm).max() <= 1e-12


def _same_as_alone(reqs, make_server, tol):
    srv = make_server(2)
    mixed, stats = srv.serve(reqs, realtime=False)
    srv.close()
    print(stats, flush=True)
    assert stats["served"] == len(reqs)
    for r in reqs:
        alone_srv = make_server(1)
        alone, _ = alone_srv.serve([Request(**{**r.__dict__, "arrival": 0.0})], realtime=False)
        alone_srv.close()
        m, a = mixed[r.rid], alone[0]
        assert a.error is None and m.error is None
        assert m.acc.shape == a.acc.shape
        d = np.abs(m.acc - a.acc).max()
        print("rid", r.rid, "N", r.N, "max |acc mix - alone|", d, flush=True)
        assert d <= tol * max(1.0, np.abs(a.acc).max())
        assert np.abs(m.norm - a.norm).max() <= tol


def test_uneven_grids_equal_alone():
    # one serve call over two grids (two lanes of different N and L_max) and
    # uneven step counts, FP64
    print("case: serving over two grids, He N=48 Lmax=6 and N=64 Lmax=8", flush=True)
    specs = [(48, 30.0, 6, 2e14, 800, 1), (64, 40.0, 8, 3e14, 700, 1.5), (48, 30.0, 6, 1e14, 900, 1.5),
             (64, 40.0, 8, 2e14, 1000, 1)]
    reqs = [Request(rid=k, atom="He", intensity=I, wavelength=w, n_cycles=c, N=n, r_max=R, lmax=L,
                    dt=0.1, arrival=0.0) for k, (n, R, L, I, w, c) in enumerate(specs)]
    _same_as_alone(reqs, lambda slots: Server(slots=slots, chunk=64, ring=4096), 1e-12)


def test_fp32_mixed_equals_alone():
    # the FP32 mixed form is column-local too: with the kinetic variant held
    # fixed (the autotuner may choose differently for one slot and for two),
    # a request gives the same numbers alone and beside another
    import functools
    from ggbtddft_amd import engine

    class NarrowFP32(engine.Propagator):
        def autotune(self, reps=10, cache=True):
            self.set_kinetic_variant("narrow")
            return "narrow", {}

    print("case: FP32 mixed serving vs alone, He and Ne N=48 Lmax=6", flush=True)
    specs = [("He", 2e14, 800, 1), ("Ne", 3e14, 700, 1), ("He", 1e14, 900, 1.5), ("Ne", 2e14, 1000, 1)]
    reqs = [Request(rid=k, atom=a, intensity=I, wavelength=w, n_cycles=c, N=48, r_max=30.0,
                    lmax=6, dt=0.1, arrival=0.0) for k, (a, I, w, c) in enumerate(specs)]
    factory = functools.partial(NarrowFP32, precision=1)
    _same_as_alone(reqs, lambda slots: Server(slots=slots, chunk=64, ring=4096, factory=factory),
                   1e-12)
