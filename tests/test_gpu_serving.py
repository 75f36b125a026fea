"""Serving on the GPU: requests seated into running lanes (continuous
batching, two lanes overlapping on the device) give the same results as each
request propagated alone, within 1e-12."""
import sys

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
