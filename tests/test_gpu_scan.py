"""Groups of contexts replayed by one graph launch (ggb_group_run) and the
batched-scan driver built on them (scan.ScanGroup): a group's results equal
each context run alone, bit for bit, across whole graphs, the plain launches
after them and a second call; the scan driver's results equal the serving
loop's for the same requests."""
import sys

import numpy as np
import pytest

from ggbtddft_amd.system import AtomSystem

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_group_equals_each_context_alone():
    _say("case: group of He N=48 Lmax=6 (1 instance) and Ne N=64 Lmax=8 (2 instances), 8-step graphs")
    from ggbtddft_amd import engine
    dt = 0.05
    specs = [(AtomSystem("He", 48, 30.0, 6, dt), 1), (AtomSystem("Ne", 64, 40.0, 8, dt), 2)]
    t = (np.arange(80) + 0.5) * dt
    fields = [np.stack([0.05 * np.sin(0.4 * t + k) for k in range(n)]) for _, n in specs]
    alone = []
    for (S, n), E in zip(specs, fields):
        with engine.Propagator(S, n_inst=n) as prop:
            prop.set_field(E)
            alone.append([prop.run(0, 20), prop.run(20, 11)])
    props = [engine.Propagator(S, n_inst=n) for S, n in specs]
    try:
        for prop, E in zip(props, fields):
            prop.set_field(E)
        grp = engine.Group(props, steps=8)
        got = [grp.run([0, 0], 20), grp.run([20, 20], 11)]   # 2 graphs + 4 steps, then 1 + 3
        grp.close()
    finally:
        for prop in props:
            prop.close()
    for call in range(2):
        for k in range(2):
            for key in ("dip", "acc", "norm"):
                assert np.array_equal(got[call][k][key], alone[k][call][key]), (call, k, key)


def test_scan_group_equals_serving():
    _say("case: batched scan driver vs serving loop, He N=48/Ne N=64, 5 requests")
    from ggbtddft_amd.scan import ScanGroup
    from ggbtddft_amd.serving import Request, Server
    specs = [("He", 48, 30.0, 6, 2e14, 800, 1), ("Ne", 64, 40.0, 8, 3e14, 700, 1.5),
             ("He", 48, 30.0, 6, 1e14, 900, 1.5), ("Ne", 64, 40.0, 8, 2e14, 1000, 1),
             ("He", 48, 30.0, 6, 4e14, 600, 1)]
    reqs = [Request(rid=k, atom=a, intensity=I, wavelength=w, n_cycles=c, N=n, r_max=R, lmax=L,
                    dt=0.1) for k, (a, n, R, L, I, w, c) in enumerate(specs)]
    srv = Server(slots=3, chunk=64, ring=4096)
    ref, _ = srv.serve(reqs, realtime=False)
    systems = dict(srv.systems)
    srv.close()
    res, stats = ScanGroup(systems, slots=3, chunk=64, ring=4096).run(reqs)
    _say(stats)
    assert stats["served"] == len(reqs)
    for a, b in zip(res, ref):
        assert a.error is None and b.error is None
        d = np.abs(a.acc - b.acc).max()
        _say("rid", a.request.rid, "max |scan - serving|", d)
        assert d <= 1e-12 * max(1.0, np.abs(b.acc).max())
        assert np.abs(a.norm - b.norm).max() <= 1e-12
