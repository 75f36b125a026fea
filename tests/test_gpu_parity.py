"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Tolerances (FP64 end to end): sub-steps 1e-12 relative to the state norm,
short runs 1e-10 relative L2 on the dipole acceleration and 1e-12 absolute on
the orbital norms.
"""
import sys

import numpy as np
import pytest

from oracle import propagate as OP, field as OF
from ggbtddft_amd.system import AtomSystem

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine():
    from ggbtddft_amd import engine
    return engine


def _pair(atom, N, R, lmax, dt, **kw):
    S = AtomSystem(atom, N, R, lmax, dt, **kw)
    O = OP.System(atom, N, R, lmax, dt, **kw)
    return S, O


def _random_state(S, seed=3, scale=1.0):
    rng = np.random.default_rng(seed)
    psi = S.psi0.copy()
    noise = rng.normal(size=psi.shape) + 1j * rng.normal(size=psi.shape)
    for k in range(S.n_orb):
        noise[:S.m[k], k, :] = 0.0
    noise *= np.exp(-S.grid.r / 5.0)[None, None, :]
    psi = psi + 0.05 * scale * noise
    for k in range(S.n_orb):
        psi[:, k, :] /= np.sqrt(np.sum(np.abs(psi[:, k, :]) ** 2))
    return psi


def test_mfma_lane_layout():
    _say("case: mfma lane layout")
    out = _engine().selftest_mfma(0)
    i = np.arange(16)[:, None]
    k = np.arange(4)[None, :]
    A = i + 0.25 * k
    B = 0.5 * np.arange(4)[:, None] - np.arange(16)[None, :]
    np.testing.assert_allclose(out[0], A @ B, rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[1], A @ B, rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("atom,N,R,lmax", [("He", 32, 30.0, 4), ("Ar", 48, 40.0, 6)])
def test_substeps(atom, N, R, lmax):
    _say("case: substeps", atom, N, lmax)
    dt = 0.05
    S, O = _pair(atom, N, R, lmax, dt)
    psi = _random_state(S)
    E = np.array([0.08, -0.05])
    with _engine().Propagator(S) as prop:
        prop.set_field(E)
        prop.upload_orbitals(psi[:, :, :])
        # potential of the uploaded state
        dv = prop.download_potential()[0]
        dv_ref = O.potential(O.theta_state(psi))
        np.testing.assert_allclose(dv, dv_ref, rtol=0, atol=1e-11 * max(1.0, np.abs(dv_ref).max()))
        # first half: Vh, Le, Lo
        prop.substep(0, step=1)
        ref = psi.copy()
        O.v_half(ref, dv_ref)
        O.laser(ref, E[1], 0)
        O.laser(ref, E[1], 1)
        got = prop.download_orbitals()
        assert np.abs(got - ref).max() < 1e-12, np.abs(got - ref).max()
        # kinetic
        prop.substep(1, step=1)
        O.kinetic(ref)
        got = prop.download_orbitals()
        assert np.abs(got - ref).max() < 1e-12, np.abs(got - ref).max()
        # second half: Lo, Le, Vh
        prop.substep(2, step=1)
        O.laser(ref, E[1], 1)
        O.laser(ref, E[1], 0)
        O.v_half(ref, dv_ref)
        got = prop.download_orbitals()
        assert np.abs(got - ref).max() < 1e-12, np.abs(got - ref).max()


@pytest.mark.parametrize("atom,N,R,lmax,nsteps", [("He", 64, 40.0, 8, 60), ("Ar", 64, 40.0, 8, 30),
                                                   ("Kr", 64, 40.0, 8, 20), ("Xe", 64, 40.0, 10, 12),
                                                   ("He", 32, 30.0, 96, 10), ("Ne", 44, 40.0, 8, 20)])
def test_short_run(atom, N, R, lmax, nsteps):
    # Kr: 13 orbitals with |m| up to 2; Xe: 4d shells, three |m| groups; He at
    # Lmax = 96: the FP64 angular tables exceed the LDS budget and are read
    # from global memory; Ne at N = 44: the last 8-node tile of the angular
    # kernel holds 4 nodes past N (copied in as node N-1, not stored)
    _say("case: short run", atom, N, lmax, nsteps)
    dt = 0.05
    S, O = _pair(atom, N, R, lmax, dt)
    w = OF.omega_from_wavelength(800.0)
    E0 = OF.e0_from_intensity(5e14)
    Em = OF.midpoint_series(E0, w, 1, dt)[:nsteps] + 0.05  # strong, nonzero from step 0
    ref = O.run(Em)
    with _engine().Propagator(S) as prop:
        prop.set_field(Em)
        out = prop.run(0, nsteps)
        psi = prop.download_orbitals()
    for key in ("dip", "acc", "norm"):
        got = out[key][:, 0, :]
        exp = ref[key]
        rel = np.linalg.norm(got - exp) / np.linalg.norm(exp)
        _say(key, "rel", rel)
        assert rel < 1e-10, (key, rel)
    assert np.abs(out["norm"][:, 0, :] - ref["norm"]).max() < 1e-12
    assert np.abs(psi - ref["psi"]).max() < 1e-10


@pytest.mark.parametrize("propagator,precision", [("cn", 0), ("cn", 1), ("cn2", 0), ("cn3", 0)])
def test_crank_nicolson_propagator(propagator, precision):
    """The reference's Crank-Nicolson kinetic propagators (CN, CN2, CN3): the
    device applies e^{-i lam' dt} with lam' = 2 arg(p(lam dt/2))/dt; against the
    oracle's run of the same form (He on the helium configuration's grid, 60
    steps): FP64 to 1e-10 / norms 1e-12, the FP32 mixed form to 1e-4 / 1e-7."""
    _say("case: %s propagator, He N=96 Lmax=12, precision %d" % (propagator, precision))
    dt, nsteps = 0.05, 60
    S, O = _pair("He", 96, 60.0, 12, dt, propagator=propagator)
    w = OF.omega_from_wavelength(800.0)
    Em = OF.midpoint_series(OF.e0_from_intensity(2e14), w, 2, dt)[:nsteps] + 0.02
    ref = O.run(Em)
    with _engine().Propagator(S, precision=precision) as prop:
        prop.set_field(Em)
        out = prop.run(0, nsteps)
    tol, ntol = (1e-10, 1e-12) if precision == 0 else (1e-4, 1e-7)
    for key in ("dip", "acc"):
        rel = np.linalg.norm(out[key][:, 0, :] - ref[key]) / np.linalg.norm(ref[key])
        _say(key, "rel", rel)
        assert rel < tol, (key, rel)
    assert np.abs(out["norm"][:, 0, :] - ref["norm"]).max() < ntol
    # and it is not the exponential propagator's run
    ref_exp = OP.System("He", 96, 60.0, 12, dt, gs=O.gs).run(Em)
    assert np.linalg.norm(ref_exp["acc"] - ref["acc"]) > 1e-6 * np.linalg.norm(ref["acc"])


@pytest.mark.parametrize("n_inst", [3, 5])
def test_batched_instances_match_alone(n_inst):
    """Instances with different fields in one context give the same results
    as each instance alone (per-column arithmetic does not depend on the mix)."""
    _say("case: batched vs alone, Ne N=48 Lmax=6, %d instances" % n_inst)
    dt = 0.05
    S = AtomSystem("Ne", 48, 30.0, 6, dt)
    nsteps = 25
    t = (np.arange(nsteps) + 0.5) * dt
    fields = np.stack([0.05 * np.sin(0.4 * t), -0.08 * np.cos(0.3 * t), 0.02 + 0.0 * t,
                       0.06 * np.sin(0.2 * t + 1.0), -0.03 + 0.0 * t])[:n_inst]
    eng = _engine()
    with eng.Propagator(S, n_inst=n_inst) as prop:
        prop.set_field(fields)
        mix = prop.run(0, nsteps)
        psi_mix = prop.download_orbitals()
    for k in range(n_inst):
        with eng.Propagator(S, n_inst=1) as prop:
            prop.set_field(fields[k])
            alone = prop.run(0, nsteps)
            psi_alone = prop.download_orbitals()
        # the batched context uses the tiled kinetic kernels (one wave runs
        # the whole contraction), the single one the narrow kernels (the
        # contraction split over four waves): the summation order differs, so
        # results agree to rounding, 1e-12 of each quantity's scale.  The
        # scale of <cos(theta)/r^2> is that of its terms, <1/r^2> ~ 2 Z^2 for a
        # 1s shell, which cancel to the small dipole response
        scale = dict(dip=np.abs(alone["dip"]).max(), acc=2.0 * S.atom.Z ** 2,
                     norm=np.abs(alone["norm"]).max())
        for key in ("dip", "acc", "norm"):
            d = np.abs(mix[key][:, k, :] - alone[key][:, 0, :]).max()
            assert d <= 1e-12 * scale[key], (k, key, d)
        cols = slice(k * S.n_orb, (k + 1) * S.n_orb)
        assert np.abs(psi_mix[:, cols, :] - psi_alone).max() < 1e-12


@pytest.mark.parametrize("atom,N,R,lmax,nsteps", [("He", 64, 40.0, 8, 60), ("Ar", 64, 40.0, 8, 30),
                                                   ("He", 32, 30.0, 96, 10), ("Ne", 40, 40.0, 8, 20)])
def test_fp32_short_run(atom, N, R, lmax, nsteps):
    """Mixed precision (FP32 tables/transforms/v_xc, FP64 state and sums)
    against the FP64 oracle: FP32-level agreement.  Ne at N = 40: the last
    16-node tile of the FP32 angular kernel holds 8 nodes past N."""
    _say("case: fp32 short run", atom, N, lmax, nsteps)
    dt = 0.05
    S, O = _pair(atom, N, R, lmax, dt)
    w = OF.omega_from_wavelength(800.0)
    Em = OF.midpoint_series(OF.e0_from_intensity(5e14), w, 1, dt)[:nsteps] + 0.05
    ref = O.run(Em)
    with _engine().Propagator(S, precision=1) as prop:
        prop.set_field(Em)
        out = prop.run(0, nsteps)
    # <cos/r^2> per orbital cancels between the hemispheres next to the
    # nucleus, where FP32 rounding of |Psi|^2 is largest: looser bound
    for key, tol in (("dip", 1e-4), ("acc", 2e-3)):
        rel = np.linalg.norm(out[key][:, 0, :] - ref[key]) / np.linalg.norm(ref[key])
        _say("fp32", key, "rel", rel)
        assert rel < tol, (key, rel)
    assert np.abs(out["norm"][:, 0, :] - ref["norm"]).max() < 1e-5


def test_potential_panel_form_equals_small_tiles():
    """From 8 instances on, the multipole moments and dV use the panel form of
    the tile product (each wave keeps three tiles over one operand panel, with
    the same products in the same order as the 16 x 16 form).  The potentials
    of a 9-instance context equal those of each instance alone (the small-tile
    form), bit for bit, at a grid with two tile groups and partial tiles
    (He N=32 L_max=40: 81 multipoles and theta nodes)."""
    _say("case: panel-form potential products, He N=32 Lmax=40, 9 instances vs alone")
    S = AtomSystem("He", 32, 30.0, 40, 0.05)
    n = 9
    psis = [_random_state(S, seed=10 + k, scale=1.0 + 0.5 * k) for k in range(n)]
    eng = _engine()
    with eng.Propagator(S, n_inst=n) as prop:
        prop.upload_orbitals(np.concatenate(psis, axis=1))
        pot = prop.download_potential()
    for k in range(n):
        with eng.Propagator(S, n_inst=1) as prop:
            prop.upload_orbitals(psis[k])
            alone = prop.download_potential()
        assert np.array_equal(pot[k], alone[0]), k
    assert np.abs(pot[0] - pot[1]).max() > 1e-6  # the instances differ


def test_graph_replay_matches_eager():
    """hipGraph replay (3 replays of 16 steps + 5 eager steps, eager work and
    synchronisations in between) reproduces plain launches bit for bit."""
    _say("case: graph replay vs eager, Ne N=48 Lmax=6")
    dt = 0.05
    S = AtomSystem("Ne", 48, 30.0, 6, dt)
    t = (np.arange(120) + 0.5) * dt
    E = 0.05 * np.sin(0.4 * t)
    eng = _engine()
    with eng.Propagator(S) as prop:
        prop.set_field(E)
        ref = [prop.run(0, 53)]
        ref.append(prop.run(53, 53))
        psi_ref = prop.download_orbitals()
    with eng.Propagator(S) as prop:
        prop.set_field(E)
        prop.set_graph(16)
        got = [prop.run(0, 53)]
        prop.download_potential()          # eager work and a sync between replays
        got.append(prop.run(53, 53))
        psi = prop.download_orbitals()
    for a, b in zip(ref, got):
        for key in ("dip", "acc", "norm"):
            assert np.array_equal(a[key], b[key]), key
    assert np.array_equal(psi, psi_ref)


@pytest.mark.parametrize("precision,N", [(0, 48), (1, 48), (1, 40), (0, 44)])
def test_kinetic_variants_agree_and_autotune(precision, N):
    """Every kinetic kernel variant gives the same run (to rounding: they sum
    K in different orders); the autotuner picks one of them and leaves the
    orbitals as they were.  N = 40 is not a multiple of the 16-row chunks of
    the FP32 tiled pass 1 through direct global-to-LDS copies (it runs the
    register-staged form there) and leaves 24 clamped rows in the last radial
    tile of pass 2.  N = 44 leaves 4 clamped nodes in the last radial tile of
    the FP64 angular kernel, whose l-space tile is copied straight into LDS."""
    _say("case: kinetic variants, Ne N=%d Lmax=6, 3 instances, precision %d" % (N, precision))
    dt = 0.05
    S = AtomSystem("Ne", N, 30.0, 6, dt)
    nsteps = 20
    t = (np.arange(nsteps) + 0.5) * dt
    fields = np.stack([0.05 * np.sin(0.4 * t), -0.08 * np.cos(0.3 * t), 0.02 + 0.0 * t])
    eng = _engine()
    variants = ["narrow", "tiled16"] + (["tiled32"] if precision else [])
    runs = {}
    for v in variants:
        with eng.Propagator(S, n_inst=3, precision=precision) as prop:
            prop.set_field(fields)
            prop.set_kinetic_variant(v)
            runs[v] = (prop.run(0, nsteps), prop.download_orbitals())
    ref_obs, ref_psi = runs["narrow"]
    tol = 1e-12 if precision == 0 else 1e-7  # FP32 products: rounding of the small remainder
    for v in variants[1:]:
        obs, psi = runs[v]
        assert np.abs(psi - ref_psi).max() < tol, v
        assert np.abs(obs["dip"] - ref_obs["dip"]).max() < tol * max(1.0, np.abs(ref_obs["dip"]).max()), v
    with eng.Propagator(S, n_inst=3, precision=precision) as prop:
        prop.set_field(fields)
        before = prop.download_orbitals()
        name, times = prop.autotune(reps=2, cache=False)
        assert name in variants and set(times) == set(variants)
        assert np.array_equal(prop.download_orbitals(), before)


@pytest.mark.parametrize("precision,n_inst", [(0, 1), (1, 12)])
def test_timing_modes(precision, n_inst):
    """Timing mode 1 stamps the two kinetic product kernels of every 16th step
    (events from the kernels' own dispatch), mode 2 every phase of every step;
    neither changes the results."""
    _say("case: timing modes, Ne N=48 Lmax=6, precision %d, %d instance(s)" % (precision, n_inst))
    dt = 0.05
    S = AtomSystem("Ne", 48, 30.0, 6, dt)
    t = (np.arange(32) + 0.5) * dt
    E = np.stack([0.05 * np.sin(0.4 * t)] * n_inst) if n_inst > 1 else 0.05 * np.sin(0.4 * t)
    eng = _engine()
    outs = []
    for mode in (0, 1, 2):
        with eng.Propagator(S, n_inst=n_inst, precision=precision) as prop:
            prop.set_field(E)
            prop.set_timing(mode)
            outs.append(prop.run(0, 32))
            tim = prop.timing()
        if mode == 1:
            assert tim["kinetic_qt"][1] == 2 and tim["kinetic_q"][1] == 2
            assert tim["kinetic_qt"][0] > 0.0 and tim["kinetic_q"][0] > 0.0
            assert all(tim[k][1] == 0 for k in ("density_moments", "hartree", "dv_phase"))
        if mode == 2:
            assert all(v[1] == 32 and v[0] > 0.0 for v in tim.values())
    for o in outs[1:]:
        for key in ("dip", "acc", "norm"):
            assert np.array_equal(o[key], outs[0][key]), key


@pytest.mark.parametrize("precision", [0, 1])
def test_idle_instances_skipped(precision):
    """Instances marked idle (ggb_set_active) skip the angular and potential
    kernels; the instances in use give exactly the results of a run with every
    instance in use."""
    _say("case: idle instance slots, Ne N=48 Lmax=6, 3 instances, precision %d" % precision)
    dt = 0.05
    S = AtomSystem("Ne", 48, 30.0, 6, dt)
    t = (np.arange(20) + 0.5) * dt
    fields = np.stack([0.05 * np.sin(0.4 * t), -0.08 * np.cos(0.3 * t), 0.02 + 0.0 * t])
    eng = _engine()
    runs = []
    for mask in ([1, 1, 1], [1, 0, 1]):
        with eng.Propagator(S, n_inst=3, precision=precision) as prop:
            prop.set_field(fields)
            prop.set_active(mask)
            runs.append((prop.run(0, 20), prop.download_orbitals()))
    (full, psi_full), (part, psi_part) = runs
    for key in ("dip", "acc", "norm"):
        assert np.array_equal(full[key][:, [0, 2]], part[key][:, [0, 2]]), key
    for k in (0, 2):
        cols = slice(k * S.n_orb, (k + 1) * S.n_orb)
        assert np.array_equal(psi_full[:, cols], psi_part[:, cols])


def test_channel_dipoles_match_oracle():
    """Per-(orbital, l) channel dipoles (ggb_set_channels, the reference's
    dipoleMoment.cuh) against the oracle, summed over l against dip, and the
    graph replay path against plain launches bit for bit."""
    _say("case: channel dipoles, Ar N=48 Lmax=6, 21 steps, 2 instances")
    dt = 0.05
    S, O = _pair("Ar", 48, 40.0, 6, dt)
    w = OF.omega_from_wavelength(800.0)
    Em = OF.midpoint_series(OF.e0_from_intensity(5e14), w, 1, dt)[:21] + 0.05
    ref = O.run(Em, channels=True)
    eng = _engine()
    E2 = np.stack([Em, 0.5 * Em])
    with eng.Propagator(S, n_inst=2) as prop:
        prop.set_field(E2)
        out = prop.run(0, 21, channels=True)
    ch = out["chan"][:, 0]
    rel = np.linalg.norm(ch - ref["chan"]) / np.linalg.norm(ref["chan"])
    _say("chan rel", rel)
    assert rel < 1e-10
    np.testing.assert_allclose(out["chan"].real.sum(axis=-1), out["dip"], rtol=0,
                               atol=1e-12 * np.abs(out["dip"]).max())
    with eng.Propagator(S, n_inst=2) as prop:
        prop.set_field(E2)
        prop.set_graph(8)
        got = prop.run(0, 21, channels=True)
    assert np.array_equal(got["chan"], out["chan"])
    assert np.array_equal(got["dip"], out["dip"])


def test_compensated_split_kinetic_product():
    """The kinetic sub-step in the FP32 mixed form against the FP64 oracle on
    a perturbed ground state: plain FP32 tiles (precision 1) and the
    compensated split-FP32 product (precision 2, hi/lo operands, FP64 chunk
    sums).  The split product removes the rounding of the operands; what is
    left is the FP32 rounding of each matrix-core result inside a 16-value
    chunk, which no split of the operands reaches.  Measured: 2.25e-8 plain,
    1.65e-8 split (of max |psi|)."""
    _say("case: kinetic sub-step, plain FP32 tiles vs split-FP32 product, Ar N=48 Lmax=6")
    dt = 0.05
    S, O = _pair("Ar", 48, 40.0, 6, dt)
    psi = _random_state(S)
    ref = psi.copy()
    O.kinetic(ref)
    scale = np.abs(ref).max()
    errs = {}
    for precision, variant in ((1, "tiled16"), (2, "split16")):
        with _engine().Propagator(S, precision=precision) as prop:
            prop.set_field(np.zeros(2))
            prop.set_kinetic_variant(variant)
            prop.upload_orbitals(psi)
            prop.substep(1, step=0)
            got = prop.download_orbitals()
        errs[variant] = np.abs(got - ref).max() / scale
        _say(variant, "max error / max |psi|", errs[variant])
    assert errs["tiled16"] < 1e-7
    assert errs["split16"] < 0.85 * errs["tiled16"]
