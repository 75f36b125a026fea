"""Benchmark: orbital-timesteps/s of the FP64 propagator on one MI355X.

Workload (the metric's configuration): argon all-electron, 7 orbitals grouped
by |m| (m=0: 1s 2s 2p0 3s 3p0, |m|=1: 2p1 3p1 at double weight), N = 512
Gauss-Lobatto radial nodes to r_max = 200, Lmax = 64, Hartree multipoles to
L = 128 on 129 theta nodes, 800 nm sin^2 pulse of 8 cycles at 2e14 W/cm^2,
dt = 0.02 (44128 steps per full run), FP64 throughout.  A "step" is one time
step of the whole split-operator update (potential half steps, laser, kinetic
propagator, potential rebuild, observables) for the instance.

value = orbital-timesteps/s = 7 * K / (wall time of K steps), state and
tables resident in HBM; the K steps are bracketed by stream synchronisation.
The dominant kernel (the kinetic propagator pass) is timed with HIP events on
its stream inside the timed region for the roofline entry.  The CPU baseline
is the FP64 oracle on the host cores for a bounded window of the same
workload.  Further entries of the JSON line: the batched FP32 leg (64
instances, ratio to the FP64 value), the serving leg (p50/p99 time to
spectrum), the device spectrum and ground state against their host
counterparts, and the precision report (full-length FP32 runs, one instance
and batched, against FP64: harmonic-intensity errors).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
ATOM, N, R_MAX, LMAX, DT = "Ar", 512, 200.0, 64, 0.02
INTENSITY, WAVELENGTH, CYCLES = 2e14, 800.0, 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(system_args, E, nsteps=100, warm=3, nsteps_1core=40):
    """FP64 oracle on the host for a bounded window (setup excluded): all
    cores (BLAS threads = min(16, nproc)), then one core on a shorter window."""
    import contextlib
    from oracle import propagate as OP

    def limited(n):
        try:
            from threadpoolctl import threadpool_limits
            return threadpool_limits(limits=n)
        except Exception:  # pragma: no cover
            return contextlib.nullcontext()

    cores = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    S = OP.System(*system_args)
    setup = time.perf_counter() - t0
    rates = {}
    for n, k in ((cores, nsteps), (1, nsteps_1core)):
        with limited(n):
            res = S.run(E[:warm], psi=S.psi0.copy())
            t0 = time.perf_counter()
            S.run(E[warm:warm + k], psi=res["psi"])
            el = time.perf_counter() - t0
        rates[n] = (S.norb * k / el, k, el)
        log("cpu baseline: %d steps on %d core(s) in %.1f s" % (k, n, el))
    v, k, el = rates[cores]
    v1, k1, el1 = rates[1]
    # the baseline is the faster of the two (numpy's threads do not help this
    # oracle, and on a shared host they can slow it down); both are reported
    best = cores if v >= v1 else 1
    vb, kb, elb = rates[best]
    return dict(value=vb, unit="orbital-timesteps/s", cores=best, kind="port",
                sample="oracle (numpy FP64), Ar N=512 Lmax=64, %d steps after %d warm-up "
                       "steps on %d thread(s), %.1f s (setup %.1f s excluded)"
                       % (kb, warm, best, elb, setup),
                value_all_cores=v, cores_all=cores,
                sample_all_cores="same, %d threads, %d steps, %.1f s" % (cores, k, el),
                value_1core=v1,
                sample_1core="same, one thread, %d steps, %.1f s" % (k1, el1),
                host_cpu=_cpu_model(), host_cpus_visible=os.cpu_count(),
                host_cpus_usable=len(os.sched_getaffinity(0)))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


FP64_MFMA_TFLOPS = 78.6
FP32_MFMA_TFLOPS = 157.3


def phase_rooflines(sysm, n_inst, precision, phases_us, tiled):
    """Algorithmic bytes (and matrix-core flops) of each timed phase of one
    step and the fraction of the MI355X peak they reach in the measured time
    (phase times from HIP events between the phases: launch gaps included)."""
    N, Nl, Nth, nh = sysm.N, sysm.n_l, sysm.n_theta, sysm.n_har
    C = n_inst * sysm.n_orb
    es = 4 if precision else 8
    nk = int(sysm.nkeep.sum())
    h = (Nth + 1) // 2
    state = Nl * C * N * 16                      # FP64 state, read or written once
    phase_tab = n_inst * N * Nth * (8 if precision else 16)
    ang_flops = 2 * 2 * h * Nl * 2 * C * N       # one l<->theta transform pair (parity-split)
    kin_flops = 2 * nk * 2 * C * N
    kin_bytes = nk * N * es + state + nk * C * 2 * es
    work = {
        "half_potential_laser": (2 * state + phase_tab + (Nl * C * N * 8 if precision and tiled else 0),
                                 ang_flops),
        "kinetic_qt": (kin_bytes, kin_flops),
        "kinetic_q": (kin_bytes, kin_flops),
        "laser_half_potential_obs": (2 * state + phase_tab + C * N * Nth * es, 1.5 * ang_flops),
        "density_moments": (C * N * Nth * es + n_inst * N * Nth * 8 + n_inst * nh * N * 8,
                            2 * nh * Nth * N * n_inst),
        "hartree": (2 * n_inst * nh * N * 8 + (2 * nh + 1) * N * 8, 0),
        "dv_phase": (n_inst * nh * N * 8 + 2 * n_inst * N * Nth * 8 + phase_tab,
                     2 * nh * Nth * N * n_inst),
    }
    out = {}
    for k, (b, f) in work.items():
        t = phases_us.get(k)
        if not t:
            continue
        peak_tf = FP32_MFMA_TFLOPS if (precision and k in ("kinetic_qt", "kinetic_q",
                                                             "half_potential_laser",
                                                             "laser_half_potential_obs")) \
            else FP64_MFMA_TFLOPS
        out[k] = {"us": t, "bytes": int(b), "GB/s": b / (t * 1e-6) / 1e9,
                  "hbm_frac": b / (t * 1e-6) / 1e9 / HBM_PEAK_GBS,
                  "flops": float(f), "TFLOP/s": f / (t * 1e-6) / 1e12,
                  "mfma_frac": f / (t * 1e-6) / 1e12 / peak_tf}
    return out


def batched_run(sysm, pulse, args, fp64_value, precision=1):
    """n Ar instances (intensities log-spaced 1e14..3e14) in one context, FP32
    mixed form (precision 1) or FP64 (0)."""
    from ggbtddft_amd.engine import Propagator
    from ggbtddft_amd.pulse import Pulse
    n = args.batched
    Is = np.geomspace(1e14, 3e14, n)
    nst = args.warmup + args.batched_steps + 100
    E = np.stack([Pulse(I, WAVELENGTH, CYCLES).midpoints(DT)[:nst] for I in Is])
    prop = Propagator(sysm, n_inst=n, precision=precision)
    prop.set_field(E)
    tuned = prop.autotune(reps=5)
    name = {0: "fp64", 1: "fp32", 2: "fp32 split"}[precision]
    log("autotuner (batched %s): %s %s" % (name, tuned[0], tuned[1]))
    prop.run(0, args.warmup)
    prop.set_timing(1)
    t0 = time.perf_counter()
    obs = prop.run(args.warmup, args.batched_steps)
    el = time.perf_counter() - t0
    tim = prop.timing()
    prop.set_timing(2)
    prop.run(args.warmup + args.batched_steps, 50)
    phases = {k: 1e3 * v[0] / max(v[1], 1) for k, v in prop.timing().items()}
    prop.close()
    assert np.all(np.isfinite(obs["acc"]))
    val = n * sysm.n_orb * args.batched_steps / el
    log("batched %s: %d instances, %d steps in %.2f s" % (name, n, args.batched_steps, el))
    tiled = tuned[0] in ("tiled16", "tiled32", "split16")
    kin_us = 1e3 * tim["kinetic_qt"][0] / max(tim["kinetic_qt"][1], 1)
    C = n * sysm.n_orb
    # matrix-core flops issued (the split product issues three FP32 products)
    kin_flops = 2 * int(sysm.nkeep.sum()) * 2 * C * sysm.N * (3 if precision == 2 else 1)
    peak = FP32_MFMA_TFLOPS if precision else FP64_MFMA_TFLOPS
    return {"value": val, "unit": "orbital-timesteps/s", "instances": n,
            "steps": args.batched_steps, "ms_per_step": 1e3 * el / args.batched_steps,
            "precision": {0: "fp64 throughout",
                          1: "fp32 tables/transforms/v_xc, fp64 state and sums",
                          2: "as fp32, kinetic product compensated (split-FP32: 3 FP32 "
                             "matrix-core products, FP64 chunk sums)"}[precision],
            "kinetic_variant": tuned[0], "autotuner_ms_per_kinetic_step": tuned[1],
            "ratio_vs_fp64_plain": val / fp64_value,
            "roofline_kinetic": {"kernel": "kinetic pass 1 (Y = e^{-i lam dt} Q^T psi), %s" % tuned[0],
                                 "bound": "mfma" if tiled else "hbm",
                                 "achieved": kin_flops / (kin_us * 1e-6) / 1e12, "peak": peak,
                                 "unit": "TFLOP/s",
                                 "frac": kin_flops / (kin_us * 1e-6) / 1e12 / peak,
                                 "flops_per_launch": kin_flops, "avg_launch_us": kin_us},
            "phase_us": phases,
            "phase_rooflines": phase_rooflines(sysm, n, precision, phases, tiled)}


def serving_requests(n, rate, seed=0):
    """The serving stream: Poisson arrivals (seed) of mixed requests over the
    grid range He N=256/L_max=32, Ne 384/48 (dt 0.05), Kr 768/96 (dt 0.02), one
    colour 800-2000 nm at 5e13-5e14 W/cm^2, and Ar 512/64 (dt 0.02) in situ:
    800 nm at 1.5e14 W/cm^2 with its second harmonic at 1e-3 relative
    intensity and a random relative phase; 2-cycle pulses."""
    from ggbtddft_amd.serving import Request, poisson_arrivals
    rng = np.random.default_rng(seed)
    arr = poisson_arrivals(rate, n, seed=seed)
    kinds = [("He", 256, 100.0, 32, 0.05, False), ("Ne", 384, 150.0, 48, 0.05, False),
             (ATOM, N, R_MAX, LMAX, DT, True), ("He", 256, 100.0, 32, 0.05, False),
             ("Ne", 384, 150.0, 48, 0.05, False), ("Kr", 768, 200.0, 96, 0.02, False)]
    reqs = []
    for k in range(n):
        I = float(np.exp(rng.uniform(np.log(5e13), np.log(5e14))))
        lam = float(rng.uniform(800.0, 2000.0))
        phase = float(rng.uniform(0.0, 2.0 * np.pi))
        atom, nr, rmax, lmax, dt, two = kinds[k % len(kinds)]
        if two:
            I, lam, second = 1.5e14, 800.0, (2, 1e-3, phase)
        else:
            second = None
        reqs.append(Request(rid=k, atom=atom, intensity=I, wavelength=lam, n_cycles=2, N=nr,
                            r_max=rmax, lmax=lmax, dt=dt, arrival=float(arr[k]), second=second))
    return reqs


def _serve(reqs, slots, chunk, systems=None, warm=True):
    from ggbtddft_amd.serving import Server
    srv = Server(slots=slots, chunk=chunk, ring=1 << 16, precision=1)
    t_sys = 0.0
    if systems:
        srv.systems.update(systems)
    if warm:
        t0 = time.perf_counter()
        for key in {r.key() for r in reqs}:   # ground states and tables before the clock starts
            srv._system(key)
        t_sys = time.perf_counter() - t0
    results, stats = srv.serve(reqs, realtime=True)
    systems_out = dict(srv.systems)
    srv.close()
    bad = [r for r in results if r.error is not None or not np.all(np.isfinite(r.spectrum))]
    stats = dict(stats)
    stats.update({"requests": len(reqs), "slots_per_lane": slots, "chunk_steps": chunk,
                  "precision": "fp32 mixed", "bad_results": len(bad),
                  "system_setup_s_before_clock": t_sys,
                  "precompute_inside_clock": not warm})
    return stats, systems_out


def serving_run(sysm_ar, args):
    """Continuous-batching serving (FP32 mixed form, arrivals in real time,
    p50/p99 arrival -> spectrum): the mixed serving stream with the new
    keys' ground states and tables computed inside the clock (cold) and
    before it (warm), and a saturated stream of in-situ argon requests on the
    benchmark grid that keeps every slot of its lane busy."""
    from ggbtddft_amd.serving import Request, poisson_arrivals
    out = {}
    reqs = serving_requests(args.serving, args.serving_rate)
    desc = ("Poisson stream (seed 0, %.1f/s): He N=256 Lmax=32, Ne N=384 Lmax=48 (dt=0.05) "
            "and Kr N=768 Lmax=96 (dt=0.02), one colour 800-2000 nm, 5e13-5e14 W/cm^2; "
            "Ar N=512 Lmax=64 (dt=0.02) in situ: 800 nm at 1.5e14 W/cm^2 + 400 nm at 1e-3 "
            "relative intensity, random relative phase; 2-cycle pulses" % args.serving_rate)
    cold, systems = _serve(reqs, args.serving_slots, args.serving_chunk,
                           systems={(ATOM, N, float(R_MAX), LMAX, float(DT)): sysm_ar}, warm=False)
    cold["workload"] = desc + " (ground states and tables of He, Ne, Kr computed inside the clock)"
    log("serving cold: p50 %.2f s, p99 %.2f s" % (cold["p50_time_to_spectrum_s"],
                                                  cold["p99_time_to_spectrum_s"]))
    warm, _ = _serve(reqs, args.serving_slots, args.serving_chunk, systems=systems, warm=True)
    warm["workload"] = desc + " (ground states and tables before the clock)"
    log("serving warm: p50 %.2f s, p99 %.2f s" % (warm["p50_time_to_spectrum_s"],
                                                  warm["p99_time_to_spectrum_s"]))
    out["mixed_stream_warm"] = warm
    out["mixed_stream_cold"] = cold
    if args.serving_saturated > 0:
        n, slots, rate = args.serving_saturated, args.serving_saturated_slots, 1000.0
        arr = poisson_arrivals(rate, n, seed=1)
        rng = np.random.default_rng(1)
        sat = [Request(rid=k, atom=ATOM, intensity=1.5e14, wavelength=800.0, n_cycles=2, N=N,
                       r_max=R_MAX, lmax=LMAX, dt=DT, arrival=float(arr[k]),
                       second=(2, 1e-3, float(rng.uniform(0.0, 2.0 * np.pi)))) for k in range(n)]
        st, _ = _serve(sat, slots, args.serving_chunk,
                       systems={(ATOM, N, float(R_MAX), LMAX, float(DT)): sysm_ar})
        st["workload"] = ("%d in-situ Ar requests (N=512 Lmax=64 dt=0.02, 800 nm 2 cycles at "
                          "1.5e14 W/cm^2 + 400 nm at 1e-3, random phase), Poisson arrivals at "
                          "%g/s, one lane of %d slots: the requests beyond the slots wait, so "
                          "every slot stays busy until the last wave drains" % (n, rate, slots))
        log("serving saturated: %.0f orbital-timesteps/s, p50 %.2f s, p99 %.2f s"
            % (st["orbital_timesteps_per_s"], st["p50_time_to_spectrum_s"],
               st["p99_time_to_spectrum_s"]))
        out["saturated_argon"] = st
    return out


def scan_run(n, cycle_scale=1.0):
    """The batched scan of n instances served as one batch (all requests at
    t = 0, one lane per grid key, each lane with a slot per request of its
    key, idle slots skipped once their request is done), FP64 then FP32 mixed;
    orbital-timesteps/s over the whole scan and the FP32 error per instance.
    cycle_scale < 1 shortens every pulse (same grids, uneven step counts)."""
    from ggbtddft_amd.scan import scan_requests
    from ggbtddft_amd.serving import Server
    reqs = scan_requests(n, cycle_scale=cycle_scale)
    keys = {}
    for r in reqs:
        keys[r.key()] = keys.get(r.key(), 0) + 1
    slots = max(keys.values())
    n_steps = [r.pulse().n_steps(r.dt) for r in reqs]
    t0 = time.perf_counter()
    systems = {}
    for key in keys:   # ground states and tables before the clock starts
        systems[key] = Server(slots=1)._system(key)
    t_sys = time.perf_counter() - t0
    out = {"instances": n, "lanes": len(keys), "slots_per_lane": slots,
           "steps_min": int(min(n_steps)), "steps_max": int(max(n_steps)),
           "system_setup_s_before_clock": t_sys,
           "workload": "batched scan (seed 0): He/Ne, 5e13-5e14 W/cm^2 log-uniform, "
                       "600-2000 nm, 4-16 cycles%s, grid by quiver radius "
                       "(N 256/384/512, L_max 32/48/64), dt 0.05, all requests at t = 0"
                       % ("" if cycle_scale == 1.0 else " scaled by %g" % cycle_scale)}
    res = {}
    from ggbtddft_amd.scan import ScanGroup
    for prec, name in ((0, "fp64"), (1, "fp32")):
        log("scan %s: %d requests on %d lanes, one graph launch per chunk" % (name, n, len(keys)))
        results, stats = ScanGroup(systems, slots=slots, chunk=256, ring=1 << 17,
                                   precision=prec).run(reqs)
        res[name] = results
        out[name] = {"wall_s": stats["wall_s"], "served": stats["served"],
                     "orbital_timesteps_per_s": stats["orbital_timesteps_per_s"],
                     "driver": "scan.ScanGroup: all lanes' chunks in one hipGraph launch"}
        log("scan %s: %d served in %.1f s, %.0f orbital-timesteps/s" % (
            name, stats["served"], stats["wall_s"], stats["orbital_timesteps_per_s"]))
    # the same scan through the serving loop (one host thread per lane)
    srv = Server(slots=slots, chunk=256, ring=1 << 17, precision=1)
    srv.systems.update(systems)
    _, stats = srv.serve(reqs, realtime=False)
    srv.close()
    out["fp32_serving_threads"] = {"wall_s": stats["wall_s"], "served": stats["served"],
                                   "orbital_timesteps_per_s": stats["orbital_timesteps_per_s"]}
    err = []
    for a, b in zip(res["fp64"], res["fp32"]):
        if a.error is None and b.error is None:
            err.append(float(np.linalg.norm(b.a_total - a.a_total) / np.linalg.norm(a.a_total)))
    out["ratio_fp32_vs_fp64"] = out["fp32"]["orbital_timesteps_per_s"] / \
        out["fp64"]["orbital_timesteps_per_s"]
    out["fp32_acc_rel_l2_median"] = float(np.median(err)) if err else None
    out["fp32_acc_rel_l2_max"] = float(np.max(err)) if err else None
    return out


def spectrum_timing(nfull):
    """Device vs host (numpy) spectrum of one full-length Ar acceleration
    series (window, padding to 17 n, FFT, |.|^2, 20 harmonic masks)."""
    from ggbtddft_amd import spectrum as spec
    rng = np.random.default_rng(0)
    x = rng.standard_normal(nfull)
    orders = np.arange(1, 41, 2, dtype=np.float64)
    w0 = 45.5633525 / WAVELENGTH
    spec.device_spectrum(x, DT, omega0=w0, orders=orders)  # warm-up (module load, first launch)
    t0 = time.perf_counter()
    _, p_d, i_d = spec.device_spectrum(x, DT, omega0=w0, orders=orders)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    w_h, p_h = spec.acceleration_spectrum(x, DT)
    i_h = spec.harmonic_intensities(w_h, p_h, w0, orders)
    t_host = time.perf_counter() - t0
    return {"samples": nfull, "fft_points": int(p_h.size), "device_ms": 1e3 * t_dev,
            "host_numpy_ms": 1e3 * t_host,
            "max_rel_diff_power": float(np.abs(p_d - p_h).max() / p_h.max()),
            "max_rel_diff_harmonics": float(np.abs(i_d / i_h - 1).max())}


def ground_state_timing():
    """LDA ground state of the benchmark grid on the device (ggb_ground_state)
    and by the host iteration (numpy, one dense eigendecomposition per l and
    iteration); returns the stats and the device ground state."""
    from ggbtddft_amd.atoms import Atom
    from ggbtddft_amd.grid import RadialGrid
    from ggbtddft_amd.groundstate import DeviceGroundState, GroundState
    DeviceGroundState(RadialGrid(48, 30.0), Atom("He"))  # warm-up (library load, first launches)
    grid, atom = RadialGrid(N, R_MAX), Atom(ATOM)
    t0 = time.perf_counter()
    dev = DeviceGroundState(grid, atom)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    host = GroundState(grid, atom)
    t_host = time.perf_counter() - t0
    stats = {"workload": "%s LDA ground state, N=%d r_max=%g, tol 1e-11" % (ATOM, N, R_MAX),
             "device_s": t_dev, "device_seconds": dev.seconds, "device_iterations": dev.iterations,
             "host_s": t_host, "host_iterations": host.iterations, "host_cores": os.cpu_count(),
             "max_abs_diff_eps": float(np.abs(dev.eps - host.eps).max()),
             "max_abs_diff_v_gs": float(np.abs(dev.v_gs - host.v_gs).max())}
    log("ground state: device %.3f s (%d it), host %.2f s (%d it)"
        % (t_dev, dev.iterations, t_host, host.iterations))
    return stats, dev


def measured_traffic():
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel
    calibration); None when the summary is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "pmc_traffic_argon.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def precision_report(sysm, pulse, n_batch=16):
    """Full-length runs: FP64 one instance, FP32 mixed one instance (narrow
    kinetic kernels), FP32 mixed in a batch of n_batch instances (tiled
    kinetic kernels, the configuration of the batched leg; instance 0 carries
    the benchmark pulse, the others 1e14-3e14 W/cm^2) and the same batch with
    the compensated split-FP32 kinetic product; errors of the FP32 spectra
    against FP64."""
    from ggbtddft_amd import spectrum as sp
    from ggbtddft_amd.engine import Propagator
    from ggbtddft_amd.pulse import Pulse
    n = pulse.n_steps(DT)
    E = pulse.midpoints(DT)
    E_end = pulse((np.arange(n) + 1.0) * DT)
    rng = np.random.default_rng(1)
    fields = [E] + [Pulse(float(I), WAVELENGTH, CYCLES).midpoints(DT)[:n]
                    for I in rng.uniform(1e14, 3e14, n_batch - 1)]
    runs = {}
    for name, prec, ninst in (("fp64", 0, 1), ("fp32", 1, 1), ("fp32_batched", 1, n_batch),
                              ("fp32_split_batched", 2, n_batch)):
        t0 = time.perf_counter()
        with Propagator(sysm, n_inst=ninst, precision=prec) as prop:
            prop.set_field(np.stack(fields[:ninst]) if ninst > 1 else E)
            # the batch takes the autotuner's pick, as the batched leg does; the
            # variants differ in their FP32 chunking, so the errors come with it
            variant = prop.autotune(reps=3)[0] if ninst > 1 else "narrow"
            o = prop.run(0, n)
        runs[name] = (sysm.total_acceleration(o["acc"][:, 0, :], o["norm"][:, 0, :], E_end),
                      o["dip"][:, 0, :] @ sysm.occ, time.perf_counter() - t0, variant)
        log("precision %s full run: %d steps, %d instance(s), %.1f s" % (name, n, ninst, runs[name][2]))
    a64, d64 = runs["fp64"][0], runs["fp64"][1]
    w, S64 = sp.acceleration_spectrum(a64, DT)
    wd, D64 = sp.dipole_spectrum(d64, DT)
    ip = -float(sysm.gs.eps.max())
    cut = sp.cutoff_energy(pulse.E0, pulse.omega0, ip) / pulse.omega0
    orders = np.arange(3, int(cut) + 8, 2)
    plateau = orders < 0.8 * cut
    I64 = sp.harmonic_intensities(w, S64, pulse.omega0, orders)
    J64 = sp.harmonic_intensities(wd, D64, pulse.omega0, orders)
    rep = {"steps": n, "cutoff_order": cut, "orders": orders.tolist(),
           "one_percent_in_log10": float(np.log10(1.01)), "wall_s_fp64": runs["fp64"][2],
           "batch_instances": n_batch}
    for name in ("fp32", "fp32_batched", "fp32_split_batched"):
        a32, d32, wall, variant = runs[name]
        _, S32 = sp.acceleration_spectrum(a32, DT)
        _, D32 = sp.dipole_spectrum(d32, DT)
        err = np.log10(sp.harmonic_intensities(w, S32, pulse.omega0, orders) / I64)
        errd = np.log10(sp.harmonic_intensities(wd, D32, pulse.omega0, orders) / J64)
        rep[name] = {"acc_rel_l2": float(np.linalg.norm(a32 - a64) / np.linalg.norm(a64)),
                     "dipole_rel_l2": float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64)),
                     "log10_intensity_error": err.tolist(),
                     "dipole_form_log10_error": errd.tolist(),
                     "max_abs_log10_err_plateau": float(np.abs(err[plateau]).max()),
                     "max_abs_log10_err_cutoff": float(np.abs(err[~plateau]).max()),
                     "kinetic_variant": variant, "wall_s": wall}
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--scan", type=int, default=64, metavar="N",
                    help="serve the batched He/Ne scan of N instances, FP64 then FP32 "
                         "(64 is the scan configuration; 0 = skip)")
    ap.add_argument("--scan-cycle-scale", type=float, default=1.0 / 16,
                    help="shorten every scan pulse by this factor (1 = the full scan, "
                         "minutes of GPU time)")
    ap.add_argument("--batched", type=int, default=64,
                    help="Ar instances for the batched FP32 leg (0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--no-batched-fp64", action="store_true",
                    help="skip the FP64 leg at the batched leg's instance count")
    ap.add_argument("--no-batched-split", action="store_true",
                    help="skip the compensated split-FP32 kinetic leg at that instance count")
    ap.add_argument("--fp64-window", type=int, default=1000,
                    help="steps of the FP64 window the batched ratios divide by (at least "
                         "this many, whatever --steps is)")
    ap.add_argument("--serving", type=int, default=24,
                    help="requests of the serving leg (0 = skip)")
    ap.add_argument("--serving-rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--serving-slots", type=int, default=8)
    ap.add_argument("--serving-saturated", type=int, default=128,
                    help="in-situ argon requests of the saturated serving run (0 = skip)")
    ap.add_argument("--serving-saturated-slots", type=int, default=64,
                    help="slots of the saturated run's lane")
    ap.add_argument("--serving-chunk", type=int, default=256,
                    help="steps per lane between seatings (a multiple of 64 keeps graph replay whole)")
    ap.add_argument("--no-precision-report", action="store_true",
                    help="skip the full-length FP64 / FP32 / batched-FP32 runs and their "
                         "harmonic-intensity errors")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on MIB MiB (profiler byte-counter "
                         "calibration) and print its bandwidth")
    args = ap.parse_args()
    if args.copy_calibration:
        import ctypes
        from ggbtddft_amd import _lib
        ms = ctypes.c_double()
        nb = args.copy_calibration << 20
        _lib.check(_lib.load().ggb_selftest_copy(0, nb, ctypes.byref(ms)), None, "ggb_selftest_copy")
        print(json.dumps({"copy_bytes": nb, "ms": ms.value,
                          "GBps_read_plus_write": 2 * nb / (ms.value * 1e-3) / 1e9}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs single-device contexts "
              "(sharding across GPUs is not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from ggbtddft_amd.engine import Propagator
    from ggbtddft_amd.pulse import Pulse
    from ggbtddft_amd.system import AtomSystem

    gs_stats, gs_dev = ground_state_timing()
    t0 = time.perf_counter()
    sysm = AtomSystem(ATOM, N, R_MAX, LMAX, DT, ground_state=gs_dev)
    log("precompute %.1f s after the ground state (sum nkeep %d)"
        % (time.perf_counter() - t0, int(sysm.nkeep.sum())))
    pulse = Pulse(INTENSITY, WAVELENGTH, CYCLES)
    nfull = pulse.n_steps(DT)
    need = args.warmup + args.steps + args.fp64_window + 200
    if need > nfull:
        raise SystemExit("warmup + steps (%d) exceeds the pulse (%d steps)" % (need, nfull))
    E = pulse.midpoints(DT)

    prop = Propagator(sysm)
    prop.set_field(E)
    tuned = prop.autotune(reps=10)
    log("autotuner (fp64 plain): %s %s" % tuned)
    prop.run(0, args.warmup)
    log("warm-up done (%d steps)" % args.warmup)
    prop.set_timing(1)
    t0 = time.perf_counter()
    obs = prop.run(args.warmup, args.steps)
    el = time.perf_counter() - t0
    log("timed %d steps in %.2f s" % (args.steps, el))
    tim = prop.timing()
    psi = prop.download_orbitals()
    s_next = args.warmup + args.steps
    # the denominator of the batched ratios: a window of at least fp64_window steps
    if args.steps >= args.fp64_window:
        ref_steps, ref_el = args.steps, el
    else:
        prop.set_timing(0)
        t0 = time.perf_counter()
        prop.run(s_next, args.fp64_window)
        ref_steps, ref_el = args.fp64_window, time.perf_counter() - t0
        s_next += args.fp64_window
    fp64_ref = sysm.n_orb * ref_steps / ref_el
    # per-phase breakdown in a separate pass (events between every phase)
    prop.set_timing(2)
    nb = 200
    prop.run(s_next, nb)
    phases = {k: v[0] / max(v[1], 1) for k, v in prop.timing().items()}
    prop.close()
    assert np.all(np.isfinite(obs["acc"])) and np.all(np.isfinite(psi))

    norb = sysm.n_orb
    value = norb * args.steps / el
    ms_step = 1e3 * el / args.steps
    # kinetic pass with the Q^T table: reads sum_l nkeep_l N doubles of Q,
    # the state (n_l C N complex) and writes Y (sum_l nkeep_l C complex)
    C = norb
    nk = int(sysm.nkeep.sum())
    bytes_qt = nk * N * 8 + sysm.n_l * C * N * 16 + nk * C * 16
    ms_qt, n_qt = tim["kinetic_qt"]
    ms_q, n_q = tim["kinetic_q"]
    avg_qt = ms_qt / max(n_qt, 1)
    avg_q = ms_q / max(n_q, 1)
    achieved = bytes_qt / (avg_qt * 1e-3) / 1e9
    traffic = measured_traffic()
    line = {
        "metric": "orbital-timesteps/s, Ar N=512 Lmax=64: batched FP32 vs FP64; p99 time-to-spectrum",
        "value": value,
        "unit": "orbital-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device LDA ground state from hydrogenic start, analytic sin^2 pulse)",
        "config": {"workload": "argon all-electron FP64 plain, one instance",
                   "atom": ATOM, "n_r": N, "r_max": R_MAX, "lmax": LMAX, "l_hartree": 2 * LMAX,
                   "n_theta": sysm.n_theta, "orbitals": norb, "dt": DT,
                   "pulse": "800 nm sin^2 8 cycles 2e14 W/cm^2", "full_run_steps": nfull,
                   "precision": "fp64", "parallelism": "single GPU"},
        "phase_ms_per_step": phases,
        "phase_rooflines": phase_rooflines(sysm, 1, 0, {k: 1e3 * v for k, v in phases.items()}, False),
        "fp64_ratio_denominator": {"value": fp64_ref, "steps": ref_steps,
                                   "ms_per_step": 1e3 * ref_el / ref_steps},
        "kinetic_variant": tuned[0], "autotuner_ms_per_kinetic_step": tuned[1],
        "roofline": {"kernel": "k_kin_fwd (kinetic propagator, Q^T pass)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["hbm_bytes_per_launch"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "bytes_per_launch": bytes_qt, "avg_launch_ms": avg_qt,
                     "avg_launch_ms_q_pass": avg_q},
    }
    if args.batched > 0:
        line["batched_fp32"] = batched_run(sysm, pulse, args, fp64_ref, precision=1)
        if not args.no_batched_fp64:
            b64 = batched_run(sysm, pulse, args, fp64_ref, precision=0)
            line["batched_fp64"] = b64
            line["batched_fp32"]["ratio_vs_fp64_batched"] = line["batched_fp32"]["value"] / b64["value"]
        if not args.no_batched_split:
            line["batched_fp32_split"] = batched_run(sysm, pulse, args, fp64_ref, precision=2)
    if args.serving > 0:
        line["serving"] = serving_run(sysm, args)
        sat = line["serving"].get("saturated_argon")
        if sat:
            sat["ratio_vs_fp64_plain"] = sat["orbital_timesteps_per_s"] / fp64_ref
    if args.scan > 0:
        line["scan"] = scan_run(args.scan, args.scan_cycle_scale)
    line["spectrum"] = spectrum_timing(nfull)
    line["ground_state"] = gs_stats
    if not args.no_precision_report:
        line["precision_report"] = precision_report(sysm, pulse)
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline((ATOM, N, R_MAX, LMAX, DT), E,
                                                nsteps=args.cpu_steps)
        except Exception as exc:  # the GPU number stands on its own
            line["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
