"""Reference-style run: an init file in, the workingData and output trees out.

Counterpart of the reference's ``init.py`` (precompute and the workingData
writers, init.py:85-380) and of ``main.cu`` around the propagation (the scan
loop, main.cu:460-481): the init keys (``config.RunConfig``), the ground state
(on the device with the GPU propagator), one propagation per scan -- the
scans of an in-situ delay scan run as instances of one batched context -- and
the total dipole acceleration with its spectrum, written in the reference's
layouts (``io``).

Time convention: observable sample k is taken after step k, which ends at
t_k of the reference's centred time grid (init.py:320); step k's field is
sampled at its midpoint t_k - dt/2.

    python -m ggbtddft_amd.run --init path/to/init --out output [--working workingData]
"""
import argparse
import functools
import math

import numpy as np

from . import io
from . import spectrum as spec
from .config import RunConfig
from .system import AtomSystem


def run(cfg, out_root, working_root=None, index=0, precision=0, device=0, factory=None,
        max_batch=64, sim_date=None):
    """Propagate every scan of `cfg` and write the output tree under
    out_root/DDDDDDDD/IIIII (and workingData under working_root when given).
    An in-situ run makes N_scan perturbed scans and then one with the
    unperturbed pulse (main.cu:460-467, h_initialiseField end = 1).
    `factory(system, n_inst=..., device=...)` makes the propagator (default:
    the GPU one).  Returns (output directory, per-scan dicts)."""
    gpu = factory is None
    system = AtomSystem(cfg.atom, cfg.N, cfg.r_max, cfg.lmax, cfg.dt, zeta=cfg.zeta,
                        interacting=cfg.interacting, ground_state="device" if gpu else None,
                        gs_device=device, propagator=cfg.propagator)
    if gpu:
        from .engine import Propagator
        factory = functools.partial(Propagator, precision=precision)
    t = cfg.time_grid()
    delays = (list(cfg.delays()) + [None]) if cfg.in_situ else [None]
    e_mid = [cfg.supergaussian_field(t - 0.5 * cfg.dt, d) for d in delays]
    e_end = [cfg.supergaussian_field(t, d) for d in delays]
    scans = []
    for b0 in range(0, len(delays), max_batch):
        nb = min(max_batch, len(delays) - b0)
        prop = factory(system, n_inst=nb, device=device)
        try:
            prop.set_field(np.stack(e_mid[b0:b0 + nb]))
            obs = prop.run(0, cfg.n_t, channels=True)
        finally:
            prop.close()
        for i in range(nb):
            s = b0 + i
            a = system.total_acceleration(obs["acc"][:, i], obs["norm"][:, i], e_end[s])
            if gpu:
                w, S, _ = spec.device_spectrum(a, cfg.dt, device=device)
            else:
                w, S = spec.acceleration_spectrum(a, cfg.dt)
            scans.append(dict(E=e_end[s], dip=obs["dip"][:, i], acc=obs["acc"][:, i],
                              norm=obs["norm"][:, i], chan=obs["chan"][:, i], acc_total=a,
                              omega=w, spectrum=S, delay=delays[s]))
    if working_root:
        io.write_working_data(working_root, system, t)
    macros = {"E0": cfg.E0, "omega0": cfg.omega0, "tau0": cfg.tau0, "CEP0": cfg.cep,
              "R_MASK": cfg.r_mask, "ETA_MASK": float(cfg.keys.get("ETA_MASK", 1e-3)),
              "PERTURBATION_AMPLITUDE": 0.01 if cfg.in_situ else 0.0,
              "COS_THETA": math.cos(30e-3)}
    base = io.write_output(out_root, index, system, t, scans, sim_date=sim_date,
                           in_situ=cfg.in_situ, macros=macros)
    return base, scans


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--init", required=True, help="reference-style init file")
    ap.add_argument("--out", required=True, help="output root (the run goes to OUT/DDDDDDDD/IIIII)")
    ap.add_argument("--working", default=None, help="also write the workingData tree here")
    ap.add_argument("--index", type=int, default=0, help="run index IIIII")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 1),
                    help="0 FP64, 1 FP32 mixed form")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    cfg = RunConfig.from_file(args.init)
    base, scans = run(cfg, args.out, args.working, args.index, args.precision, args.device)
    print("%s: %d scan(s) of %d steps" % (base, len(scans), cfg.n_t))


if __name__ == "__main__":
    main()
