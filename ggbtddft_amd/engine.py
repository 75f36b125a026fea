"""Propagator: owns one libggbtd context (one HIP stream on one device).

Counterpart of the reference's driver around h_timePropagation
(main.cu:218-594): allocation, loaders, field, the step loop and the
observable read-back, with device memory resident across calls.
"""
import ctypes

import numpy as np

from . import _lib


_TUNED = {}  # (N, n_l, nk_max, columns, precision) -> (variant, timings) of the autotuner


def problem_for(system, n_inst=1, device=0, precision=0):
    """ggb_problem of a context holding n_inst instances of `system`."""
    p = _lib.Problem()
    p.abi_version = _lib.GGB_ABI_VERSION
    p.n_r, p.n_l, p.n_theta = system.N, system.n_l, system.n_theta
    p.n_har, p.n_orb, p.n_inst = system.n_har, system.n_orb, int(n_inst)
    p.n_m, p.nk_max = system.angular.n_m, system.nk_max
    p.interacting = 1 if system.interacting else 0
    p.precision, p.device, p.dt = int(precision), int(device), system.dt
    return p


def footprint(system, n_inst=1, field_steps=0, obs_steps=0, precision=0):
    """Device bytes of a context (ggb_footprint) plus its field and observable buffers."""
    lib = _lib.load()
    b = ctypes.c_uint64()
    p = problem_for(system, n_inst, 0, precision)
    _lib.check(lib.ggb_footprint(ctypes.byref(p), ctypes.byref(b)), None, "ggb_footprint")
    return int(b.value) + 8 * int(n_inst) * int(field_steps) + \
        24 * system.n_orb * int(n_inst) * int(obs_steps)


class Propagator:
    """Batched propagator for n_inst instances of one AtomSystem.

    Instances share the grid, the ground state and the kinetic tables; each
    has its own orbitals, field and potentials.
    """

    def __init__(self, system, n_inst=1, device=0, precision=0):
        self.lib = _lib.load()
        self.precision = int(precision)
        self.sys = system
        self.n_inst = int(n_inst)
        self.C = system.n_orb * self.n_inst
        p = problem_for(system, self.n_inst, device, self.precision)
        self._problem = p
        ctx = ctypes.c_void_p()
        _lib.check(self.lib.ggb_create(ctypes.byref(p), ctypes.byref(ctx)), None, "ggb_create")
        self.ctx = ctx
        self._arrays = system.static_arrays()
        a = self._arrays
        s = _lib.Static()
        for name in ("r", "wr", "v_gs", "mask", "Q", "lam", "B", "PL", "chi", "omega", "occ"):
            setattr(s, name, _lib.dptr(a[name]))
        s.nkeep = _lib.iptr(a["nkeep"])
        s.m = _lib.iptr(a["m"])
        s.rho_frozen = None
        if all(k in a for k in ("phi", "l0", "eps")):
            s.phi = _lib.dptr(a["phi"])
            s.l0 = _lib.iptr(a["l0"])
            s.eps = _lib.dptr(a["eps"])
        self._check(self.lib.ggb_upload_static(self.ctx, ctypes.byref(s)), "ggb_upload_static")
        self.n_field = 0
        self.upload_orbitals()

    def _check(self, rc, what):
        _lib.check(rc, self.ctx, what)

    # ---- state -----------------------------------------------------------
    def upload_orbitals(self, psi=None):
        """psi: [n_l, C, N] complex (default: the ground state in every instance)."""
        if psi is None:
            psi = np.concatenate([self.sys.psi0] * self.n_inst, axis=1)
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        assert psi.shape == (self.sys.n_l, self.C, self.sys.N)
        buf = psi.view(np.float64)
        self._check(self.lib.ggb_upload_orbitals(self.ctx, _lib.dptr(buf)), "ggb_upload_orbitals")

    def download_orbitals(self):
        out = np.empty((self.sys.n_l, self.C, self.sys.N), dtype=np.complex128)
        self._check(self.lib.ggb_download_orbitals(self.ctx, _lib.dptr(out.view(np.float64))),
                    "ggb_download_orbitals")
        return out

    def download_potential(self):
        out = np.empty((self.n_inst, self.sys.N, self.sys.n_theta))
        self._check(self.lib.ggb_download_potential(self.ctx, _lib.dptr(out)),
                    "ggb_download_potential")
        return out

    # ---- field and propagation ---------------------------------------------
    def set_field(self, E):
        """E: [n_inst, n_steps] midpoint field samples (or [n_steps] for all)."""
        E = np.asarray(E, dtype=np.float64)
        if E.ndim == 1:
            E = np.tile(E, (self.n_inst, 1))
        E = np.ascontiguousarray(E)
        assert E.shape[0] == self.n_inst
        self._check(self.lib.ggb_set_field(self.ctx, _lib.dptr(E), E.shape[1]), "ggb_set_field")
        self.n_field = E.shape[1]

    def field_ring(self, capacity):
        """Zeroed ring field buffer: step n reads sample n mod capacity."""
        self._check(self.lib.ggb_field_ring(self.ctx, int(capacity)), "ggb_field_ring")
        self.n_field = int(capacity)

    def set_field_slice(self, inst, start, E):
        E = np.ascontiguousarray(E, dtype=np.float64)
        self._check(self.lib.ggb_set_field_slice(self.ctx, int(inst), int(start), E.size,
                                                 _lib.dptr(E)), "ggb_set_field_slice")

    def upload_instance(self, inst, psi=None):
        """Seat orbitals [n_l, n_orb, N] (default: ground state) in instance slot
        `inst` and rebuild the potentials (other instances are unchanged)."""
        if psi is None:
            psi = self.sys.psi0
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        assert psi.shape == (self.sys.n_l, self.sys.n_orb, self.sys.N)
        c0 = int(inst) * self.sys.n_orb
        self._check(self.lib.ggb_upload_columns(self.ctx, c0, self.sys.n_orb,
                                                _lib.dptr(psi.view(np.float64))),
                    "ggb_upload_columns")
        self._check(self.lib.ggb_rebuild_potential(self.ctx), "ggb_rebuild_potential")

    def set_active(self, active):
        """Mark the instance slots in use (bool per instance); idle ones skip
        the angular and potential kernels."""
        m = np.ascontiguousarray(np.asarray(active, dtype=bool).astype(np.int32))
        assert m.shape == (self.n_inst,)
        self._check(self.lib.ggb_set_active(self.ctx, _lib.iptr(m)), "ggb_set_active")

    def run(self, step0, nsteps, observables=True, channels=False):
        """Propagate nsteps steps; returns dict(dip, acc, norm) of shape
        [nsteps, n_inst, n_orb] (None when observables=False) and, with
        channels, chan [nsteps, n_inst, n_orb, n_l] complex: the per-(orbital,
        l) channel dipoles of the reference's output files."""
        if channels != getattr(self, "_channels", False):
            self._check(self.lib.ggb_set_channels(self.ctx, 1 if channels else 0), "ggb_set_channels")
            self._channels = bool(channels)
        if not observables:
            self._check(self.lib.ggb_run(self.ctx, step0, nsteps, None, None, None), "ggb_run")
            return None
        dip = np.empty((nsteps, self.C))
        acc = np.empty((nsteps, self.C))
        nrm = np.empty((nsteps, self.C))
        self._check(self.lib.ggb_run(self.ctx, step0, nsteps, _lib.dptr(dip), _lib.dptr(acc),
                                     _lib.dptr(nrm)), "ggb_run")
        sh = (nsteps, self.n_inst, self.sys.n_orb)
        out = dict(dip=dip.reshape(sh), acc=acc.reshape(sh), norm=nrm.reshape(sh))
        if channels and nsteps > 0:
            ch = np.empty((nsteps, self.C, self.sys.n_l), dtype=np.complex128)
            self._check(self.lib.ggb_get_channels(self.ctx, nsteps, _lib.dptr(ch.view(np.float64))),
                        "ggb_get_channels")
            out["chan"] = ch.reshape(sh + (self.sys.n_l,))
        return out

    def set_kinetic_variant(self, variant):
        """Kinetic kernel variant: index or name in _lib.KIN_VARIANTS."""
        v = _lib.KIN_VARIANTS.index(variant) if isinstance(variant, str) else int(variant)
        self._check(self.lib.ggb_set_kinetic_variant(self.ctx, v), "ggb_set_kinetic_variant")

    def autotune(self, reps=10, cache=True):
        """Pick the fastest kinetic variant for this context's shape (timed on
        the device with the current instance mask, orbitals restored);
        remembered per shape in this process unless cache is False.
        Returns (name, {name: ms per kinetic step})."""
        key = (self.sys.N, self.sys.n_l, self.sys.nk_max, self.C, self.precision)
        if cache and key in _TUNED:
            name, times = _TUNED[key]
            self.set_kinetic_variant(name)
            return name, times
        chosen = ctypes.c_int32()
        ms = np.zeros(len(_lib.KIN_VARIANTS))
        self._check(self.lib.ggb_autotune(self.ctx, int(reps), ctypes.byref(chosen), _lib.dptr(ms)),
                    "ggb_autotune")
        times = {_lib.KIN_VARIANTS[v]: float(ms[v]) for v in range(1, len(ms)) if ms[v] > 0}
        name = _lib.KIN_VARIANTS[chosen.value]
        if cache:
            _TUNED[key] = (name, times)
        return name, times

    def set_graph(self, steps):
        """Replay the step loop from a hipGraph of `steps` steps (0 = off)."""
        self._check(self.lib.ggb_set_graph(self.ctx, int(steps)), "ggb_set_graph")

    def set_timing(self, mode):
        """0 off, 1 kinetic passes, 2 every phase (HIP events on the context stream)."""
        self._check(self.lib.ggb_set_timing(self.ctx, int(mode)), "ggb_set_timing")

    def timing(self):
        """{phase: (total ms, launches)} accumulated since set_timing."""
        ms = np.zeros(_lib.NPHASE)
        n = np.zeros(_lib.NPHASE, dtype=np.int64)
        self._check(self.lib.ggb_get_timing(self.ctx, _lib.dptr(ms),
                                            n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
                    "ggb_get_timing")
        return {name: (float(ms[k]), int(n[k])) for k, name in enumerate(_lib.PHASES)}

    def substep(self, which, step=0):
        self._check(self.lib.ggb_debug_substep(self.ctx, which, step), "ggb_debug_substep")

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.ggb_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Group:
    """Several propagators (e.g. one per grid of a batched scan) advanced
    together: whole runs of `steps` steps of every member come from one
    hipGraph launch (ggb_group_run), each member on its own stream."""

    def __init__(self, props, steps=64):
        self.lib = _lib.load()
        self.props = list(props)
        arr = (ctypes.c_void_p * len(self.props))(*[p.ctx.value for p in self.props])
        g = ctypes.c_void_p()
        rc = self.lib.ggb_group_create(arr, len(self.props), int(steps), ctypes.byref(g))
        if rc != 0:
            raise _lib.GGBError("ggb_group_create failed: %s"
                                % self.lib.ggb_group_last_error(None).decode())
        self.g = g

    def run(self, step0, nsteps):
        """Advance member k from step0[k] by nsteps; returns one dict(dip, acc,
        norm) per member, [nsteps, n_inst, n_orb] each."""
        n = len(self.props)
        s0 = np.ascontiguousarray(np.asarray(step0, dtype=np.int32))
        assert s0.shape == (n,)
        outs = [dict(dip=np.empty((nsteps, p.C)), acc=np.empty((nsteps, p.C)),
                     norm=np.empty((nsteps, p.C))) for p in self.props]
        ptrs = {k: (_lib._dp * n)(*[_lib.dptr(o[k]) for o in outs]) for k in ("dip", "acc", "norm")}
        rc = self.lib.ggb_group_run(self.g, _lib.iptr(s0), int(nsteps), ptrs["dip"], ptrs["acc"],
                                    ptrs["norm"])
        if rc != 0:
            raise _lib.GGBError("ggb_group_run failed (%s): %s"
                                % (_lib.ERRORS.get(rc, rc), self.lib.ggb_group_last_error(self.g).decode()))
        for p, o in zip(self.props, outs):
            sh = (nsteps, p.n_inst, p.sys.n_orb)
            for k in o:
                o[k] = o[k].reshape(sh)
        return outs

    def close(self):
        if getattr(self, "g", None):
            self.lib.ggb_group_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def selftest_mfma(device=0):
    """[2, 16, 16]: the FP64 and FP32 16x16x4 matrix-core products."""
    lib = _lib.load()
    out = np.empty((2, 16, 16))
    _lib.check(lib.ggb_selftest_mfma(device, _lib.dptr(out)), None, "ggb_selftest_mfma")
    return out
