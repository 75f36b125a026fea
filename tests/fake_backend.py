"""Oracle-backed stand-in for ggbtddft_amd.engine.Propagator (test
infrastructure only): same methods the serving loop uses, each instance
stepped by the FP64 oracle on the CPU."""
import numpy as np

from oracle import propagate as OP


class OraclePropagator:
    def __init__(self, system, n_inst=1, device=0):
        self.sys = system
        self.n_inst = n_inst
        self.O = OP.System(system.atom.symbol, system.N, system.r_max, system.lmax, system.dt)
        self.psi = [self.O.psi0.copy() for _ in range(n_inst)]
        self.dV = [self.O.potential(self.O.theta_state(p)) for p in self.psi]
        self.E = np.zeros((n_inst, 1))

    def set_field(self, E):
        E = np.asarray(E, dtype=np.float64)
        self.E = np.tile(E, (self.n_inst, 1)) if E.ndim == 1 else E.copy()

    def field_ring(self, capacity):
        self.E = np.zeros((self.n_inst, capacity))

    def set_field_slice(self, inst, start, E):
        cap = self.E.shape[1]
        for j, v in enumerate(np.asarray(E)):
            self.E[inst, (start + j) % cap] = v

    def upload_instance(self, inst, psi=None):
        self.psi[inst] = (self.O.psi0 if psi is None else psi).copy()
        self.dV[inst] = self.O.potential(self.O.theta_state(self.psi[inst]))

    def run(self, step0, nsteps, channels=False):
        no = self.O.norb
        out = {k: np.zeros((nsteps, self.n_inst, no)) for k in ("dip", "acc", "norm")}
        if channels:
            out["chan"] = np.zeros((nsteps, self.n_inst, no, self.O.nl), dtype=np.complex128)
        cap = self.E.shape[1]
        for s in range(nsteps):
            for i in range(self.n_inst):
                self.O.step(self.psi[i], self.dV[i], self.E[i, (step0 + s) % cap])
                Psi = self.O.theta_state(self.psi[i])
                d, a, n = self.O.observables(Psi)
                out["dip"][s, i], out["acc"][s, i], out["norm"][s, i] = d, a, n
                if channels:
                    out["chan"][s, i] = self.O.channel_dipoles(self.psi[i])
                self.dV[i] = self.O.potential(Psi)
        return out

    def close(self):
        pass
