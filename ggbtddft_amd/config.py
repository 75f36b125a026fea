"""Run configuration: the reference's `init` keys, read without executing them.

The reference `exec`s each line of src/init (init.py:75-77) and fills defaults
(init.py:85-294).  Here each `KEY = expression` line is parsed with ``ast`` and
evaluated by a small whitelist (numbers, + - * / **, parentheses, np.pi,
np.sqrt, bare atom/method names as strings); anything else is an error.

Key meanings and conversions follow init.py: dt in attoseconds
(x ATTOSECOND_AU, :222-227), I0 in W/cm^2 (E0 = sqrt(I0/3.509e16), :255-256),
tau0 in femtoseconds (x FEMTOSECOND_AU, :234-237), R_MASK =
max(ABC_R_SCALE E0/omega0^2, 16) and R_MAX = 1.5 R_MASK (:279-284).
"""
import ast
import math

import numpy as np

from .atoms import ATTOSECOND_AU, FEMTOSECOND_AU, INTENSITY_AU

DEFAULTS = {
    "TD_SIMULATION": 0, "IN_SITU": 0, "FROM_SCRATCH": 1, "N_scan": 64, "USE_MULLER": 0,
    "N_r": 64, "zeta": 0.20, "N_l": 32, "N_t": 2048, "dt": 10.0, "I0": 1e10,
    "CEP0": math.pi / 2.0, "USE_HARTREE": 1, "N_har": 2, "EXP_METHOD": "EIG",
    "ABC_R_SCALE": 1.0,
}

_FUNCS = {"sqrt": math.sqrt, "exp": math.exp, "log": math.log, "power": math.pow}
_CONST = {"pi": math.pi, "e": math.e}


def _eval(node):
    if isinstance(node, ast.Expression):
        return _eval(node.body)
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float, str)):
        return node.value
    if isinstance(node, ast.Name):
        if node.id in ("True", "False"):
            return node.id == "True"
        return node.id                       # bare names (atom, method) -> strings
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _eval(node.operand)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp):
        a, b = _eval(node.left), _eval(node.right)
        ops = {ast.Add: lambda x, y: x + y, ast.Sub: lambda x, y: x - y,
               ast.Mult: lambda x, y: x * y, ast.Div: lambda x, y: x / y,
               ast.Pow: lambda x, y: x ** y}
        for t, f in ops.items():
            if isinstance(node.op, t):
                return f(a, b)
    if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) \
            and node.value.id in ("np", "numpy", "math") and node.attr in _CONST:
        return _CONST[node.attr]
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute) \
            and isinstance(node.func.value, ast.Name) \
            and node.func.value.id in ("np", "numpy", "math") and node.func.attr in _FUNCS \
            and not node.keywords:
        return _FUNCS[node.func.attr](*[_eval(a) for a in node.args])
    raise ValueError("unsupported expression in init file: %s" % ast.dump(node))


def read_init(text):
    """Parse init-file text into a dict of raw values (no defaults applied)."""
    out = {}
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line or "=" not in line:
            continue
        key, expr = line.split("=", 1)
        key = key.strip()
        if not key.isidentifier():
            raise ValueError("bad key %r" % key)
        out[key] = _eval(ast.parse(expr.strip(), mode="eval"))
    return out


class RunConfig:
    """Reference init keys mapped to this propagator's parameters (atomic units)."""

    def __init__(self, keys):
        k = dict(DEFAULTS)
        k.update(keys)
        if "atom" not in k:
            raise ValueError("init: 'atom' is required (init.py:208-213)")
        self.keys = k
        self.atom = str(k["atom"])
        self.N = int(k["N_r"])
        self.zeta = float(k["zeta"])
        self.lmax = int(k["N_l"]) - 1
        self.n_t = int(k["N_t"])
        self.dt = float(k["dt"]) * ATTOSECOND_AU
        self.I0 = float(k["I0"])
        self.E0 = math.sqrt(self.I0 / INTENSITY_AU)
        if "omega0" in k:
            self.omega0 = float(k["omega0"])
        else:  # init.py:263-269
            co, ip = 70.0 / 27.2114, 12.8 / 27.2114
            self.omega0 = 0.890224690738243 * math.sqrt(self.I0 / INTENSITY_AU) / math.sqrt(co - ip)
        self.tau0 = float(k["tau0"]) * FEMTOSECOND_AU if "tau0" in k else 0.25 * self.n_t * self.dt
        self.cep = float(k["CEP0"])
        self.in_situ = bool(k["IN_SITU"])
        self.n_scan = int(k["N_scan"]) if self.in_situ else 1
        self.interacting = bool(k["USE_HARTREE"])
        r_mask = max(float(k["ABC_R_SCALE"]) * self.E0 / self.omega0 ** 2, 16.0)
        self.r_mask = float(k.get("R_MASK", r_mask))
        self.r_max = 1.5 * self.r_mask
        # EXP_METHOD (init.py:55-58, 189; internal.py:252-318): CN = 0, CN2 = 1,
        # CN3 = 2, EIG = 3 (the default), anything else expm; EIG and expm are
        # both exp(-i H dt)
        em = k["EXP_METHOD"]
        em = {0: "CN", 1: "CN2", 2: "CN3", 3: "EIG"}.get(em, em) if isinstance(em, (int, float)) else str(em)
        self.propagator = {"CN": "cn", "CN2": "cn2", "CN3": "cn3"}.get(str(em).upper(), "exp")

    @classmethod
    def from_file(cls, path):
        with open(path) as f:
            return cls(read_init(f.read()))

    def time_grid(self):
        """t_k = (k - N_t/2) dt (init.py:320)."""
        return (np.arange(self.n_t) - 0.5 * self.n_t) * self.dt

    def supergaussian_field(self, t, delay=None, amplitude=1e-2, cos_theta=math.cos(30e-3)):
        """The reference's pulse E0 exp(-(2t/tau0)^10) cos(omega0 t + CEP)
        (h_initialiseField.cuh:20-35), with the in-situ perturbation when
        delay is given."""
        t = np.asarray(t, dtype=np.float64)
        E = self.E0 * np.exp(-(2.0 * t / self.tau0) ** 10) * np.cos(self.omega0 * t + self.cep)
        if delay is not None:
            tp = t * cos_theta
            E = E + amplitude * self.E0 * np.exp(-(2.0 * tp / self.tau0) ** 10) * \
                np.cos(self.omega0 * (t + delay) + self.cep)
        return E

    def delays(self):
        """TD = linspace(-T0, T0, N_scan) * 4 (init.py:271-272)."""
        T0 = 2.0 * math.pi / self.omega0
        return np.linspace(-T0, T0, self.n_scan) * 4.0
