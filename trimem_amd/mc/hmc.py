"""Samplers over :mod:`trimem_amd.core`: Hamiltonian moves of the vertices,
edge-flip sweeps of the triangulation, and their random alternation.

The classes carry trimem.mc's names, constructor arguments, option names and
defaults, and console output (src/trimem/mc/hmc.py), so trimem's drivers run
on them unchanged.  Given the same seed they draw numpy's global random
stream in the same order -- per Hamiltonian move one ``normal`` of the
state's shape and then one ``uniform``; per alternation step one
``choice(2)`` -- and the leapfrog groups its arithmetic like trimem's, so a
host chain is reproducible bit for bit.

:class:`DeviceMeshHMC` keeps the whole move on the GPU
(:class:`trimem_amd.core.DeviceHMC`): positions, momenta and forces stay in
HBM for the trajectory and the Metropolis test.
"""
from collections import Counter

import numpy as np

from .. import core as _core


def leapfrog(x, p, force, mass, dt, n_steps):
    """``n_steps`` velocity-Verlet steps of (x, p) under ``force``.

    One force evaluation at the start, then per step a drift with the
    current force and a kick with the mean of the old and new force.  The
    grouping ``x + (p dt + a dt^2 / 2) / m`` and ``p + (a + a') dt / 2`` is
    trimem's (hmc.py:14-26); it decides the last bits of every step, which
    the device integrator reproduces.
    """
    dt_sq = dt**2
    a = force(x)
    for _ in range(n_steps):
        x = x + (p * dt + 0.5 * a * dt_sq) / mass
        a_new = force(x)
        p = p + 0.5 * (a + a_new) * dt
        a = a_new
    return x, p


def get_step_counters():
    """A fresh counter of moves and flips (hmc.py:28-35)."""
    return Counter(move=0, flip=0)


# trimem's samplers default their ``counter`` argument to a single Counter
# made at import time (hmc.py:90, 216, 266, 346), so every sampler built
# without one counts into the same object; this module keeps that.
_DEFAULT_COUNTER = get_step_counters()


def _settings(defaults, given, rule, title):
    """The caller's options over the defaults, printed as a table under a
    banner (the console output of trimem's samplers)."""
    merged = dict(defaults)
    merged.update(given)
    print(rule)
    print(title)
    width = max(len(str(k)) for k in merged)
    for key, value in merged.items():
        print(f"  {key: <{width}}: {value}")
    return merged


class Cooling:
    """Temperature after n total Monte Carlo steps: T_init until the start
    step, then T_init exp(-factor (n - start)), never below T_min
    (hmc.py:174-177)."""

    def __init__(self, t_init, t_min, factor, start):
        self.t_init, self.t_min, self.factor, self.start = t_init, t_min, factor, start

    def __call__(self, n_total):
        t = np.exp(-self.factor * (n_total - self.start)) * self.t_init
        return max(min(t, self.t_init), self.t_min)


class HMC:
    """Metropolis-corrected Hamiltonian moves of a state ``x`` under the
    density exp(-nlog_prob(x) / T).

    Args:
        x (ndarray[float]): initial state.
        nlog_prob (callable): ``nlog_prob(x) -> float``.
        grad_nlog_prob (callable): its gradient, shaped like ``x``.

    Keyword Args:
        callback: kept as ``self.callback`` and not called by :meth:`run`,
            as in trimem (its HMC stores a lambda that returns the callback,
            hmc.py:98).
        counter (Counter): shared count of ``move`` and ``flip`` steps.
        options (dict): ``mass`` (1.0), ``time_step`` (1e-4),
            ``num_integration_steps`` (100), ``initial_temperature`` (1.0),
            ``minimal_temperature`` (1e-6), ``cooling_factor`` (0.0),
            ``cooling_start_step`` (0), ``info_step`` (100).

    ``x`` holds the chain's state and ``acc`` the moves accepted since the
    last info line.
    """

    DEFAULTS = {
        "mass": 1.0,
        "time_step": 1.0e-4,
        "num_integration_steps": 100,
        "initial_temperature": 1.0,
        "minimal_temperature": 1.0e-6,
        "cooling_factor": 0.0,
        "cooling_start_step": 0,
        "info_step": 100,
    }

    def __init__(self, x, nlog_prob, grad_nlog_prob, callback=None, counter=None, options={}):
        opts = _settings(self.DEFAULTS, options, "\n---------------------------------------",
                         "Hamiltonian Monte Carlo Initialization:")
        self.m = opts["mass"]
        self.dt = opts["time_step"]
        self.L = opts["num_integration_steps"]
        self.cooling = Cooling(opts["initial_temperature"], opts["minimal_temperature"],
                               opts["cooling_factor"], opts["cooling_start_step"])
        self.info_every = opts["info_step"]
        self.energy = nlog_prob
        self.gradient = grad_nlog_prob
        self.callback = callback
        self.counter = _DEFAULT_COUNTER if counter is None else counter
        self.T = self.cooling.t_init
        self.x = x
        self.acc = 0          # accepted moves since the last info line
        self.window = 0       # moves since the last info line

    def hamiltonian(self, x, p):
        """E(x) + |p|^2 / 2m (hmc.py:129-131)."""
        q = p.ravel()
        return self.energy(x) + 0.5 * q.dot(q) / self.m

    def _move(self):
        p = np.random.normal(size=self.x.shape) * np.sqrt(self.m * self.T)
        x_new, p_new = leapfrog(self.x, p, lambda y: -self.gradient(y), self.m, self.dt, self.L)
        dh = (self.hamiltonian(x_new, p_new) - self.hamiltonian(self.x, p)) / self.T
        threshold = 1.0 if dh <= 0 else np.exp(-dh)
        if np.random.uniform() <= threshold:
            self.x = x_new
            self.acc += 1
        self.window += 1

    def _accepted_in_window(self):
        return self.acc

    def _new_window(self):
        self.acc = 0
        self.window = 0

    def info(self):
        """Every ``info_step`` total steps: the move count, the acceptance
        rate since the last such line and the temperature."""
        if self.info_every and sum(self.counter.values()) % self.info_every == 0:
            rate = self._accepted_in_window() / self.window if self.window else 0.0
            print("\n-- HMC-Step ", self.counter["move"])
            print("----- acc-rate:   ", rate)
            print("----- temperature:", self.T)
            self._new_window()

    def step(self):
        """Set the temperature for the current total step count, make one
        move and count it."""
        self.T = self.cooling(sum(self.counter.values()))
        self._move()
        self.counter["move"] += 1

    def run(self, N):
        for _ in range(N):
            self.step()
            self.info()


class MeshHMC(HMC):
    """HMC of a :class:`~trimem_amd.mc.mesh.Mesh`'s vertices; the mesh holds
    the state after every step.  The initial state is the mesh's points
    view itself, as in trimem (hmc.py:221), so evaluators that write the
    trajectory into the mesh also move it."""

    def __init__(self, mesh, nlog_prob, grad_nlog_prob, callback=None, counter=None, options={}):
        super().__init__(mesh.x, nlog_prob, grad_nlog_prob, callback, counter, options)
        self.mesh = mesh

    def step(self):
        super().step()
        self.mesh.x = self.x


class DeviceMeshHMC(HMC):
    """MeshHMC with the trajectory and the Metropolis test on the GPU.

    The positions live in the mesh's device copy; ``mesh.x`` (and this
    object's ``x``) copy them back on access.  Energy and forces are those
    of ``estore`` (a :class:`trimem_amd.core.EnergyManager` of ``mesh``), in
    its precision mode.

    Keyword Args:
        rng (str): ``"host"`` takes the momenta and the uniform from numpy's
            global stream in :class:`MeshHMC`'s order (the same chain bit for
            bit); ``"device"`` draws them on the GPU from ``(seed, step)``.
        seed (int): key of the device draws.
        model (str): ``"energy"`` (default) or ``"harmonic"`` (E = x.x / 2,
            trimem's sampler test energy, with ``estore=None``).
        (other arguments as :class:`HMC`)
    """

    def __init__(self, mesh, estore, callback=None, counter=None, options={}, rng="host", seed=0,
                 model="energy"):
        if rng not in ("host", "device"):
            raise ValueError(f"Wrong rng: {rng}")
        self.mesh = None
        self.dev = None
        super().__init__(None, None, None, callback, counter, options)
        self.mesh = mesh
        self.estore = estore
        self.rng = rng
        self.seed = int(seed)
        self.n_draws = 0
        self.dev = _core.DeviceHMC(mesh.trimesh, estore, model)
        self._acc_base = 0

    @property
    def x(self):
        return None if self.mesh is None else self.mesh.x

    @x.setter
    def x(self, values):
        if values is not None:
            self.mesh.x = values

    @property
    def acc(self):
        return 0 if self.dev is None else self._accepted_in_window()

    @acc.setter
    def acc(self, value):
        if self.dev is not None and value == 0:
            self._acc_base = self.dev.result()["n_accepted"]

    def _move(self):
        kw = dict(mass=self.m, time_step=self.dt, num_integration_steps=self.L, temperature=self.T,
                  time_step_sq=self.dt**2)
        if self.rng == "host":
            shape = (self.mesh.trimesh.n_vertices(), 3)
            p = np.random.normal(size=shape) * np.sqrt(self.m * self.T)
            self.dev.step(momenta=p, uniform=np.random.uniform(), **kw)
        else:
            self.dev.step(seed=self.seed, step=self.n_draws, **kw)
        self.n_draws += 1
        self.window += 1

    def result(self):
        """Outcome of the last move (dh, energies, acceptance) and the running
        counts; synchronises with the GPU."""
        return self.dev.result()

    def _accepted_in_window(self):
        return self.dev.result()["n_accepted"] - self._acc_base

    def _new_window(self):
        self._acc_base = self.dev.result()["n_accepted"]
        self.window = 0


class MeshFlips:
    """Edge-flip sweeps of a mesh as Monte Carlo steps.

    Options: ``flip_type`` ("parallel" (default), "serial" or "none"),
    ``flip_ratio`` (0.1: the fraction of the edges proposed per sweep) and
    ``info_step`` (100).  A zero ratio or type "none" makes steps that
    flip nothing; any other unknown type raises ValueError.
    """

    DEFAULTS = {"flip_type": "parallel", "flip_ratio": 0.1, "info_step": 100}
    SWEEPS = {"serial": _core.flip, "parallel": _core.pflip}

    def __init__(self, mesh, estore, counter=None, options={}):
        opts = _settings(self.DEFAULTS, options, "\n--------------------------------",
                         "Flips Monte Carlo Initialization:")
        self.mesh = mesh
        self.estore = estore
        self.info_every = opts["info_step"]
        self.fr = opts["flip_ratio"]
        kind = opts["flip_type"]
        if kind == "none" or self.fr == 0.0:
            self._sweep = None
        elif kind in self.SWEEPS:
            self._sweep = self.SWEEPS[kind]
        else:
            raise ValueError("Wrong flip-type: {}".format(kind))
        self.counter = _DEFAULT_COUNTER if counter is None else counter
        self.acc = 0          # accepted flips since the last info line
        self.window = 0       # sweeps since the last info line

    def info(self):
        """Every ``info_step`` total steps: the flip count, accepted flips per
        edge and sweep since the last such line, and the ratio."""
        if self.info_every and sum(self.counter.values()) % self.info_every == 0:
            rate = self.acc / (self.window * self.mesh.trimesh.n_edges()) if self.window else 0.0
            print("\n-- MCFlips-Step ", self.counter["flip"])
            print("----- flip-accept: ", rate)
            print("----- flip-rate:   ", self.fr)
            self.acc = 0
            self.window = 0

    def step(self):
        if self._sweep is not None:
            self.acc += self._sweep(self.mesh.trimesh, self.estore, self.fr)
        self.window += 1
        self.counter["flip"] += 1

    def run(self, N):
        for _ in range(N):
            self.step()
            self.info()


class MeshMonteCarlo:
    """Each step is a Hamiltonian move or a flip sweep with probability 1/2
    (one ``np.random.choice(2)`` per step); the two samplers count into one
    counter, and ``callback(x, counter)`` follows every step."""

    def __init__(self, hmc, flips, counter=None, callback=None):
        self.hmc = hmc
        self.flips = flips
        self.counter = _DEFAULT_COUNTER if counter is None else counter
        self.hmc.counter = self.counter
        self.flips.counter = self.counter
        self.callback = callback

    def step(self):
        (self.hmc if np.random.choice(2) == 0 else self.flips).step()

    def run(self, N):
        for _ in range(N):
            self.step()
            self.hmc.info()
            self.flips.info()
            if self.callback is not None:
                self.callback(self.flips.mesh.x, self.counter)
