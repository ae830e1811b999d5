"""Hamiltonian Monte Carlo and edge-flip Monte Carlo for triangulated
membranes (the counterpart of src/trimem/mc/hmc.py).

:class:`HMC`, :class:`MeshHMC`, :class:`MeshFlips` and :class:`MeshMonteCarlo`
take the reference's arguments and options and consume numpy's global
random stream in the reference's order (per HMC step: ``normal`` of the
state's shape, then one ``uniform``; per combined step: one ``choice(2)``),
so a seeded run draws the same numbers and, given the same energies and
gradients, reproduces the reference's chain.

:class:`DeviceMeshHMC` runs the same HMC step with positions, momenta and
forces resident in HBM (:class:`trimem_amd.core.DeviceHMC`): one trajectory
is a sequence of GPU kernels with no host round trip per force evaluation.
With ``rng="host"`` it takes its momenta and uniform from numpy's global
stream exactly like :class:`MeshHMC`, and then yields the same positions;
with ``rng="device"`` both are drawn on the GPU from ``(seed, step)``.
"""
from collections import Counter

import numpy as np

from .. import core as _core


def _vv_integration(x0, p0, force, m, dt, N):
    """Velocity Verlet in momentum form (hmc.py:14-26).

    One force evaluation at x0, then N steps of drift, force, kick.  The
    arithmetic is grouped as in the reference (x + (p dt + 0.5 a dt**2) / m,
    p + 0.5 (a + an) dt) so that results agree bit for bit.
    """
    dt2 = dt**2
    x, p = x0, p0
    a = force(x)
    for _ in range(N):
        x = x + (p * dt + 0.5 * a * dt2) / m
        an = force(x)
        p = p + 0.5 * (a + an) * dt
        a = an
    return x, p


def get_step_counters():
    """Step counter shared by moves and flips (hmc.py:28-35)."""
    return Counter(move=0, flip=0)


# The reference's constructors default to one Counter created at import time,
# which every sampler built without an explicit counter shares (hmc.py:90,
# 216, 266, 346).  The same object plays that role here.
_SHARED_COUNTER = get_step_counters()

_hmc_default_options = {
    "mass": 1.0,
    "time_step": 1.0e-4,
    "num_integration_steps": 100,
    "initial_temperature": 1.0,
    "minimal_temperature": 1.0e-6,
    "cooling_factor": 0.0,
    "cooling_start_step": 0,
    "info_step": 100,
}

_mc_flip_default_options = {
    "flip_type": "parallel",
    "flip_ratio": 0.1,
    "info_step": 100,
}


def _print_options(title, options):
    print(title)
    width = max(len(str(k)) for k in options.keys())
    for k, v in options.items():
        print(f"  {k: <{width}}: {v}")


class HMC:
    """Hamiltonian Monte Carlo with optional exponential cooling
    (hmc.py:48-191).

    Args:
        x (ndarray[float]): initial state.
        nlog_prob (callable): negative log density, ``nlog_prob(x) -> float``.
        grad_nlog_prob (callable): its gradient, same shape as ``x``.

    Keyword Args:
        callback (callable): stored like the reference's; note that, as in
            the reference (hmc.py:98), :meth:`run` does not invoke it.
        counter (collections.Counter): step counter (``move``, ``flip``).
        options (dict): ``mass``, ``time_step``, ``num_integration_steps``,
            ``initial_temperature``, ``minimal_temperature``,
            ``cooling_factor``, ``cooling_start_step``, ``info_step``.
    """

    def __init__(self, x, nlog_prob, grad_nlog_prob, callback=None, counter=None, options={}):
        self.nlog_prob = nlog_prob
        self.grad_nlog_prob = grad_nlog_prob
        self.callback = callback
        self.cb = lambda x, s: None

        options = {**_hmc_default_options, **options}
        self.m = options["mass"]
        self.dt = options["time_step"]
        self.L = options["num_integration_steps"]
        self.Tinit = options["initial_temperature"]
        self.Tmin = options["minimal_temperature"]
        self.fT = options["cooling_factor"]
        self.cN = options["cooling_start_step"]
        self.istep = options["info_step"]

        print("\n---------------------------------------")
        _print_options("Hamiltonian Monte Carlo Initialization:", options)

        self.counter = _SHARED_COUNTER if counter is None else counter
        self.i = 0
        self.acc = 0
        self.T = self.Tinit
        self.x = x

    def _hamiltonian(self, x, p):
        """E(x) + p.p / (2 m) (hmc.py:129-131)."""
        return self.nlog_prob(x) + 0.5 * p.ravel().dot(p.ravel()) / self.m

    def _step(self):
        """One Metropolis-corrected trajectory (hmc.py:133-158)."""
        p = np.random.normal(size=self.x.shape) * np.sqrt(self.m * self.T)
        xn, pn = _vv_integration(self.x, p, lambda y: -self.grad_nlog_prob(y), self.m, self.dt, self.L)
        dh = (self._hamiltonian(xn, pn) - self._hamiltonian(self.x, p)) / self.T
        alpha = 1.0 if dh <= 0 else np.exp(-dh)
        if np.random.uniform() <= alpha:
            self.x = xn
            self.acc += 1
        self.i += 1

    def _update_temperature(self):
        """T = clamp(Tinit exp(-fT (i - cN)), Tmin, Tinit) with i the total
        step count (hmc.py:174-177)."""
        i = sum(self.counter.values())
        Tn = np.exp(-self.fT * (i - self.cN)) * self.Tinit
        self.T = max(min(Tn, self.Tinit), self.Tmin)

    def _acceptance_count(self):
        return self.acc

    def _reset_acceptance(self):
        self.acc = 0

    def info(self):
        """Acceptance rate and temperature every ``info_step`` total steps
        (hmc.py:160-169)."""
        i_total = sum(self.counter.values())
        if self.istep and i_total % self.istep == 0:
            ar = self._acceptance_count() / self.i if self.i != 0 else 0.0
            print("\n-- HMC-Step ", self.counter["move"])
            print("----- acc-rate:   ", ar)
            print("----- temperature:", self.T)
            self._reset_acceptance()
            self.i = 0

    def step(self):
        """Update the temperature, make one step, count it (hmc.py:171-183)."""
        self._update_temperature()
        self._step()
        self.counter["move"] += 1

    def run(self, N):
        """N steps with info output (hmc.py:185-190)."""
        for _ in range(N):
            self.step()
            self.info()
            self.cb(self.x, self.counter)


class MeshHMC(HMC):
    """HMC whose state is a :class:`trimem_amd.mc.mesh.Mesh` (hmc.py:193-233):
    after every step the accepted state is written to the mesh."""

    def __init__(self, mesh, nlog_prob, grad_nlog_prob, callback=None, counter=None, options={}):
        super().__init__(mesh.x, nlog_prob, grad_nlog_prob, callback, counter, options)
        self.mesh = mesh

    def step(self):
        super().step()
        self.mesh.x = self.x


class DeviceMeshHMC(HMC):
    """MeshHMC with the trajectory and the Metropolis test on the GPU.

    The positions live in the mesh's device copy; ``mesh.x`` (and this
    object's ``x``) copy them back on access.  Energy and forces are those of
    ``estore`` (a :class:`trimem_amd.core.EnergyManager` of ``mesh``), in its
    precision mode.

    Keyword Args:
        rng (str): ``"host"`` draws momenta and the uniform from numpy's
            global stream in :class:`MeshHMC`'s order (bit-identical chains
            for the same seed); ``"device"`` draws them on the GPU.
        seed (int): key of the device draws (``rng="device"``).
        model (str): ``"energy"`` (default) or ``"harmonic"`` (E = x.x / 2,
            the reference sampler test's energy, with ``estore=None``).
        (other arguments as :class:`HMC`)
    """

    def __init__(self, mesh, estore, callback=None, counter=None, options={}, rng="host", seed=0,
                 model="energy"):
        if rng not in ("host", "device"):
            raise ValueError(f"Wrong rng: {rng}")
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
        return None if getattr(self, "mesh", None) is None else self.mesh.x

    @x.setter
    def x(self, values):
        if values is not None:
            self.mesh.x = values

    def _step(self):
        kw = dict(mass=self.m, time_step=self.dt, num_integration_steps=self.L, temperature=self.T,
                  time_step_sq=self.dt**2)
        if self.rng == "host":
            shape = (self.mesh.trimesh.n_vertices(), 3)
            p = np.random.normal(size=shape) * np.sqrt(self.m * self.T)
            u = np.random.uniform()
            self.dev.step(momenta=p, uniform=u, **kw)
        else:
            self.dev.step(seed=self.seed, step=self.n_draws, **kw)
        self.n_draws += 1
        self.i += 1

    def result(self):
        """Outcome of the last step (dh, energies, acceptance) and running
        counts; synchronises with the GPU."""
        return self.dev.result()

    def _acceptance_count(self):
        return self.dev.result()["n_accepted"] - self._acc_base

    def _reset_acceptance(self):
        self._acc_base = self.dev.result()["n_accepted"]

    @property
    def acc(self):
        return 0 if getattr(self, "dev", None) is None else self._acceptance_count()

    @acc.setter
    def acc(self, value):
        if getattr(self, "dev", None) is not None and value == 0:
            self._reset_acceptance()


class MeshFlips:
    """Edge-flip sweeps as Monte Carlo steps (hmc.py:242-322).

    Options: ``flip_type`` ("parallel", "serial" or "none"), ``flip_ratio``
    (fraction of edges proposed per sweep) and ``info_step``.
    """

    def __init__(self, mesh, estore, counter=None, options={}):
        self.mesh = mesh
        self.estore = estore
        options = {**_mc_flip_default_options, **options}
        self.istep = options["info_step"]
        self.fr = options["flip_ratio"]
        self.ft = options["flip_type"]

        print("\n--------------------------------")
        _print_options("Flips Monte Carlo Initialization:", options)

        if self.ft == "none" or self.fr == 0.0:
            self._flips = lambda: 0
        elif self.ft == "serial":
            self._flips = lambda: _core.flip(self.mesh.trimesh, self.estore, self.fr)
        elif self.ft == "parallel":
            self._flips = lambda: _core.pflip(self.mesh.trimesh, self.estore, self.fr)
        else:
            raise ValueError("Wrong flip-type: {}".format(self.ft))

        self.i = 0
        self.acc = 0
        self.counter = _SHARED_COUNTER if counter is None else counter

    def info(self):
        """Flip acceptance per edge every ``info_step`` total steps
        (hmc.py:300-310)."""
        i_total = sum(self.counter.values())
        if self.istep and i_total % self.istep == 0:
            n_edges = self.mesh.trimesh.n_edges()
            ar = self.acc / (self.i * n_edges) if self.i != 0 else 0.0
            print("\n-- MCFlips-Step ", self.counter["flip"])
            print("----- flip-accept: ", ar)
            print("----- flip-rate:   ", self.fr)
            self.acc = 0
            self.i = 0

    def step(self):
        self.acc += self._flips()
        self.i += 1
        self.counter["flip"] += 1

    def run(self, N):
        for _ in range(N):
            self.step()
            self.info()


class MeshMonteCarlo:
    """Alternates HMC moves and flip sweeps with probability 1/2 each
    (hmc.py:325-372); all three objects share one step counter."""

    def __init__(self, hmc, flips, counter=None, callback=None):
        self.hmc = hmc
        self.flips = flips
        self.cb = (lambda x, s: None) if callback is None else callback
        self.counter = _SHARED_COUNTER if counter is None else counter
        self.hmc.counter = self.counter
        self.flips.counter = self.counter

    def step(self):
        if np.random.choice(2) == 0:
            self.hmc.step()
        else:
            self.flips.step()

    def run(self, N):
        for _ in range(N):
            self.step()
            self.hmc.info()
            self.flips.info()
            self.cb(self.flips.mesh.x, self.counter)
