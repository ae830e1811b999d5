"""Energy, gradient and per-step callback as functions of the vertex
positions (the counterpart of src/trimem/mc/evaluators.py).

Generic samplers and minimisers act on an array ``x``; these evaluators
write ``x`` into the mesh and call the :class:`trimem_amd.core.EnergyManager`
(evaluators.py:98-150), and the callback injects trimem's periodic work:
info output, trajectory output, checkpoints, the neighbour-list refresh and
the continuation of the reference properties (evaluators.py:152-183).
"""
import time
from datetime import datetime, timedelta

import numpy as np

_eval_default_options = {
    "info_step": 100,
    "output_step": 1000,
    "cpt_step": 0,
    "refresh_step": 10,
    "flatten": False,
    "num_steps": None,
    "write_cpt": lambda m, e, s: None,
}


class EnergyEvaluators:
    """``fun(x)``, ``grad(x)`` and ``callback(x, steps)`` over one mesh.

    Args:
        mesh (Mesh): state; its vertices are set to ``x`` before every call.
        estore (EnergyManager): energy and gradient backend.
        output: object with ``write_points_cells(points, cells)``.
        options (dict): ``info_step``, ``output_step``, ``cpt_step``,
            ``refresh_step``, ``flatten``, ``num_steps``, ``write_cpt``
            (defaults as evaluators.py:14-22).
    """

    def __init__(self, mesh, estore, output, options):
        self.mesh = mesh
        self.estore = estore
        self.output = output
        options = {**_eval_default_options, **options}

        print("\n------------------------")
        print("Energy Evaluators Setup:")
        width = max(len(str(k)) for k in options.keys())
        for k, v in options.items():
            print(f"  {k: <{width}}: {v}")

        self.info_step = max(options["info_step"], 0)
        self.out_step = max(options["output_step"], 0)
        self.cpt_step = max(options["cpt_step"], 0)
        self.refresh_step = max(options["refresh_step"], 0)
        self.write_cpt = options["write_cpt"]
        self._ravel = np.ravel if options["flatten"] else (lambda y: y)

    def _set(self, x):
        self.mesh.x = x.reshape(self.mesh.x.shape)

    def fun(self, x):
        """Energy of the mesh with vertices ``x``."""
        self._set(x)
        return self.estore.energy(self.mesh.trimesh)

    def grad(self, x):
        """Gradient of the energy at ``x`` (raveled if ``flatten``)."""
        self._set(x)
        return self._ravel(self.estore.gradient(self.mesh.trimesh))

    def callback(self, x, steps):
        """Per-step hook: info every ``info_step``, output every
        ``output_step``, checkpoint every ``cpt_step``, neighbour-list
        refresh every ``refresh_step`` total steps, then the continuation
        update of the reference properties (evaluators.py:152-183)."""
        self._set(x)
        i = sum(steps.values())
        if self.info_step and i % self.info_step == 0:
            print("\n-- Energy-Evaluation-Step ", i)
            self.estore.print_info(self.mesh.trimesh)
        if self.out_step and i % self.out_step == 0:
            self.output.write_points_cells(self.mesh.x, self.mesh.f)
        if self.cpt_step and i % self.cpt_step == 0:
            self.write_cpt(self.mesh, self.estore, steps)
        if self.refresh_step and i % self.refresh_step == 0:
            self.estore.update_repulsion(self.mesh.trimesh)
        self.estore.update_reference_properties()


class TimingEnergyEvaluators(EnergyEvaluators):
    """EnergyEvaluators that also print the measured seconds per step and an
    estimated end time every ``info_step`` steps (evaluators.py:186-214)."""

    def __init__(self, mesh, estore, output, options):
        super().__init__(mesh, estore, output, options)
        self.timestamps = []
        self.n = options["num_steps"] // self.info_step
        self.start = datetime.now()

    def callback(self, x, steps):
        super().callback(x, steps)
        i = sum(steps.values())
        if self.info_step and i % self.info_step == 0:
            self.timestamps.append(time.time())
            if len(self.timestamps) == 2:
                tspan = self.timestamps[1] - self.timestamps[0]
                speed = tspan / self.info_step
                finish = self.start + timedelta(seconds=tspan) * self.n
                print("\n-- Performance measurements")
                print(f"----- estimated speed: {speed:.3e} s/step")
                print(f"----- estimated end:   {finish}")
                self.timestamps.pop(0)
