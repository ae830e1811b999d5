"""Energy, gradient and per-step callback of a mesh as functions of the
vertex positions, for samplers and minimisers that work on an array ``x``
(trimem.mc's EnergyEvaluators: src/trimem/mc/evaluators.py).

Each call first writes ``x`` into the mesh, then asks the
:class:`trimem_amd.core.EnergyManager`.  The callback runs the periodic
work of a simulation -- info, trajectory output, checkpoints, neighbour-list
refresh -- on the total step count, then advances the continuation of the
reference properties.  Option names and defaults are trimem's.
"""
import numpy as np

from .hmc import _settings


class EnergyEvaluators:
    """``fun(x)``, ``grad(x)`` and ``callback(x, steps)`` of one mesh.

    Args:
        mesh (Mesh): its vertices are set to ``x`` before every call.
        estore (EnergyManager): energy and gradient of the mesh.
        output: writer with ``write_points_cells(points, cells)``.
        options (dict): ``info_step`` (100), ``output_step`` (1000),
            ``cpt_step`` (0), ``refresh_step`` (10), ``flatten`` (False:
            gradients keep the (N,3) shape), ``num_steps`` (None),
            ``write_cpt`` (``write_cpt(mesh, estore, steps)``, a no-op by
            default).  A period of 0 (or below) switches its task off.
    """

    DEFAULTS = {
        "info_step": 100,
        "output_step": 1000,
        "cpt_step": 0,
        "refresh_step": 10,
        "flatten": False,
        "num_steps": None,
        "write_cpt": lambda m, e, s: None,
    }

    def __init__(self, mesh, estore, output, options):
        opts = _settings(self.DEFAULTS, options, "\n------------------------", "Energy Evaluators Setup:")
        self.mesh = mesh
        self.estore = estore
        self.output = output
        self.flatten = bool(opts["flatten"])
        write_cpt = opts["write_cpt"]
        # (period, task) in the order the tasks run within one callback
        self.tasks = [
            (max(opts["info_step"], 0), self._print_info),
            (max(opts["output_step"], 0), lambda steps: self.output.write_points_cells(self.mesh.x, self.mesh.f)),
            (max(opts["cpt_step"], 0), lambda steps: write_cpt(self.mesh, self.estore, steps)),
            (max(opts["refresh_step"], 0), lambda steps: self.estore.update_repulsion(self.mesh.trimesh)),
        ]

    def _load(self, x):
        self.mesh.x = np.reshape(x, self.mesh.x.shape)

    def fun(self, x):
        """Energy of the mesh with vertices ``x``."""
        self._load(x)
        return self.estore.energy(self.mesh.trimesh)

    def grad(self, x):
        """Gradient of the energy at ``x``, raveled with ``flatten``."""
        self._load(x)
        g = self.estore.gradient(self.mesh.trimesh)
        return g.ravel() if self.flatten else g

    def _print_info(self, steps):
        print("\n-- Energy-Evaluation-Step ", sum(steps.values()))
        self.estore.print_info(self.mesh.trimesh)

    def callback(self, x, steps):
        """Run every task whose period divides the total step count, then
        advance the continuation (update_reference_properties)."""
        self._load(x)
        n = sum(steps.values())
        for period, task in self.tasks:
            if period and n % period == 0:
                task(steps)
        self.estore.update_reference_properties()
