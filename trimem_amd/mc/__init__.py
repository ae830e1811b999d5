"""Monte Carlo drivers over :mod:`trimem_amd.core` (the counterpart of
``trimem.mc``: src/trimem/mc/hmc.py, mesh.py, evaluators.py)."""
