"""Writes tests/golden/si_rta_subset.npz: every 20th mode of the reference's
Si_rta.db (read with the non-executing decoder of openbte_amd.rta) plus the
database's stored bulk conductivity tensor and the full-database mode sums
used as known answers.  Run from the repository root:

    python tests/golden/make_si_subset.py /root/reference/openbte/materials
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from openbte_amd.rta import load_rta  # noqa: E402

path = sys.argv[1]
d = load_rta('Si_rta', path=path)
C, tau, v = d.heat_capacity, d.scattering_time, d.group_velocity
kappa_modes = np.einsum('k,ki,kj,k->ij', C, v, v, tau)
sel = np.arange(0, len(C), 20)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'si_rta_subset.npz')
np.savez_compressed(out, heat_capacity=C[sel], scattering_time=tau[sel], group_velocity=v[sel],
                    frequency=d.frequency[sel], kappa_db=d.thermal_conductivity, kappa_modes_full=kappa_modes,
                    n_modes_full=len(C), stride=20)
print(out, len(sel), np.diag(d.thermal_conductivity), np.diag(kappa_modes))
