"""Write the golden trajectory fixtures of the FP64 oracle.

    python tests/golden/make_golden.py

Each fixture holds the inputs (shape, parameters, seed, steps, cadence) and
the oracle's outputs after the run: psi, psibar (reference layout), the
controller state and the written observables.  The reference publishes no
vectors for this path, so these pin regressions of the oracle and give the
GPU tests a fixed target.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cl_var_dt_amd import CLV_CONJ_DRIFT_ACTION, CLV_CONJ_DRIFT_REFERENCE, Shape, default_params  # noqa: E402
from oracle import pyoracle  # noqa: E402

CASES = {
    "traj_1d_16x8": dict(shape=(16, 1, 1, 8), steps=60, eval_every=5,
                         params=dict(mu=1.0, G=0.1, c1=-0.05, Q=0.1, P=0.02, eps=1 / 8, DT=1e-3,
                                     seed=5, conj_drift_mode=CLV_CONJ_DRIFT_REFERENCE)),
    "traj_3d_8x4x2x4": dict(shape=(8, 4, 2, 4), steps=40, eval_every=4,
                            params=dict(mu=0.7, G=0.2, c1=0.05, Q=0.0, P=0.1, eps=0.25, DT=2e-3,
                                        seed=17, conj_drift_mode=CLV_CONJ_DRIFT_ACTION)),
}


def run(case):
    c = CASES[case]
    p = default_params(**c["params"])
    o = pyoracle.Oracle(Shape(*c["shape"]), p, nthreads=1)
    o.init_fields(1.0, 0.1)
    o.propagate(c["steps"], c["eval_every"])
    psi, psib = o.get_fields()
    st = o.stats()
    obs, F = o.observables()
    return dict(psi=psi, psibar=psib,
                stats=np.array([st.dt, st.max_drift, st.mean_drift, st.t_langevin, st.steps]),
                obs=np.array([obs.count, obs.n_tot, *obs.n_comp, obs.energy, obs.density, obs.Sz]),
                Fx=F[0], Fy=F[1], Fz=F[2])


if __name__ == "__main__":
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run(name))
        print("wrote", name)
