"""Step loop of SolverDEM::run (src/CabanaDEM_Solver.hpp:79-97) over the CPU
oracle, one step at a time so tests can record time series.  Test helper."""
from __future__ import annotations

import numpy as np

import oracle_py as O


def solver_run(st, w, prm_kw, run_kw, record=None, gravity=(0., 0., 0.), mode=0, rebuild_every=1):
    prm = O.params(dt=prm_kw["dt"], cor_pp=prm_kw["cor_pp"], cor_pw=prm_kw["cor_pw"],
                   fr_pp=prm_kw["fr_pp"], fr_pw=prm_kw["fr_pw"], gravity=gravity, mode=mode)
    dt = prm_kw["dt"]
    counts = ids = None
    for step in range(run_kw["nsteps"]):
        O.stage(st, 1, dt)                                   # :81
        O.stage(st, 2, dt)                                   # :83
        if counts is None or step % rebuild_every == 0:      # :87-88
            counts, ids = O.neighbors_cell(st.x, run_kw["rc"], run_kw["lo"], run_kw["hi"])
        ov = O.compute_forces(st, w, counts, ids, prm)       # :90-91
        assert ov == 0, "tracked-contact capacity exceeded"
        O.stage(st, 3, dt)                                   # :94
        if record is not None:
            record(step, st)
    return st
