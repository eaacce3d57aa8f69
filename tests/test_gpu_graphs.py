"""GMRES-IR sweep cycles captured as HIP graphs (ccab_set_graphs): each
cycle's launches (m Arnoldi steps over the cycle's systems, back
substitution, update) are captured once per shape and replayed.  Replays
equal each other bit for bit across repeated sweeps with unrelated eager work
and synchronisations between them, and agree with the eager sweep to the
solve tolerance (the eager path shrinks its launches as systems finish, which
changes the GEMV's column split and so the rounding)."""
import math

import numpy as np
import pytest
import torch

from cc_abstemious_amd import capi, incident_wave, mesh, solve
from cc_abstemious_amd.preprocess import IncidentWave, PreData, Solver

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[case] {msg}", flush=True)


def _pd(level, nfreq):
    fs = list(np.linspace(0.5, 8.0, nfreq) * 343.0 / (2 * math.pi))
    return PreData(mesh.icosphere(level), fs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR,
                   capi.BC_NORMAL_VELOCITY, (0.0, 0.0),
                   IncidentWave(capi.WAVE_PLANE, (0.3, -0.4, 0.866), (1.0, 0.0)),
                   Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-10, restart=30,
                          ir_tolerance=1e-3, ir_max_steps=60, ref_mgs=False,
                          stop_rule=capi.STOP_RELATIVE_RESIDUAL, matrix_prec=capi.C32,
                          residual_matrix=capi.C32))


@pytest.mark.parametrize("level,nfreq", [(3, 16), (4, 6)])
def test_graph_cycles_replay_and_match_eager(level, nfreq):
    pd = _pd(level, nfreq)
    n = pd.get_num_eqn()
    _note(f"GMRES-IR graph cycles n={n} nfreq={nfreq} (Gram-Schmidt "
          f"{'multi-workgroup' if n >= 4096 else 'one workgroup'} per system)")
    ctx = capi.Context()
    phi_e, _, st_e = solve.sweep(pd, ctx=ctx)
    ctx.set_graphs(True)
    outs = []
    for rep in range(3):
        phi, _, st = solve.sweep(pd, ctx=ctx)    # rep 0 captures, later reps replay
        assert all(s["converged"] for s in st), rep
        outs.append((phi.clone(), [s["iterations"] for s in st]))
        # unrelated eager work and a synchronisation between the replays
        pd1 = _pd(2, 1)
        pd1.solver = Solver(kind=capi.SOLVER_DIRECT)
        solve.solve_for_surface(pd1, incident_wave.get_incident_surface(pd1, ctx=ctx), ctx=ctx)
        torch.cuda.synchronize()
    ctx.set_graphs(False)
    for rep, (phi, its) in enumerate(outs):
        assert torch.equal(phi, outs[0][0]), rep
        assert its == outs[0][1], rep
        e = float(torch.linalg.vector_norm(phi - phi_e) / torch.linalg.vector_norm(phi_e))
        print(f"replay {rep}: vs eager {e:.2e}, iterations {sum(its)} vs "
              f"{sum(s['iterations'] for s in st_e)}")
        assert e < 1e-9
    ctx.close()
