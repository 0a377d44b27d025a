"""The 64-case ensemble workload through the batched launch, checked against
the oracle.

64 seeded heat and acoustic cases (sides drawn from {16, 24, 32, 48, 64, 96,
128}, FP32, the cases the benchmark times): a few steps in every tile form
equal the oracle's heat_run / wave_run of each case bit for bit, and a
500-step run keeps each heat case's total (zero-flux walls) and bounds each
acoustic case's energy (symplectic steps, reflective walls)."""
import numpy as np
import pytest

from cfdarcov2_amd import workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (tile_rows, chunk, group_bytes, form): 2.5D march tiles, single-step row
# tiles, two-step row tiles in one group and in cache-sized groups
SHAPES = [(8, 16, 0, 0), (8, 16, 0, 1), (12, 16, 0, 2), (10, 16, 192 << 20, 2)]


def _snapshot(ens, nsteps):
    torch.cuda.synchronize()
    return [[t.cpu().numpy() for t in ens.state(i, nsteps)] for i in range(len(ens.cases))]


def test_ensemble_workload_equals_oracle(hip_ctx, oracle_lib):
    from cfdarcov2_amd.ensemble import Ensemble
    specs, _ = workloads.ensemble_config()
    assert len(specs) == 64 and max(s.n for s in specs) == 128
    steps = 4
    snaps = {}
    ens = Ensemble(hip_ctx, specs)
    try:
        for shape in SHAPES:
            print(f"\n[case] ensemble workload shape={shape} steps={steps}", flush=True)
            ens.reset()
            ens.configure(*shape)
            ens.run(steps, use_graph=True)
            snaps[shape] = _snapshot(ens, steps)
    finally:
        ens.close()
    for i, sp in enumerate(specs):
        om = oracle_lib.OracleMesh(sp.n, sp.n, sp.n, sp.spacing, sp.spacing, sp.spacing, dtype=np.float32)
        f0 = np.asarray(sp.field0, dtype=np.float32)
        if sp.kind == "heat":
            ref = [om.heat_run(f0, sp.coef, sp.dt, steps)]
        else:
            z = np.zeros_like(f0)
            ref = om.wave_run(f0, z, z, z, sp.coef, sp.dt, steps)
        del om
        for shape, snap in snaps.items():
            for f, (g, r) in enumerate(zip(snap[i], ref)):
                assert np.array_equal(g, r), (i, sp.kind, sp.n, shape, f, float(np.abs(g - r).max()))


def test_ensemble_workload_500_steps_properties(hip_ctx):
    from cfdarcov2_amd.ensemble import Ensemble
    specs, steps = workloads.ensemble_config()
    assert steps == 500
    ens = Ensemble(hip_ctx, specs)
    try:
        print(f"\n[case] ensemble workload {steps} steps, two-step row tiles", flush=True)
        ens.configure(12, 16, 0, 2)
        ens.run(steps, use_graph=True)
        final = _snapshot(ens, steps)
    finally:
        ens.close()
    for i, sp in enumerate(specs):
        f0 = np.asarray(sp.field0, dtype=np.float32).astype(np.float64)
        got = [a.astype(np.float64) for a in final[i]]
        assert all(np.all(np.isfinite(a)) for a in got), i
        if sp.kind == "heat":
            # zero-flux walls: sum_i V T_i is conserved up to FP32 rounding
            assert abs(got[0].sum() - f0.sum()) <= 1e-4 * f0.sum(), (i, got[0].sum(), f0.sum())
            assert got[0].max() < f0.max()          # diffusion lowers the peak
        else:
            c = float(np.sqrt(-sp.coef))
            z = np.zeros_like(f0)
            E0 = workloads.wave_energy(f0, z, z, z, c)
            E = workloads.wave_energy(*got, c)
            assert 0.8 * E0 < E < 1.1 * E0, (i, E / E0)
