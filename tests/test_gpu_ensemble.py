"""GPU checks of the batched ensemble: every case's fields after a batched run
(plain launches, one group or several, and HIP-graph replays) equal the same
case advanced alone by its single-case fused kernel, bit for bit."""
import numpy as np
import pytest

from cfdarcov2_amd import workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _specs():
    # uneven sides, both kinds, sides that do not divide the tile shapes
    return workloads.ensemble_config(n_cases=10, steps=0, seed=3, spacing=1.0 / 64,
                                     sides=(5, 16, 23, 40, 67, 33))[0]


def _snapshot(ens, nsteps):
    torch.cuda.synchronize()
    return [[t.clone() for t in ens.state(i, nsteps)] for i in range(len(ens.cases))]


def _compare(got, ref):
    for i, (g, r) in enumerate(zip(got, ref)):
        for f, (a, b) in enumerate(zip(g, r)):
            assert torch.equal(a, b), (i, f, float((a - b).abs().max()))


@pytest.mark.parametrize("shape", [(4, 8, 0), (8, 5, 0), (16, 32, 0), (4, 16, 1 << 20)])
def test_batched_equals_sequential(hip_ctx, shape):
    from cfdarcov2_amd.ensemble import Ensemble
    print(f"\n[case] ensemble batched shape={shape}", flush=True)
    ens = Ensemble(hip_ctx, _specs())
    try:
        ens.run_sequential(7)
        ref = _snapshot(ens, 7)
        ens.reset()
        ens.configure(*shape)
        if shape[2]:
            assert len(ens.groups) > 1
        ens.run(7, use_graph=False)
        _compare(_snapshot(ens, 7), ref)
    finally:
        ens.close()


def test_graph_replays(hip_ctx):
    """Three replays of a captured 6-step graph, with eager work and
    synchronisations between them, each checked against the plain runs."""
    from cfdarcov2_amd.ensemble import Ensemble
    print("\n[case] ensemble graph replays", flush=True)
    ens = Ensemble(hip_ctx, _specs())
    try:
        refs = []
        for k in range(3):
            ens.run_sequential(6)
            refs.append(_snapshot(ens, 6))
        ens.reset()
        ens.configure(8, 8, 0)
        for k in range(3):
            ens.run(6, use_graph=True)
            scratch = torch.randn(1 << 20, device="cuda")
            scratch.mul_(2.0).sum().item()
            _compare(_snapshot(ens, 6), refs[k])
    finally:
        ens.close()


def test_errors(hip_ctx):
    from cfdarcov2_amd import capi
    from cfdarcov2_amd.ensemble import Ensemble
    ens = Ensemble(hip_ctx, _specs()[:2])
    try:
        ens.configure()
        with pytest.raises(capi.CfdError):
            ens.run(3, use_graph=True)        # odd graph runs are refused
        with pytest.raises(capi.CfdError):
            capi.check(ens.lib.cfd_batch_set_shape(ens.groups[0], 12, 8))
    finally:
        ens.close()
    # a non-uniform mesh cannot join a batch
    m = capi.HexMesh(hip_ctx, 4, 4, 4, 0.1, 0.1, 0.1, dtype="f32")
    assert m.info().uniform == 0
    t = [torch.zeros(64, device="cuda") for _ in range(2)]
    arr = (capi.BatchCase * 1)()
    arr[0].mesh = m.h
    arr[0].kind = 0
    arr[0].state_a[0] = t[0].data_ptr()
    arr[0].state_b[0] = t[1].data_ptr()
    h = capi.ctypes.c_void_p()
    assert ens.lib.cfd_batch_create(hip_ctx.h, arr, 1, capi.ctypes.byref(h)) == -4
    m.close()


def _row_specs():
    # z extents that suit the row form (multiples of 4 up to 256), uneven
    # sides that do not divide the tile shapes, both kinds
    return workloads.ensemble_config(n_cases=10, steps=0, seed=5, spacing=1.0 / 64,
                                     sides=(4, 16, 24, 36, 68, 40, 8))[0]


@pytest.mark.parametrize("shape", [(4, 8, 0), (8, 5, 0), (4, 1, 0), (8, 16, 1 << 20)])
def test_batched_row_form_equals_sequential(hip_ctx, shape):
    from cfdarcov2_amd.ensemble import Ensemble
    print(f"\n[case] ensemble batched row form shape={shape}", flush=True)
    ens = Ensemble(hip_ctx, _row_specs())
    try:
        ens.run_sequential(7)
        ref = _snapshot(ens, 7)
        ens.reset()
        ens.configure(*shape, form=1)
        ens.run(7, use_graph=False)
        _compare(_snapshot(ens, 7), ref)
        # graph replays
        ens.reset()
        ens.run_sequential(6)
        ref6 = _snapshot(ens, 6)
        for _ in range(2):
            ens.reset()
            ens.run(6, use_graph=True)
            torch.randn(1 << 16, device="cuda").sum().item()
            _compare(_snapshot(ens, 6), ref6)
    finally:
        ens.close()


def test_batched_row_form_refuses_unsuited_cases(hip_ctx):
    from cfdarcov2_amd import capi
    from cfdarcov2_amd.ensemble import Ensemble
    ens = Ensemble(hip_ctx, _specs())  # z extents 5, 23, ... are not multiples of 4
    try:
        with pytest.raises(capi.CfdError) as e:
            ens.configure(4, 8, 0, form=1)
        assert e.value.code == -4
    finally:
        ens.close()


@pytest.mark.parametrize("shape", [(8, 8, 0), (10, 5, 0), (12, 3, 1 << 20), (12, 16, 0)])
def test_batched_two_step_equals_sequential(hip_ctx, shape):
    """Row form with two steps per launch: 9 steps (two pairs of two-step
    launches and one single step) and graph replays of 10 steps equal the
    single-case runs bit for bit."""
    from cfdarcov2_amd.ensemble import Ensemble
    print(f"\n[case] ensemble batched two-step shape={shape}", flush=True)
    ens = Ensemble(hip_ctx, _row_specs())
    try:
        ens.run_sequential(9)
        ref = _snapshot(ens, 9)
        ens.reset()
        ens.configure(*shape, form=2)
        ens.run(9, use_graph=False)
        _compare(_snapshot(ens, 9), ref)
        ens.reset()
        ens.run_sequential(10)
        ref10 = _snapshot(ens, 10)
        for _ in range(2):
            ens.reset()
            ens.run(10, use_graph=True)
            torch.randn(1 << 16, device="cuda").sum().item()
            _compare(_snapshot(ens, 10), ref10)
    finally:
        ens.close()



def test_benchmark_configuration_500_steps_equals_single_case_runs(hip_ctx):
    """The benchmark's whole run: 500 steps of the 64-case configuration
    (heat even, acoustic odd, seeded sides 16 to 128, FP32) as one captured
    graph of the two-step batched form equal the 64 single-case runs of the
    same length bit for bit (the oracle checks of this configuration are in
    test_gpu_ensemble_config.py)."""
    from cfdarcov2_amd.ensemble import Ensemble
    specs, steps = workloads.ensemble_config()
    assert len(specs) == 64 and steps == 500
    print("\n[case] ensemble 64-case configuration, 500 steps batched vs single-case", flush=True)
    ens = Ensemble(hip_ctx, specs)
    try:
        ens.configure(12, 16, 0, form=2)
        ens.run(steps, use_graph=True)
        got = _snapshot(ens, steps)
        ens.reset()
        ens.run_sequential(steps)
        _compare(got, _snapshot(ens, steps))
    finally:
        ens.close()
