"""Fixed-cycle FAS solves replayed from cached HIP graphs against the same
solves launched eagerly.

Two identical coupled models step side by side, one with graph replay and one
without, with eager torch work and synchronisations between the calls; every
replay must give bitwise the same state, adjoint and gradients as the eager
model (all kernels are deterministic).  With launch timing on, both models run
eagerly (events inside graphs cannot be timed) and must still agree."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _model(n, levels, seed):
    from glide_amd import IceDynamics
    from glide_amd.synthetic import dome_problem, install
    p = dome_problem(n, 1500.0, seed=seed)
    model = IceDynamics(ny=n, nx=n, dx=1500.0, n_levels=levels, device="cuda")
    install(model.mg, p)
    for s in (model.forward_solver, model.adjoint_solver):
        s.fas_options.set(maximum_vcycles=2, relative_tolerance=0.0, absolute_tolerance=0.0, report_norms=False)
        s.fas_options.set(finest_steps=6, pre_steps=3, post_steps=4, coarsest_steps=12)
    return model


def _state(model):
    g = model.mg[0]
    return [t.detach().clone() for t in (g.state.u.data, g.state.v.data, g.state.H.data, g.state.mask.data,
                                         g.adjoint.lambda_u.data, g.adjoint.lambda_v.data, g.adjoint.lambda_H.data,
                                         g.sliding.beta.grad, g.geometry.bed.grad)]


@pytest.mark.parametrize("timing", [False, True])
def test_graph_replays_match_eager(timing):
    from glide_amd import _lib
    print(f"graph replay timing={timing}", flush=True)
    eager, graphed = _model(32, 3, 3), _model(32, 3, 3)
    _lib.timing_enable(timing)
    try:
        t = 0.0
        for step in range(4):
            for model, use_graph in ((eager, False), (graphed, True)):
                _lib.graphs_enable(use_graph)
                model.forward(t, 10.0, update_geometry=True)
                g = model.mg[0]
                model.backward(t, 10.0, dJdu=g.state.u.data, dJdv=g.state.v.data, dJdH=None)
                torch.cuda.synchronize()
                # eager work of the program between replays
                g.state.H_prev.data.mul_(1.0)
                torch.cuda.synchronize()
            for a, b in zip(_state(eager), _state(graphed)):
                assert torch.equal(a, b), f"step {step}"
            t += 10.0
        if timing:
            c, ms, cells = _lib.timing_read("vanka_cell")
            assert c > 0 and ms > 0.0
    finally:
        _lib.graphs_enable(True)
        _lib.timing_enable(False)


def test_graph_cache_follows_configuration_changes():
    """A changed option (here the V-cycle count) is a different graph: results
    match an eager run with the same options."""
    from glide_amd import _lib
    print("graph cache config change", flush=True)
    eager, graphed = _model(32, 3, 4), _model(32, 3, 4)
    try:
        for vcycles in (2, 1, 2):
            for model, use_graph in ((eager, False), (graphed, True)):
                _lib.graphs_enable(use_graph)
                model.forward_solver.fas_options.set(maximum_vcycles=vcycles)
                model.forward(0.0, 10.0, update_geometry=False)
                torch.cuda.synchronize()
            for a, b in zip(_state(eager)[:4], _state(graphed)[:4]):
                assert torch.equal(a, b), f"vcycles {vcycles}"
    finally:
        _lib.graphs_enable(True)
