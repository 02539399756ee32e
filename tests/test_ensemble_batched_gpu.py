"""Batched ensemble step (glide_ensemble_step; no counterpart in the
reference, which runs one domain per process): every stage is one launch
over a table of the domains taking part in it, the V-cycles run level by
level and the whole step replays as one HIP graph.

* Valley glaciers of 48, 40, 30 and 10 cells (4, 3, 2 and 1 levels: every
  descent, ascent and coarsest-level branch) stepped by the batched ensemble
  give bitwise the state, multipliers and gradients of each domain's own
  mixed solves (the per-domain ensemble path), over two steps of different
  lengths (the second re-records the tables).
* One batched step is a few hundred launches, independent of the domain count
  (the per-domain path issues roughly a thousand per domain).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = (48, 40, 30, 10)
DT = 1.0


def N(t):
    return t.detach().cpu().numpy()


def _results(model):
    g = model.mg[0]
    return {
        "u": N(g.state.u.data), "v": N(g.state.v.data), "H": N(g.state.H.data), "H_prev": N(g.state.H_prev.data),
        "mask": N(g.state.mask.data), "lambda_u": N(g.adjoint.lambda_u.data), "lambda_v": N(g.adjoint.lambda_v.data),
        "lambda_H": N(g.adjoint.lambda_H.data), "g_beta": N(g.sliding.beta.grad), "g_bed": N(g.geometry.bed.grad),
        "g_H_prev": N(g.state.H_prev.grad), "g_smb": N(g.forcing.smb.grad),
    }


def _run(batched):
    from glide_amd.ensemble import Ensemble, valley_glacier
    probs = [valley_glacier(n, seed=70 + k) for k, n in enumerate(SIZES)]
    ens = Ensemble(probs, precision="mixed", iterations=3, streams=2, batched=batched)
    for q, dt in enumerate((DT, 2.0 * DT)):
        print(f"  {'batched' if batched else 'per-domain'} step {q} (dt {dt})", flush=True)
        ens.step(q * DT, dt)
    torch.cuda.synchronize()
    return ens, [_results(m) for m in ens.models]


def test_batched_ensemble_equals_per_domain_solves():
    from glide_amd.ensemble import levels_for
    print(f"batched ensemble of {SIZES} (levels {[levels_for(n) for n in SIZES]}) against per-domain solves",
          flush=True)
    _, ref = _run(False)
    ens, bat = _run(True)
    for k, n in enumerate(SIZES):
        for name, a in ref[k].items():
            b = bat[k][name]
            assert np.all(np.isfinite(b)), (n, name)
            assert np.array_equal(a, b), (n, name, float(np.max(np.abs(a - b))))
    assert ens.launches_per_step > 0
    ens.close()


def test_batched_launch_count_independent_of_domains():
    from glide_amd.ensemble import Ensemble, valley_glacier
    counts = []
    for sizes in ((48, 40), (48, 40, 48, 40, 48, 40)):
        probs = [valley_glacier(n, seed=90 + k) for k, n in enumerate(sizes)]
        ens = Ensemble(probs, precision="mixed", iterations=3, batched=True)
        ens.step(0.0, DT)
        torch.cuda.synchronize()
        counts.append(ens.launches_per_step)
        for m in ens.models:
            assert np.all(np.isfinite(N(m.mg[0].sliding.beta.grad)))
        ens.close()
    print(f"launches per batched step: {counts}", flush=True)
    assert counts[0] == counts[1]
    assert counts[0] < 3000


def test_tiles_leave_results_unchanged_and_autotune_picks_candidates(tmp_path):
    from glide_amd.ensemble import Ensemble, autotune_tiles, valley_glacier
    probs = [valley_glacier(n, seed=70 + k) for k, n in enumerate(SIZES)]
    a = Ensemble(probs, precision="mixed", iterations=3, batched=True)
    b = Ensemble(probs, precision="mixed", iterations=3, batched=True)
    assert a.levels == 4
    b.set_tiles([1, 2, 8, 1, 2])
    print("  default tiles against [1, 2, 8, 1, 2]", flush=True)
    a.step(0.0, DT)
    b.step(0.0, DT)
    torch.cuda.synchronize()
    for ma, mb in zip(a.models, b.models):
        ra, rb = _results(ma), _results(mb)
        for name in ra:
            assert np.array_equal(ra[name], rb[name]), name
    a.close()
    b.close()
    cache = str(tmp_path / "tiles.json")
    print("  autotune", flush=True)
    tiles, timings = autotune_tiles(probs, iterations=3, reps=1, candidates=(1, 4), cache=cache)
    assert len(tiles) == 5 and all(t in (1, 4) for t in tiles)
    assert all(np.isfinite(ms) and ms > 0 for ms in timings.values())
    again, _ = autotune_tiles(probs, iterations=3, reps=1, candidates=(1, 4), cache=cache)
    assert again == tiles
