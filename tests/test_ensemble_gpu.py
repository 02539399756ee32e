"""Ensemble of uneven glacier domains stepped together (glide_amd.ensemble;
no counterpart in the reference, which runs one domain per process).

* Three valley glaciers of 48, 64 and 96 cells stepped as an ensemble over
  concurrent streams give bitwise the results of stepping each alone (the
  mixed form with fixed iterations, graph-replayed).
* The FP64 form's ensemble step matches the FP64 oracle per domain: the
  forward solve of 3 reference-form V-cycles (u, v, H) and the adjoint's
  basal-drag gradient, to 1e-10 relative.
"""
import numpy as np
import pytest
import torch

from tests.util import rel_err

pytestmark = pytest.mark.gpu

SIZES = (48, 64, 96)
DT = 1.0


def N(t):
    return t.detach().cpu().numpy()


def _results(model):
    g = model.mg[0]
    return [N(x) for x in (g.state.u.data, g.state.v.data, g.state.H.data, g.adjoint.lambda_u.data,
                           g.adjoint.lambda_v.data, g.adjoint.lambda_H.data, g.sliding.beta.grad)]


def test_ensemble_step_equals_single_domain_steps():
    from glide_amd.ensemble import Ensemble, valley_glacier
    print(f"ensemble of {SIZES} (mixed, 3 iterations) against single-domain steps", flush=True)
    probs = [valley_glacier(n, seed=40 + k) for k, n in enumerate(SIZES)]
    ens = Ensemble(probs, precision="mixed", iterations=3, streams=3)
    for q in range(2):
        ens.step(q * DT, DT)
    torch.cuda.synchronize()
    together = [_results(m) for m in ens.models]
    for k, p in enumerate(probs):
        one = Ensemble([p], precision="mixed", iterations=3, streams=1)
        for q in range(2):
            one.step(q * DT, DT)
        torch.cuda.synchronize()
        for a, b in zip(_results(one.models[0]), together[k]):
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b), SIZES[k]


def test_fp64_ensemble_matches_the_oracle(orc):
    from glide_amd.ensemble import Ensemble, levels_for, valley_glacier
    print(f"ensemble of {SIZES} (FP64, 3 V-cycles) against the oracle", flush=True)
    probs = [valley_glacier(n, seed=50 + k) for k, n in enumerate(SIZES)]
    ens = Ensemble(probs, precision="fp64", iterations=3, streams=3)
    ens.step(0.0, DT)
    torch.cuda.synchronize()
    cfg = orc.OracleMG.fas_config(vcycles=3, rtol=0.0, atol=0.0)
    for p, m in zip(probs, ens.models):
        om = orc.OracleMG(p.n, p.n, p.dx, levels_for(p.n))
        for k in ("bed", "B", "beta", "smb"):
            om.set(k, getattr(p, k))
        om.set("H", p.H)
        om.set("H_prev", p.H)
        om.fascd_solve(DT, cfg)
        g = m.mg[0]
        for k in ("u", "v", "H"):
            assert rel_err(N(getattr(g.state, k).data), om.get(k)) < 1e-10, (p.n, k)
        om.put("af_u", -om.get("u"))
        om.put("af_v", -om.get("v"))
        om.put("af_H", np.zeros((p.n, p.n)))
        om.fas_adjoint_solve(DT, cfg)
        om.gradients(DT)
        assert rel_err(N(g.sliding.beta.grad), om.get("g_beta")) < 1e-10, p.n
