"""Coupling from the past on the GPU (rt_domino_cftp) as an exact sampler: domino placement
frequencies of independent samples against the exact placement probabilities
(randomtilings_amd.placement, Kasteleyn), uniform and weighted heat-bath rules."""
import numpy as np
import pytest

import randomtilings_amd as rt
from randomtilings_amd import placement as PL
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p12", [None, (0.5, 0.5, 0.8, 0.2)])
def test_cftp_samples_match_exact_placement(p12):
    n = 10
    d = S.call("domino.AztecDiamond", args=[n])
    tmax, tmin = S.call("domino.MaxTiling", d), S.call("domino.MinTiling", d)
    ctx = rt.Context(0)
    ctx.load_params(default_table(tmax.size // 2))
    rule = rt.RT_DOMINO_CFTP if p12 is None else rt.RT_DOMINO_WEIGHTED
    print("cftp placement", n, p12, flush=True)
    samples = [ctx.domino_cftp(tmax, tmin, 7919 * s + 5, rule=rule, p12=p12, mode=rt.RT_CFTP_DOUBLING)[0]
               for s in range(400)]
    ctx.close()
    z, live = PL.zscores(samples, d, (0.5,) * 4 if p12 is None else p12)
    print("positions", live, "max |z|", float(np.max(np.abs(z))), "mean z^2", float(np.mean(z ** 2)), flush=True)
    assert live > 100
    assert np.max(np.abs(z)) < 5.0
    assert abs(np.mean(z ** 2) - 1) < 0.35
