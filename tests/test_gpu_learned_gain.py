"""What the learned proposal gains over the plain search, and what the trained
network gains over the same method without training.

* The restricted proposal (the network's top orientations in its top
  regions, full search as fallback): on a held-out part (a seed outside the
  training episodes), the star scene at 130^3 with 96 orientations is packed
  to exhaustion with the plain search, the committed trained weights and the
  same network untrained.  The trained network packs more parts than the
  untrained one, close to the plain search, in far less time per step.
* The lazy proposal (Proposal.lazy_step): with every orientation allowed per
  step it is the plain search, choice for choice; with its default setting
  (12 orientations, prior weight 2) it packs at least 99% of the plain
  search's parts at three times its steps/s or more on the three held-out
  256^3 parts, the target the learned form is held to.
Every placement of every run is a feasible placement found by the exact
search."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _weights_or_skip():
    from spectrapack_amd import learned as L
    if not os.path.exists(L.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights committed")
    return L.load_weights()


def test_trained_proposal_beats_untrained_and_is_faster_than_plain(ctx):
    import spectrapack_amd as sp
    from spectrapack_amd import learned as L
    from spectrapack_amd import workloads as W
    trained = sp.Proposal(ctx, max_batch=1)
    trained.set_parameters(_weights_or_skip())
    untrained = sp.Proposal(ctx, max_batch=1)
    untrained.set_parameters(L.initial_parameters(1))
    sc = L.Scene(ctx, W.star_so3_box(9100))
    print("[gain] plain", flush=True)
    plain = L.run_episode(sc)
    print(f"[gain] plain {plain}", flush=True)
    learned = L.run_episode(sc, trained, 8, 512)
    print(f"[gain] learned {learned}", flush=True)
    base = L.run_episode(sc, untrained, 8, 512)
    print(f"[gain] untrained {base}", flush=True)
    assert learned["parts"] >= 0.93 * plain["parts"]
    assert learned["parts"] > base["parts"]
    assert learned["fallbacks"] < base["fallbacks"]
    assert learned["steps_per_s"] > 2.0 * plain["steps_per_s"]
    sc.close()


def test_lazy_step_with_every_orientation_is_the_plain_search(ctx):
    """Two sessions on the same scene: one takes lazy steps that may search
    every blocked orientation, the other plain steps.  Every step places the
    same orientation at the same corner, both end with the same occupancy,
    and after the first step the lazy steps search far fewer orientations
    than the 96 of a plain step."""
    import spectrapack_amd as sp
    from spectrapack_amd import learned as L
    from spectrapack_amd import workloads as W
    prop = sp.Proposal(ctx, max_batch=1)
    prop.set_parameters(_weights_or_skip())
    w = W.star_so3_box(9101)
    lazy, plain = L.Scene(ctx, w), L.Scene(ctx, w)
    prop.set_prior(lazy.prior, 2.0)
    prop.lazy_reset()
    searched = []
    steps = 0
    while True:
        a = prop.lazy_step(lazy.session, lazy.desc, 96)
        b = plain.session.step(stamp=True)
        assert (a["orientation"], tuple(a["translation"])) == (b["orientation"], tuple(b["translation"])), steps
        steps += 1
        if b["orientation"] < 0:
            break
        searched.append(a["searched"])
    print(f"[lazy] {steps} steps, searched per step {np.mean(searched[1:]):.1f}", flush=True)
    assert steps > 50
    assert np.array_equal(lazy.session.environment(), plain.session.environment())
    assert np.mean(searched[1:]) < 48
    lazy.close()
    plain.close()


def test_lazy_proposal_meets_the_packing_target(ctx):
    """Held-out parts 9000-9002 at 256^3 (the bench's learned workload): the
    lazy proposal with 12 orientations and prior weight 2 packs >= 99% of the
    plain search's parts at >= 3x its steps/s, summed over the three."""
    import spectrapack_amd as sp
    from spectrapack_amd import learned as L
    from spectrapack_amd import workloads as W
    prop = sp.Proposal(ctx, max_batch=1)
    prop.set_parameters(_weights_or_skip())
    tot = dict(pp=0, ps=0, pt=0.0, lp=0, ls=0, lt=0.0)
    for seed in (9000, 9001, 9002):
        sc = L.Scene(ctx, W.star_so3_fine_box(seed))
        pl = L.run_episode(sc)
        lz = L.run_episode_lazy(sc, prop, 12, prior_weight=2.0)
        print(f"[target] seed {seed}: plain {pl['parts']} parts {pl['steps_per_s']:.0f}/s, lazy {lz['parts']} parts "
              f"{lz['steps_per_s']:.0f}/s, {lz['searched_per_step']:.1f} searched per step", flush=True)
        tot["pp"] += pl["parts"]
        tot["ps"] += pl["steps"]
        tot["pt"] += pl["seconds"]
        tot["lp"] += lz["parts"]
        tot["ls"] += lz["steps"]
        tot["lt"] += lz["seconds"]
        sc.close()
    assert tot["lp"] >= 0.99 * tot["pp"], tot
    assert tot["ls"] / tot["lt"] >= 3.0 * tot["ps"] / tot["pt"], tot
