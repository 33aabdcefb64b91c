"""What the trained proposal network gains over the method without it.

On a held-out part (a seed outside the training episodes), the star scene at
130^3 with 96 orientations is packed to exhaustion three ways: the plain
search (every orientation every step), the learned proposal with the committed
trained weights, and the same network untrained (its initial weights).  Every
placement of every run is a feasible placement found by the exact search.  The
trained network packs more parts than the untrained one, close to the plain
search, in far less time per step."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_proposal_beats_untrained_and_is_faster_than_plain(ctx):
    import spectrapack_amd as sp
    from spectrapack_amd import learned as L
    from spectrapack_amd import workloads as W
    if not os.path.exists(L.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights committed")
    trained = sp.Proposal(ctx, max_batch=1)
    trained.set_parameters(L.load_weights())
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
