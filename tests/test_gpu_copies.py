"""Placed copies on the device against the oracle's BlockedField
(conservative_fields.cpp:1666-1887): per-cell placed counts, removal by copy
id, and the steps after a removal, in full and incremental sessions."""
import numpy as np
import pytest

from test_oracle import sol02_blocker

pytestmark = pytest.mark.gpu


def test_sol02_counts_on_the_device(ctx, oracle):
    """conservative_fields_tests.cpp:63-91 through a session: two copies at the
    same pose count 2 at (5,3,3), removal counts down, a missing id fails."""
    import spectrapack_amd as sp
    _, mask, fp, kernel = sol02_blocker(oracle)
    s = sp.Session(ctx, (7, 7, 7), (0, 0, 0), 1)
    s.set_environment(mask)
    s.set_kernel(0, kernel["cells"], kernel["shape"], kernel["first"])
    s.set_pair_clearance(1.5, 1.0)
    print("[copies] sol02", flush=True)
    s.stamp(0, (3, 3, 3))
    s.stamp(0, (3, 3, 3))
    counts = s.placed_counts()
    assert counts[3, 3, 5] == 2 and counts[3, 5, 5] == 2 and counts[5, 5, 5] == 0
    assert np.array_equal(counts, 2 * fp.astype(np.uint32))
    assert [(c["copy_id"], c["orientation"], c["translation"], c["removed"]) for c in s.copies()] == \
        [(0, 0, (3, 3, 3), False), (1, 0, (3, 3, 3), False)]
    for bad in (7, 2):
        with pytest.raises(sp.Failure) as e:
            s.remove_copy(bad)
        assert e.value.code == "FIELD_COPY_ID"
    s.remove_copy(0)
    assert s.placed_counts()[3, 3, 5] == 1 and s.environment()[3, 3, 5] == 1
    with pytest.raises(sp.Failure) as e:
        s.remove_copy(0)
    assert e.value.code == "FIELD_COPY_ID"
    s.remove_copy(1)
    assert s.placed_counts().sum() == 0 and s.environment().sum() == 0
    assert [c["removed"] for c in s.copies()] == [True, True]
    s.close()


@pytest.mark.parametrize("incremental", [False, True])
@pytest.mark.parametrize("clearance", [0.0, 1.5])
def test_steps_after_removals_match_oracle(ctx, oracle, incremental, clearance):
    """Steps, removals of some placed copies (whose clearance zones overlap
    their neighbours'), then more steps: every choice, overlap field and
    occupancy equals the oracle's BlockedField scene."""
    import blocked as B
    import spectrapack_amd as sp
    from spectrapack_amd import workloads as W
    # 62^3 cells: the update window (36^3 with the 1.5 mm clearance's halo)
    # is under half the 64^3 FFT grid, so incremental steps are effective
    dims = (60.0, 60.0, 60.0)
    first, shape = W.box_window(dims, 1.0)
    mask = oracle.box_container(first, shape, 1.0, dims)
    part = W.cuboid(10.0, 6.0, 4.0)
    kernels = []
    for q in oracle.cube_catalog()[:4]:
        f, s_, c = oracle.voxelize_object(part.vertices, part.triangles, q, 1.0)
        kernels.append(dict(first=f, shape=s_, cells=c))
    stencil = oracle.clearance_stencil(1.0, clearance) if clearance else None
    s = sp.Session(ctx, shape, first, len(kernels))
    s.set_environment(mask)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    if clearance:
        s.set_pair_clearance(clearance, 1.0)
    s.keep_fields(True)
    if incremental:
        s.set_incremental(True)
    ref = B.BlockedField(mask)
    placed = []
    plan = ["step"] * 9 + ["remove 1", "remove 4", "step", "step", "remove 0", "step", "step", "step", "remove 8",
                           "remove 2", "step", "step"]
    for i, act in enumerate(plan):
        print(f"[copies] incremental={incremental} clearance={clearance} action {i}: {act}", flush=True)
        if act == "step":
            r = s.step(stamp=True)
            o, t, fields = oracle.place_step(ref.blocked(), first, kernels, direct=True, fields=True)
            assert r["orientation"] == o and tuple(r["translation"]) == tuple(t), (i, r, o, t)
            for k in range(len(kernels)):
                got = s.overlap_field(k)
                assert (got is None and fields[k] is None) or np.array_equal(got, fields[k]), (i, k)
            fp = oracle.stamp(np.zeros_like(mask), first, kernels[o], t, stencil)
            ref.add(str(len(placed)), fp)
            placed.append((o, tuple(int(v) for v in t)))
        else:
            cid = int(act.split()[1])
            s.remove_copy(cid)
            ref.remove(str(cid))
        assert np.array_equal(s.environment(), ref.blocked()), i
        assert np.array_equal(s.placed_counts(), ref.counts), i
    assert [(c["orientation"], c["translation"]) for c in s.copies()] == placed
    s.close()
