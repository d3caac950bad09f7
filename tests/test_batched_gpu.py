"""Batched tables: one persistent launch per phase over many tables gives
the same Find/Contains outputs and failure counts as per-table StaticMap
calls and as the oracle; graph replays repeat them exactly."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def mx():
    import mxco
    return mxco


def workload(num_tables=96, cap_log2=(6, 14), find_log2=(2, 12)):
    import mxco
    from mxco import workloads
    return workloads.batched_tables(num_tables, cap_log2, (0.5, 0.85), find_log2, seed=7,
                                    valid_extent=lambda c: mxco.valid_extent(c, 5, "linear"))


@pytest.mark.parametrize("group", [0, 1, 2, 4, 8])
def test_batched_equals_plain_and_oracle(group):
    import mxco
    w = workload()
    T = len(w["caps"])
    b = mxco.BatchedStaticMap(w["caps"], -1, -1)
    assert [b.capacity(t) for t in range(T)] == w["caps"]
    ins = b.plan(mxco.BATCH_INSERT, [(t, w["ins_off"][t], w["n_ins"][t]) for t in range(T)], group)
    fnd = b.plan(mxco.BATCH_FIND, [(t, w["find_off"][t], w["n_find"][t]) for t in range(T)], group)
    con = b.plan(mxco.BATCH_CONTAINS, [(t, w["find_off"][t], w["n_find"][t]) for t in range(T)], group)
    out = torch.empty_like(w["queries"])
    flags = torch.empty(w["queries"].numel(), dtype=torch.uint8, device="cuda")
    ins.run(w["pairs"])
    fnd.run(w["queries"], out)
    con.run(w["queries"], flags)
    assert ins.failures() == [0] * T
    assert torch.equal(out, w["expected"])
    assert torch.equal(flags.bool(), w["expected"] != -1)
    # per-table plain calls and the oracle on a few tables
    pairs_h = w["pairs"].cpu().numpy()
    q_h = w["queries"].cpu().numpy()
    out_h = out.cpu().numpy()
    for t in range(0, T, 17):
        a, n = w["ins_off"][t], w["n_ins"][t]
        qa, qn = w["find_off"][t], w["n_find"][t]
        m = mxco.StaticMap(w["caps"][t], -1, -1)
        assert m.insert(w["pairs"][a:a + n]) == 0
        assert torch.equal(m.find(w["queries"][qa:qa + qn]), out[qa:qa + qn])
        ref = O.OpenAddressing(np.int32, np.int32, w["caps"][t], -1, -1)
        ref.insert(pairs_h[a:a + n, 0], pairs_h[a:a + n, 1])
        assert ref.find(q_h[qa:qa + qn])[0].tobytes() == out_h[qa:qa + qn].tobytes()


def test_batched_duplicates_counted_per_item():
    import mxco
    caps = [64, 1000, 5]
    b = mxco.BatchedStaticMap(caps, -1, -1)
    keys = np.concatenate([np.array([1, 2, 2, 3, -1], np.int32), np.arange(100, dtype=np.int32),
                           np.arange(12, dtype=np.int32)])
    pairs = torch.from_numpy(O.pack_pairs(keys, keys)).cuda()
    ins = b.plan(mxco.BATCH_INSERT, [(0, 0, 5), (1, 5, 100), (2, 105, 12)])
    ins.run(pairs)
    # table 0: one duplicate + the reserved key; table 2 holds 5 slots: 7 fail
    assert ins.failures() == [2, 0, 7]


def test_graph_replays_match_eager():
    import mxco
    w = workload(48)
    T = len(w["caps"])
    b = mxco.BatchedStaticMap(w["caps"], -1, -1)
    ins = b.plan(mxco.BATCH_INSERT, [(t, w["ins_off"][t], w["n_ins"][t]) for t in range(T)])
    fnd = b.plan(mxco.BATCH_FIND, [(t, w["find_off"][t], w["n_find"][t]) for t in range(T)])
    out = torch.empty_like(w["queries"])

    def record(s):
        b.clear(s)
        ins.run(w["pairs"], stream=s)
        fnd.run(w["queries"], out, stream=s)

    g = mxco.Graph(record)
    for rep in range(4):
        out.fill_(123)
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        assert torch.equal(out, w["expected"]), rep
        assert ins.failures(g.stream) == [0] * T
        # eager work between replays
        b.clear()
        torch.cuda.synchronize()


def test_tuner_keeps_results():
    import mxco
    w = workload(32)
    T = len(w["caps"])
    b = mxco.BatchedStaticMap(w["caps"], -1, -1)
    ins = b.plan(mxco.BATCH_INSERT, [(t, w["ins_off"][t], w["n_ins"][t]) for t in range(T)])
    fnd = b.plan(mxco.BATCH_FIND, [(t, w["find_off"][t], w["n_find"][t]) for t in range(T)])
    ins.tune(w["pairs"])
    # the tuner leaves the tables as one run leaves them: every key inserted once
    assert ins.failures() == [0] * T
    out = torch.empty_like(w["queries"])
    fnd.tune(w["queries"], out)
    g, bpc = fnd.config()
    assert g in (1, 2, 4, 8) and bpc > 0
    blk, staged = fnd.launch_shape()
    assert blk in (64, 128, 256) and (blk == 256 or not staged)
    assert ins.launch_shape()[0] in (64, 128, 256)
    out.fill_(0)
    fnd.run(w["queries"], out)
    assert torch.equal(out, w["expected"])


@pytest.mark.parametrize("block", [64, 128, 256])
def test_every_launch_shape_gives_the_same_tables(block):
    """Insert and Find plans under 64-, 128- and 256-thread workgroups (the
    plain probe kernel, and the wave-staged Find at 256) with each lane-group
    width: the same failures and Find outputs."""
    import mxco
    w = workload(48)
    T = len(w["caps"])
    for g in (1, 2, 4, 8):
        b = mxco.BatchedStaticMap(w["caps"], -1, -1)
        ins = b.plan(mxco.BATCH_INSERT, [(t, w["ins_off"][t], w["n_ins"][t]) for t in range(T)])
        fnd = b.plan(mxco.BATCH_FIND, [(t, w["find_off"][t], w["n_find"][t]) for t in range(T)])
        ins.set_config(g, 4)
        ins.set_launch_shape(block, False)
        ins.run(w["pairs"])
        assert ins.failures() == [0] * T
        for staged in ((False, True) if block == 256 else (False,)):
            fnd.set_config(g, 4)
            fnd.set_launch_shape(block, staged)
            out = torch.zeros_like(w["queries"])
            fnd.run(w["queries"], out)
            assert torch.equal(out, w["expected"]), (g, block, staged)
    with pytest.raises(ValueError):
        ins.set_launch_shape(96, False)
    with pytest.raises(ValueError):
        fnd.set_launch_shape(128, True)
