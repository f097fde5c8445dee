"""Batched scoring (ac_batch_*) and graph-replayed iterations, through the C ABI.

Bar: each pattern of a batched run equals a plain run of the same pattern (ac_set_graph + ac_run) of the same build
to |a - b| <= 1e-8 + 1e-5 |b|, and the FP64 oracle to the same tolerance; iterations replayed from a captured HIP
graph are bit-identical to launched ones (same kernels, same order), on every replay.
"""
import sys

import numpy as np
import pytest

from application_classification_amd import workloads as W
from application_classification_amd.matcher import AC_OPT_GRAPH, AC_OPT_SLAB_EDGES, AC_OPT_WALK_BATCH, Matcher
from application_classification_amd._lib import ACError, AC_ERR_ARG, AC_ERR_STATE
from oracle import oracle as O
from tests.conftest import allclose_report

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-5, 1e-8


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def g64(g):
    return O.Graph64.of(g.vfeat, g.efeat, g.src, g.dst)


def assert_close(mu, ref, what):
    bad, ratio = allclose_report(mu, ref, RTOL, ATOL)
    say(f"{what}: max err/tol = {ratio:.4f}, violations = {bad}")
    assert bad == 0, f"{what}: {bad} log-probabilities outside rtol={RTOL}, atol={ATOL} (max err/tol {ratio:.3f})"


@pytest.fixture(scope="module")
def small():
    return W.small(0)


@pytest.fixture(scope="module")
def uneven(small):
    """Twelve patterns of 3..24 vertices (edges in [V, 2V]) planted in the small data graph."""
    return W.uneven_patterns(small.data, 12, seed=5, v_range=(3, 24))


def plain(data, patt, n_iter=-1, graph=False):
    with Matcher(0) as m:
        m.set_option(AC_OPT_GRAPH, int(graph))
        m.set_data(data)
        m.set_pattern(patt)
        return m.run(n_iter)


@pytest.mark.parametrize("slab_edges,graph", [(128, True), (16, False), (256, True)])
def test_batch_matches_plain_runs(small, uneven, slab_edges, graph):
    d = small.data
    with Matcher(0) as m:
        m.set_option(AC_OPT_SLAB_EDGES, slab_edges)
        m.set_option(AC_OPT_GRAPH, int(graph))
        m.set_data(d)
        m.batch_set_patterns(uneven)
        m.batch_run()
        info = m.batch_info()
        say(f"slab_edges={slab_edges} graph={graph}: {info}")
        assert info["patterns"] == len(uneven)
        assert info["updates"] == sum(d.num_edges * p.num_edges * p.num_nodes for p in uneven)
        assert info["updates_computed"] >= info["updates"]
        for i, p in enumerate(uneven):
            assert_close(m.batch_mu(i), plain(d, p), f"pattern {i} ({p.num_nodes}/{p.num_edges})")


def test_batch_with_a_pattern_without_edges(small, uneven):
    """An edge-less pattern among others: it only adds vertices to its slab (its columns keep the vertex costs'
    renormalisation), and every member matches its plain run."""
    d = small.data
    e0 = W.Graph(uneven[2].vfeat, np.zeros((0, 13), np.float32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    pats = [uneven[0], e0, uneven[1]]
    with Matcher(0) as m:
        m.set_data(d)
        m.batch_set_patterns(pats)
        m.batch_run()
        for i, p in enumerate(pats):
            assert_close(m.batch_mu(i), plain(d, p), f"pattern {i} ({p.num_nodes}/{p.num_edges})")


def test_batch_against_oracle(small, uneven):
    d = small.data
    with Matcher(0) as m:
        m.set_data(d)
        m.batch_set_patterns(uneven[:4])
        m.batch_run()
        for i, p in enumerate(uneven[:4]):
            assert_close(m.batch_mu(i), O.run(g64(d), g64(p)), f"pattern {i} vs oracle")


def test_batch_explicit_iterations(small, uneven):
    """Per-pattern iteration counts, including 0 (prep state) and counts past the vertex count."""
    d = small.data
    pats = uneven[:5]
    iters = [0, 1, 3, pats[3].num_nodes + 2, 2]
    with Matcher(0) as m:
        m.set_option(AC_OPT_GRAPH, 1)
        m.set_data(d)
        m.batch_set_patterns(pats)
        m.batch_run(iters)
        for i, p in enumerate(pats):
            assert_close(m.batch_mu(i), plain(d, p, iters[i]), f"pattern {i}, {iters[i]} iterations")


def test_batch_rmat_planted(uneven):
    """R-MAT scale 12 (split in-edge lists, vertices without edges) with uneven planted patterns."""
    w = W.rmat(12, 16, 8, 10, seed=1)
    pats = W.uneven_patterns(w.data, 8, seed=2, v_range=(4, 16))
    with Matcher(0) as m:
        m.set_option(AC_OPT_GRAPH, 1)
        m.set_data(w.data)
        m.batch_set_patterns(pats)
        m.batch_run()
        for i, p in enumerate(pats):
            mu = m.batch_mu(i)
            assert_close(mu, plain(w.data, p), f"rmat12 pattern {i}")
            np.testing.assert_array_equal(mu.argmax(axis=0), p.vmap)


def test_batched_workload_rmat20_256_patterns():
    """The batched benchmark's workload exactly: 256 patterns (V in [4, 32], E in [V, 2V]) against R-MAT scale 20
    with 32-dim features, autotuned slabs, graph-replayed iterations; every 32nd pattern against a sequential plain
    run on the same resident data, and every pattern's planted vertices recovered."""
    from application_classification_amd.matcher import AC_OPT_GRAPH

    g, pats = W.batched(0)
    assert len(pats) == 256 and all(p.num_nodes <= p.num_edges <= 2 * p.num_nodes for p in pats)
    say(f"batched workload: {g.num_edges} edges, {sum(p.num_edges for p in pats)} pattern edges")
    with Matcher(0) as m:
        m.set_data(g)
        ref = {}
        for i in range(0, len(pats), 32):
            m.set_pattern(pats[i])
            ref[i] = m.run()
        m.batch_set_patterns(pats)
        (cap, wb), _ = m.batch_autotune(1)
        m.set_option(AC_OPT_GRAPH, 1)
        say(f"autotuned: slab edges {cap}, walk batch {wb}")
        m.batch_run()
        info = m.batch_info()
        say(f"{info}")
        for i, r in ref.items():
            assert_close(m.batch_mu(i), r, f"pattern {i} ({pats[i].num_nodes}/{pats[i].num_edges})")
        missed = [i for i, p in enumerate(pats) if not np.array_equal(m.batch_mu(i).argmax(axis=0), p.vmap)]
        assert not missed, f"planted vertices not recovered for patterns {missed[:10]}"


def test_graph_replay_is_bit_identical(small):
    """Iteration pairs replayed from a captured graph give exactly the launched result, run after run (each run
    replays the pair several times, with other work and synchronisations between runs)."""
    d, p = small.data, small.pattern
    eager = plain(d, p, 9, graph=False)
    with Matcher(0) as m:
        m.set_option(AC_OPT_GRAPH, 1)
        m.set_data(d)
        m.set_pattern(p)
        for rep in range(3):
            np.testing.assert_array_equal(m.run(9), eager)
            np.testing.assert_array_equal(plain(d, p, 4, graph=False), plain(d, p, 4, graph=True))


def test_batch_graph_replay_both_parities_bit_exact(small, uneven):
    """Snapshots after 4 and 9 iterations in one slab: the snapshot at 4 forces iteration 3 to be launched alone, so
    the pairs before it start at odd iterations (1-2) and the pairs after it at even ones (4-5, 6-7).  Each parity
    replays its own captured pair (the scalar slots a pair reads and writes depend on its parity), so the graph run
    is bit-identical to the launched one at both snapshots."""
    d = small.data
    pats = uneven[:2]
    assert sum(p.num_edges for p in pats) <= 256
    out = {}
    for graph in (0, 1):
        with Matcher(0) as m:
            m.set_option(AC_OPT_SLAB_EDGES, 256)
            m.set_option(AC_OPT_GRAPH, graph)
            m.set_data(d)
            m.batch_set_patterns(pats)
            m.batch_run([4, 9])
            assert m.batch_info()["slabs"] == 1
            out[graph] = [m.batch_mu(i) for i in range(len(pats))]
    for i in range(len(pats)):
        say(f"pattern {i}: graph replay bit-identical = {np.array_equal(out[1][i], out[0][i])}")
        np.testing.assert_array_equal(out[1][i], out[0][i])
    assert_close(out[1][1], plain(d, pats[1], 9), "9 iterations vs a plain run")


def test_batch_ignores_the_adjoint_tape(small, uneven):
    """AC_OPT_TAPE configures ac_run's adjoint tape; a batched run on the same context runs untaped (its slabs never
    feed ac_backward) and leaves the option as it was."""
    from application_classification_amd.matcher import AC_OPT_TAPE

    d = small.data
    pats = uneven[:3]
    res = []
    for tape in (0, 4):
        with Matcher(0) as m:
            m.set_option(AC_OPT_TAPE, tape)
            m.set_data(d)
            m.batch_set_patterns(pats)
            m.batch_run()
            assert m.get_option(AC_OPT_TAPE) == tape
            res.append([m.batch_mu(i) for i in range(len(pats))])
    for a, b in zip(res[0], res[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("B", [1, 2, 4])
def test_walk_batch_sizes_agree(small, B):
    """Every edge-walk batch size walks the same bundle layout to the same (exact) maxima."""
    d, p = small.data, small.pattern
    ref = plain(d, p)
    with Matcher(0) as m:
        m.set_option(AC_OPT_WALK_BATCH, B)
        assert m.get_option(AC_OPT_WALK_BATCH) == B
        m.set_data(d)
        m.set_pattern(p)
        np.testing.assert_array_equal(m.run(), ref)


def test_autotune_picks_a_valid_configuration(small, uneven):
    d = small.data
    with Matcher(0) as m:
        m.set_data(d)
        m.batch_set_patterns(uneven)
        (cap, B), est = m.batch_autotune(2)
        say(f"autotune: slab edges {cap}, walk batch {B}, estimated {est:.3f} ms")
        assert cap in (64, 128, 256) and B in (1, 2, 4) and est > 0
        assert m.get_option(AC_OPT_SLAB_EDGES) == cap and m.get_option(AC_OPT_WALK_BATCH) == B
        m.batch_run()
        assert_close(m.batch_mu(0), plain(d, uneven[0]), "after autotune")


def test_batch_errors(small, uneven):
    d = small.data
    with Matcher(0) as m:
        with pytest.raises(ACError) as e:
            m.batch_run()
        assert e.value.code == AC_ERR_STATE
        m.set_data(d)
        bad = W.Graph(uneven[0].vfeat, uneven[0].efeat, uneven[0].src.copy(), uneven[0].dst.copy())
        bad.dst[0] = bad.num_nodes  # endpoint outside the pattern
        with pytest.raises(ACError) as e:
            m.batch_set_patterns([uneven[1], bad])
        assert e.value.code == AC_ERR_ARG
        with pytest.raises(ACError) as e:
            m.set_option(AC_OPT_SLAB_EDGES, 1000)
        assert e.value.code == AC_ERR_ARG
        m.batch_set_patterns(uneven[:2])
        with pytest.raises(ACError) as e:
            m.batch_mu(0)
        assert e.value.code == AC_ERR_STATE


def test_batch_cnull_normalised_per_pattern(small, uneven):
    """With AC_OPT_CNULL each pattern of a slab gets the Cnull of its own edges (the log-softmax runs over one
    pattern's edges, not over the slab), so batched results equal plain runs with the option set."""
    from application_classification_amd.matcher import AC_OPT_CNULL

    d = W.Graph(small.data.vfeat, small.data.efeat * 2.0, small.data.src, small.data.dst)
    pats = [W.Graph(p.vfeat, p.efeat * 2.0, p.src, p.dst) for p in uneven[:5]]
    with Matcher(0) as m:
        m.set_option(AC_OPT_CNULL, 1)
        m.set_option(AC_OPT_SLAB_EDGES, 256)
        m.set_data(d)
        m.batch_set_patterns(pats)
        m.batch_run()
        assert m.batch_info()["slabs"] < len(pats)
        for i, p in enumerate(pats):
            with Matcher(0) as m1:
                m1.set_option(AC_OPT_CNULL, 1)
                m1.set_data(d)
                m1.set_pattern(p)
                assert_close(m.batch_mu(i), m1.run(), f"Cnull pattern {i}")
