"""Adjoint (ac_set_weights / ac_backward) through the C ABI, checked against the FP64 oracle.

Bar: the gradient of the planted cross-entropy  L = -mean_p MU[vmap[p]][p]  with respect to the per-dimension cdist
weights agrees with FP64 central finite differences (h = 1e-4) of the oracle's loss to 1e-3 of the gradient's
largest component; the weighted forward agrees with the weighted oracle to the plain path's bar
(|a - b| <= 1e-8 + 1e-5 |b|); every checkpoint interval gives the same gradient.
"""
import sys

import numpy as np
import pytest

from application_classification_amd import workloads as W
from application_classification_amd.matcher import AC_OPT_TAPE, Matcher
from application_classification_amd._lib import ACError, AC_ERR_STATE
from oracle import oracle as O
from tests.conftest import allclose_report

pytestmark = pytest.mark.gpu

H = 1e-4
GRAD_TOL = 1e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def g64(g):
    return O.Graph64.of(g.vfeat, g.efeat, g.src, g.dst)


def loss_grad_mu(dn, vmap):
    pn = len(vmap)
    g = np.zeros((dn, pn), np.float32)
    g[np.asarray(vmap), np.arange(pn)] = -1.0 / pn
    return g


def oracle_loss(d64, p64, vmap, wv, we, n_iter):
    mu = O.run_weighted(d64, p64, wv, we, n_iter=n_iter)
    return -float(np.mean(mu[np.asarray(vmap), np.arange(len(vmap))]))


def fd_grad(d64, p64, vmap, wv, we, n_iter):
    gv = np.zeros_like(wv)
    ge = np.zeros_like(we)
    for w, g in ((wv, gv), (we, ge)):
        for k in range(w.size):
            w[k] += H
            lp = oracle_loss(d64, p64, vmap, wv, we, n_iter)
            w[k] -= 2 * H
            lm = oracle_loss(d64, p64, vmap, wv, we, n_iter)
            w[k] += H
            g[k] = (lp - lm) / (2 * H)
    return gv, ge


def gpu_grad(data, patt, wv, we, n_iter, K=4):
    with Matcher(0) as m:
        m.set_option(AC_OPT_TAPE, K)
        m.set_data(data)
        m.set_weights(wv, we)
        m.set_pattern(patt)
        mu = m.run(n_iter)
        gv, ge = m.backward(loss_grad_mu(data.num_nodes, patt.vmap))
        return mu, gv.astype(np.float64), ge.astype(np.float64)


def check_grad(g, fd, what):
    scale = max(np.abs(fd).max(), 1e-12)
    err = np.abs(g - fd).max() / scale
    say(f"{what}: |grad - fd|_inf / |fd|_inf = {err:.2e} (|fd|_inf = {scale:.3e})")
    assert err <= GRAD_TOL, f"{what}: gradient off by {err:.2e} of its scale"


@pytest.fixture(scope="module")
def small():
    return W.small(0)


def weights(seed, dv, de):
    r = np.random.default_rng(seed)
    return r.uniform(0.5, 1.5, dv), r.uniform(0.5, 1.5, de)


def test_weighted_forward_matches_oracle(small):
    d, p = small.data, small.pattern
    wv, we = weights(1, 13, 13)
    with Matcher(0) as m:
        m.set_data(d)
        m.set_weights(wv, we)
        m.set_pattern(p)
        mu = m.run()
        bad, ratio = allclose_report(mu, O.run_weighted(g64(d), g64(p), wv, we), 1e-5, 1e-8)
        say(f"weighted forward: max err/tol {ratio:.4f}")
        assert bad == 0
        m.set_weights(None, None)  # back to the reference's cdist
        np.testing.assert_array_equal(m.run(), Matcher_plain(d, p))


def Matcher_plain(d, p):
    with Matcher(0) as m:
        m.set_data(d)
        m.set_pattern(p)
        return m.run()


@pytest.mark.parametrize("n_iter,K", [(0, 4), (3, 4), (-1, 4), (0, 1), (3, 1)])
def test_gradient_matches_finite_differences_small(small, n_iter, K):
    """K = 1 keeps every state (the iterations write straight into their checkpoint slots; zero iterations read the
    prep's slot), K = 4 recomputes segments from checkpoints."""
    d, p = small.data, small.pattern
    wv, we = weights(2, 13, 13)
    iters = p.num_nodes if n_iter < 0 else n_iter
    _, gv, ge = gpu_grad(d, p, wv, we, iters, K=K)
    fv, fe = fd_grad(g64(d), g64(p), p.vmap, wv.copy(), we.copy(), iters)
    check_grad(np.concatenate([gv, ge]), np.concatenate([fv, fe]), f"small, {iters} iterations, K = {K}")


@pytest.mark.parametrize("K", [1, 3])
def test_gradient_for_a_pattern_without_edges(small, K):
    """A pattern with vertices but no edges: every iteration only renormalises the vertex costs (the taped run copies
    its state from slot to slot).  The edge weights do not enter, so their gradient is exactly zero; the vertex
    weights' gradient matches FP64 central differences of the oracle."""
    d, p = small.data, small.pattern
    patt = W.Pattern(p.vfeat, np.zeros((0, 13), np.float32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                     vmap=p.vmap)
    wv, we = weights(4, 13, 13)
    _, gv, ge = gpu_grad(d, patt, wv, we, 3, K=K)
    assert np.all(ge == 0.0)
    fv, _ = fd_grad(g64(d), g64(patt), patt.vmap, wv.copy(), we.copy(), 3)
    check_grad(gv, fv, f"pattern without edges, K = {K}")


@pytest.mark.parametrize("K", [1, 3])
def test_gradient_for_data_without_edges(small, K):
    """A data graph without edges (every vertex isolated, no edge rows): the maxima are the fills, the edge weights
    do not enter (gradient exactly zero), and the vertex weights' gradient matches FP64 central differences of the
    oracle."""
    d, p = small.data, small.pattern
    data = W.Graph(d.vfeat, np.zeros((0, 13), np.float32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    wv, we = weights(5, 13, 13)
    _, gv, ge = gpu_grad(data, p, wv, we, 2, K=K)
    assert np.all(ge == 0.0)
    fv, _ = fd_grad(g64(data), g64(p), p.vmap, wv.copy(), we.copy(), 2)
    check_grad(gv, fv, f"data without edges, K = {K}")


def test_gradient_matches_finite_differences_rmat_subgraph():
    """R-MAT scale 11 (2,048 vertices: hubs, split in-edge lists, vertices without edges), 32-dim features,
    a planted 12-vertex / 16-edge pattern; all 64 weights."""
    w = W.rmat(11, 32, 12, 16, seed=4)
    d, p = w.data, w.pattern
    wv, we = weights(3, 32, 32)
    _, gv, ge = gpu_grad(d, p, wv, we, p.num_nodes)
    fv, fe = fd_grad(g64(d), g64(p), p.vmap, wv.copy(), we.copy(), p.num_nodes)
    check_grad(np.concatenate([gv, ge]), np.concatenate([fv, fe]), "rmat11 d32")


@pytest.mark.parametrize("K", [1, 3, 100])
def test_checkpoint_intervals_agree(small, K):
    d, p = small.data, small.pattern
    wv, we = weights(5, 13, 13)
    mu4, gv4, ge4 = gpu_grad(d, p, wv, we, -1, K=4)
    muk, gvk, gek = gpu_grad(d, p, wv, we, -1, K=K)
    np.testing.assert_array_equal(mu4, muk)  # taping does not change the forward
    g4, gk = np.concatenate([gv4, ge4]), np.concatenate([gvk, gek])
    assert np.abs(g4 - gk).max() <= 1e-5 * np.abs(g4).max()


def test_taped_forward_equals_plain_forward(small):
    d, p = small.data, small.pattern
    mu, _, _ = gpu_grad(d, p, np.ones(13), np.ones(13), -1)
    np.testing.assert_array_equal(mu, Matcher_plain(d, p))


def test_profiled_taped_run_reports_every_stage(small):
    """ac_run_profiled of a run with a checkpoint at every iteration (whose iterations write straight into the
    checkpoint slots) times each stage, as a plain profiled run does, and returns the same MU."""
    d, p = small.data, small.pattern
    st = {}
    mus = []
    for tape in (0, 1):
        with Matcher(0) as m:
            m.set_option(AC_OPT_TAPE, tape)
            m.set_data(d)
            m.set_pattern(p)
            st[tape] = m.run_profiled()
            mus.append(m.mu())
            say(f"tape={tape}: {st[tape]}")
    np.testing.assert_array_equal(mus[0], mus[1])
    for s in ("vertex_a", "reduce_a", "edge_walk", "vertex_c", "reduce_c", "output"):
        assert st[1][s] > 0.0, s
    # The output stage is one pass, not the whole loop.
    assert st[1]["output"] < st[1]["edge_walk"] + st[1]["vertex_a"] + st[1]["vertex_c"]


def test_packed_costs_give_identical_gradients():
    """The taped forward and the recompute walks of ac_backward read the 24-bit cost codes (AC_OPT_PACK_COSTS) where
    bundles pack; the forward and both gradients equal those with FP32 costs exactly.  64-dim features, so most
    bundles pack."""
    from application_classification_amd.matcher import AC_OPT_PACK_COSTS

    w = W.rmat(11, 64, 10, 14, seed=6)
    d, p = w.data, w.pattern
    wv, we = weights(7, 64, 64)
    res = []
    for pack in (0, 1):
        with Matcher(0) as m:
            m.set_option(AC_OPT_PACK_COSTS, pack)
            m.set_option(AC_OPT_TAPE, 3)
            m.set_data(d)
            m.set_weights(wv, we)
            m.set_pattern(p)
            mu = m.run()
            nb, packed = m.cost_packing()
            gv, ge = m.backward(loss_grad_mu(d.num_nodes, p.vmap))
            res.append((mu, gv, ge))
            say(f"pack={pack}: {packed} of {nb} bundles packed")
            if pack:
                assert packed > nb // 2
    for a, b in zip(res[0], res[1]):
        np.testing.assert_array_equal(a, b)


def test_backward_is_reproducible():
    """R-MAT scale 12: hub sources are the recorded argmax in-edge of many destinations at once, and split hubs spread
    over several CSR pieces.  A source's edge terms are summed in fixed point (integer additions, whose result does
    not depend on their order), so two backward passes (and two separate contexts) give bit-identical gradients."""
    w = W.rmat(12, 16, 10, 14, seed=8)
    d, p = w.data, w.pattern
    outdeg = np.bincount(d.src, minlength=d.num_nodes)
    assert outdeg.max() > 128  # split CSR runs
    wv, we = weights(9, 16, 16)
    gm = loss_grad_mu(d.num_nodes, p.vmap)
    res = []
    for ctx in range(2):
        with Matcher(0) as m:
            m.set_option(AC_OPT_TAPE, 4)
            m.set_data(d)
            m.set_weights(wv, we)
            m.set_pattern(p)
            m.run()
            for rep in range(2):
                res.append(np.concatenate(m.backward(gm)))
    for r in res[1:]:
        np.testing.assert_array_equal(r, res[0])


def test_full_tape_equals_recomputed_segments():
    """With a checkpoint at every iteration (AC_OPT_TAPE = 1) the run's own walks record their argmax rows and
    ac_backward recomputes nothing; with longer intervals it recomputes each segment from its checkpoint.  The
    recording walk leaves the same maxima as the plain one (split hubs included: R-MAT 12, in-lists of more than
    128 edges), so the two tapes are the same and so are the forward output and both weight gradients, bit for bit."""
    w = W.rmat(12, 16, 10, 14, seed=8)
    d, p = w.data, w.pattern
    assert np.bincount(d.dst, minlength=d.num_nodes).max() > 128  # split in-edge lists
    wv, we = weights(9, 16, 16)
    gm = loss_grad_mu(d.num_nodes, p.vmap)
    res = {}
    for K in (1, 4):
        with Matcher(0) as m:
            m.set_option(AC_OPT_TAPE, K)
            m.set_data(d)
            m.set_weights(wv, we)
            m.set_pattern(p)
            mu = m.run()
            res[K] = (mu, np.concatenate(m.backward(gm)))
    say(f"full tape vs K = 4: MU bit-identical = {np.array_equal(res[1][0], res[4][0])}, "
        f"gradients bit-identical = {np.array_equal(res[1][1], res[4][1])}")
    np.testing.assert_array_equal(res[1][0], res[4][0])
    np.testing.assert_array_equal(res[1][1], res[4][1])


def test_backward_scales_exactly_with_the_output_gradient():
    """The adjoint is linear in the output gradient, and every scale inside it follows the gradients: the fixed-point
    scale of the source sums is a power of two set by the largest edge-term gradient.  Multiplying grad_mu by 2^k
    therefore multiplies both weight gradients by 2^k, exactly wherever no intermediate value is subnormal (printed:
    whether the result is bit-identical); the check allows 1e-6 of the gradients' scale for such values.  A wrong
    fixed-point scale errs by powers of two at the hub sources."""
    w = W.rmat(12, 16, 10, 14, seed=8)
    d, p = w.data, w.pattern
    wv, we = weights(9, 16, 16)
    gm = loss_grad_mu(d.num_nodes, p.vmap).astype(np.float32)
    with Matcher(0) as m:
        m.set_option(AC_OPT_TAPE, 4)
        m.set_data(d)
        m.set_weights(wv, we)
        m.set_pattern(p)
        m.run()
        base = np.concatenate(m.backward(gm))
        assert np.abs(base).max() > 0
        for k in (20, -20):
            got = np.concatenate(m.backward(np.ldexp(gm, k)))
            want = np.ldexp(base, k)
            err = np.abs(got.astype(np.float64) - want).max() / np.abs(want).max()
            say(f"2^{k} x grad_mu: bit-identical = {np.array_equal(got, want)}, max |diff| / scale = {err:.2e}")
            assert err <= 1e-6


def test_weights_are_learned():
    """Gradient descent on the per-dimension weights lowers the planted cross-entropy: a pattern planted with
    feature noise sigma = 0.3 (an uncertain match), weights from 1; each step w <- max(w - lr g, 0) with lr halved
    until the loss falls.  A gradient pointing uphill would leave no step size that lowers it."""
    base = W.small(0).data
    p = W.planted_pattern(base, 10, 12, seed=12, sigma=0.3)
    wv, we = np.ones(13), np.ones(13)

    def loss_and_grad(m, wv, we):
        m.set_weights(wv, we)
        mu = m.run()
        loss = -float(np.mean(mu[p.vmap, np.arange(p.num_nodes)].astype(np.float64)))
        gv, ge = m.backward(loss_grad_mu(base.num_nodes, p.vmap))
        return loss, gv.astype(np.float64), ge.astype(np.float64)

    with Matcher(0) as m:
        m.set_option(AC_OPT_TAPE, 4)
        m.set_data(base)
        m.set_pattern(p)
        loss, gv, ge = loss_and_grad(m, wv, we)
        losses = [loss]
        lr = 1.0 / max(np.abs(np.concatenate([gv, ge])).max(), 1e-12)
        for step in range(5):
            for _ in range(12):
                nv, ne = np.maximum(wv - lr * gv, 0.0), np.maximum(we - lr * ge, 0.0)
                nl, ngv, nge = loss_and_grad(m, nv, ne)
                if nl < loss:
                    break
                lr *= 0.5
            else:
                pytest.fail(f"step {step}: no step size lowers the loss {loss:.6f}")
            wv, we, loss, gv, ge = nv, ne, nl, ngv, nge
            losses.append(loss)
        say("planted cross-entropy over the steps: " + ", ".join(f"{x:.5f}" for x in losses))
    assert losses[-1] < 0.9 * losses[0]


def test_backward_requires_a_taped_run(small):
    with Matcher(0) as m:
        m.set_data(small.data)
        m.set_pattern(small.pattern)
        m.run()
        with pytest.raises(ACError) as e:
            m.backward(np.zeros((small.data.num_nodes, small.pattern.num_nodes), np.float32))
        assert e.value.code == AC_ERR_STATE


def test_autograd_function(small):
    torch = pytest.importorskip("torch")
    from application_classification_amd.autograd import match_log_probs, planted_cross_entropy

    d, p = small.data, small.pattern
    wv, we = weights(6, 13, 13)
    with Matcher(0) as m:
        m.set_option(AC_OPT_TAPE, 4)
        m.set_data(d)
        m.set_pattern(p)
        tv = torch.tensor(wv, dtype=torch.float32, requires_grad=True)
        te = torch.tensor(we, dtype=torch.float32, requires_grad=True)
        loss = planted_cross_entropy(match_log_probs(m, tv, te), p.vmap)
        loss.backward()
    _, gv, ge = gpu_grad(d, p, wv.astype(np.float32), we.astype(np.float32), -1)
    np.testing.assert_allclose(tv.grad.numpy(), gv, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(te.grad.numpy(), ge, rtol=1e-5, atol=1e-7)
