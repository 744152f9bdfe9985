"""Batched solve: every graph of a batch must stop at the same iteration
count as its own plain solve, with bit-identical beliefs, whatever the tile
size, grid size, chunk length or hipGraph use."""
import numpy as np
import pytest

import fastlbp_amd as fl
from fastlbp_amd.batch import Batch
from tests.helpers import random_alarm_clamps, same_bits

pytestmark = pytest.mark.gpu


def batch_graphs(seed, n_graphs, n_variants):
    rng = np.random.default_rng(seed)
    graphs = []
    for g in range(n_graphs):
        n = int(np.exp(rng.uniform(np.log(300), np.log(30000))))
        graphs.append(fl.Graph(fl.generate(n, seed * 1000 + g).factors))
    base = fl.generate(8000, seed * 1000 + 999)
    for v in range(n_variants):
        labels, values = random_alarm_clamps(rng, base.alarms, 50)
        graphs.append(fl.Graph(fl.clamp(base.factors, labels, values)))
    return graphs


def plain(graph, tol, maxiter, damping):
    p = fl.Plan(graph)
    p.init()
    r = p.run(tol, maxiter, damping)
    return r.iters, r.maxdiff, p.beliefs()


@pytest.mark.parametrize("opts", [dict(), dict(tile_factors=256, tile_vars=512, grid_blocks=37, chunk=5),
                                  dict(use_graph=0, chunk=16), dict(tile_factors=64, tile_vars=96, grid_blocks=3),
                                  dict(layout=1), dict(layout=2),
                                  dict(layout=1, use_graph=0, chunk=9)])
def test_batch_matches_plain_solves(opts):
    graphs = batch_graphs(1, 12, 6)
    tol, maxiter, damping = 1e-9, 300, 0.3
    b = Batch(graphs, **opts)
    b.init()
    iters, md, tot = b.run(tol, maxiter, damping)
    for g, graph in enumerate(graphs):
        it, mdp, bp = plain(graph, tol, maxiter, damping)
        assert iters[g] == it, (g, iters[g], it)
        assert same_bits([md[g]], [mdp])
        assert same_bits(b.beliefs(g), bp), g
    assert tot.sweeps == int(iters.max())


def test_batch_maxiter_cap_and_undamped():
    graphs = batch_graphs(2, 5, 3)
    b = Batch(graphs, chunk=7)
    b.init()
    iters, md, _ = b.run(1e-14, 23, 0.0)
    for g, graph in enumerate(graphs):
        it, _, bp = plain(graph, 1e-14, 23, 0.0)
        assert iters[g] == it and same_bits(b.beliefs(g), bp)


def test_tile_bounds_outside_the_documented_range_are_rejected():
    graphs = batch_graphs(2, 2, 3)
    for bad in (dict(tile_factors=4096, tile_vars=4096), dict(tile_factors=32), dict(tile_vars=1024)):
        with pytest.raises(fl.FlbpError):
            Batch(graphs, **bad)
    b = Batch(graphs)
    with pytest.raises(fl.FlbpError):
        b.set_opts(tile_factors=1024)


def test_autotune_keeps_results():
    graphs = batch_graphs(3, 8, 4)
    b = Batch(graphs)
    best = b.autotune(0.3, trial_sweeps=4)
    assert best["tile_factors"] in (256, 512) and best["layout"] in (1, 2)
    iters, _, _ = b.run(1e-9, 300, 0.3)
    for g, graph in enumerate(graphs):
        it, _, bp = plain(graph, 1e-9, 300, 0.3)
        assert iters[g] == it and same_bits(b.beliefs(g), bp)


def test_undamped_then_damped_runs():
    """Undamped sweeps skip the factor-major f->v copy that damping reads;
    a damped run after them must rebuild it and continue exactly as the
    plain solves do."""
    graphs = batch_graphs(4, 6, 3)
    b = Batch(graphs, chunk=6)
    b.init()
    b.run(1e-14, 5, 0.0)
    iters, md, _ = b.run(1e-9, 80, 0.4)
    for g, graph in enumerate(graphs):
        p = fl.Plan(graph)
        p.init()
        p.run(1e-14, 5, 0.0)
        r = p.run(1e-9, 80, 0.4)
        assert iters[g] == r.iters and same_bits([md[g]], [r.maxdiff])
        assert same_bits(b.beliefs(g), p.beliefs()), g


def test_layout_switch_mid_run():
    """set_opts(layout=...) moves the stored messages: a run continued under
    the other layout matches plain solves that never switched."""
    graphs = batch_graphs(5, 5, 2)
    b = Batch(graphs, chunk=4, layout=2)
    b.init()
    b.run(1e-14, 6, 0.0)
    b.set_opts(layout=1)
    b.run(1e-14, 11, 0.3)
    b.set_opts(layout=2)
    b.run(1e-14, 14, 0.0)
    b.set_opts(layout=1)
    iters, md, _ = b.run(1e-9, 90, 0.3)
    for g, graph in enumerate(graphs):
        p = fl.Plan(graph)
        p.init()
        p.run(1e-14, 6, 0.0)
        p.run(1e-14, 11, 0.3)
        p.run(1e-14, 14, 0.0)
        r = p.run(1e-9, 90, 0.3)
        assert iters[g] == r.iters and same_bits([md[g]], [r.maxdiff])
        assert same_bits(b.beliefs(g), p.beliefs()), g
