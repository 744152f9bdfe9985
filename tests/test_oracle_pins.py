"""Pins of the oracle against what the mathematics fixes (CPU only).

- Every f->v type code (csrc/kernel/causal_bp_parall.cu:94-398) against
  exact sum-product enumeration over the factor's variables, in rational
  arithmetic, including zero messages, head masks, zero mass and the e1
  cancellation regime (a dropped term, a swapped state or a wrong index
  fails these).
- The product + leave-one-out v->f message (20-92) against the exact
  leave-one-out product, zero counts included.
- Whole-graph PARALL BP on random trees against exact marginals by joint
  enumeration (BP is exact on trees).
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import fastlbp_amd as fl
import oracle

REL = 1e-13


def cpd(kind, x, h, p, q, clamped, mask):
    """Factor value for a full assignment x (tuple of 0/1) of its slots."""
    p, q = Fraction(p), Fraction(q)
    xh = x[h]
    if kind == 0:
        v = p if xh == 1 else 1 - p
        return v
    body = [x[j] for j in range(len(x)) if j != h]
    if kind == 1:      # AND: P(h=1) = p if every body var is 1 else q
        p1 = p if all(b == 1 for b in body) else q
        v = p1 if xh == 1 else 1 - p1
    else:              # OR: P(h=0) = p if every body var is 0 else q
        p0 = p if all(b == 0 for b in body) else q
        v = p0 if xh == 0 else 1 - p0
    if clamped:
        v *= Fraction(mask[xh])
    return v


def exact_fv(kind, m, t, h, p, q, clamped=False, mask=(1.0, 1.0)):
    d = len(m)
    M = [Fraction(0), Fraction(0)]
    for x in itertools.product((0, 1), repeat=d):
        w = cpd(kind, x, h, p, q, clamped, mask)
        for j in range(d):
            if j != t:
                w *= Fraction(m[j][x[j]])
        M[x[t]] += w
    s = M[0] + M[1]
    if s == 0:
        if clamped and t == h and kind in (1, 2):
            if mask[0] == 0:
                return (0.0, 1.0)
            if mask[1] == 0:
                return (1.0, 0.0)
        return (0.5, 0.5)
    return (float(M[0] / s), float(M[1] / s))


def reference_zero_head_quirk(kind, m, t, h, p, q, mask):
    """Body targets whose head-side mass Q is exactly zero: the reference
    skips the loop over the other body messages (`if (prod_10 != 0)`,
    causal_bp_parall.cu:294,320,346,372), and scale(0, 0) is 0/0.  Returns the
    reference's closed-form answer there, or None when Q != 0."""
    if kind == 0 or t == h:
        return None
    p, q = Fraction(p), Fraction(q)
    hx, hy = (m[h][1], m[h][0]) if kind == 1 else (m[h][0], m[h][1])
    mx, my = (mask[1], mask[0]) if kind == 1 else (mask[0], mask[1])
    P = (p - q) * (Fraction(hx) * Fraction(mx) - Fraction(hy) * Fraction(my))
    Q = q * Fraction(hx) * Fraction(mx) + (1 - q) * Fraction(hy) * Fraction(my)
    if Q != 0:
        return None
    if P > 0:
        return (0.0, 1.0) if kind == 1 else (1.0, 0.0)
    return (float("nan"), float("nan"))


def code_of(kind, clamped, is_head):
    if kind == 0:
        return 0
    if kind == 1:
        return (3 if is_head else 7) if clamped else (1 if is_head else 5)
    return (4 if is_head else 8) if clamped else (2 if is_head else 6)


def close(a, b, rel=REL, abs_=1e-13):
    """Messages are normalized pairs, so an absolute floor of ~1e-13 is the
    precision bar; components far below it come from cancellation inside
    the reference's own closed forms (the e1 test pins that regime)."""
    return all(abs(x - y) <= max(rel * max(abs(x), abs(y)), abs_) for x, y in zip(a, b))


def random_msgs(rng, d, specials):
    m = []
    for _ in range(d):
        r = rng.random()
        if specials and r < 0.15:
            m.append((0.0, 1.0))
        elif specials and r < 0.3:
            m.append((1.0, 0.0))
        elif specials and r < 0.4:
            m.append((1e-20 * rng.random(), 1.0))
        elif specials and r < 0.5:
            m.append((1.0, 1e-20 * rng.random()))
        else:
            a = rng.random()
            m.append((a, 1.0 - a) if rng.random() < 0.5 else (rng.random(), rng.random()))
    return m


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("clamped", [False, True])
@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 6])
def test_fv_codes_match_enumeration(kind, clamped, d):
    rng = np.random.default_rng(1000 * kind + 10 * d + clamped)
    masks = [(1.0, 1.0), (0.0, 1.0), (1.0, 0.0), (0.0, 0.0)]
    for trial in range(60):
        m = random_msgs(rng, d, specials=trial % 2 == 1)
        h = int(rng.integers(0, d))
        p = float(rng.random()) if trial % 3 else float(rng.choice([0.0, 1.0]))
        q = float(rng.random()) if trial % 4 else 0.0
        mask = masks[trial % 4] if clamped else (1.0, 1.0)
        for t in range(d):
            code = code_of(kind, clamped, t == h)
            got = oracle.fv_message(code, [a for a, _ in m], [b for _, b in m], t, h, p, q, *mask)
            quirk = reference_zero_head_quirk(kind, m, t, h, p, q, mask if clamped else (1.0, 1.0))
            if quirk is not None:
                assert np.array_equal(np.array(got), np.array(quirk), equal_nan=True), (got, quirk)
                continue
            want = exact_fv(kind, m, t, h, p, q, clamped, mask)
            assert close(got, want, rel=1e-12), (kind, clamped, d, t, h, p, q, mask, m, got, want)


def test_singleton_code():
    for p in (0.0, 0.3, 0.99, 1.0):
        assert oracle.fv_message(0, [0.2], [0.8], 0, 0, p) == (1 - p, p)


def test_e1_cancellation_term():
    """Body messages (eps, 1) with eps below half an ulp of 1: Z - pi1 is lost
    in floating point and only the e1 term (causal_bp_parall.cu:127-141)
    keeps P(h=1) = q * sum(eps) instead of 0."""
    eps = [1e-20, 3e-21, 7e-22]
    m0 = [0.5] + eps            # slot 0 = head
    m1 = [0.5] + [1.0] * 3
    got = oracle.fv_message(1, m0, m1, 0, 0, 0.0, 0.5)
    want = exact_fv(1, list(zip(m0, m1)), 0, 0, 0.0, 0.5)
    assert want[1] > 1e-21 and close(got, want, rel=1e-9, abs_=0.0)
    got_or = oracle.fv_message(2, m1, m0, 0, 0, 0.0, 0.5)   # OR mirror: states swapped
    want_or = exact_fv(2, list(zip(m1, m0)), 0, 0, 0.0, 0.5)
    assert want_or[0] > 1e-21 and close(got_or, want_or, rel=1e-9, abs_=0.0)


def exact_vf(m, src):
    P = [Fraction(1), Fraction(1)]
    for j, (a, b) in enumerate(m):
        if j != src:
            P[0] *= Fraction(a)
            P[1] *= Fraction(b)
    s = P[0] + P[1]
    if s == 0:
        return (0.5, 0.5)
    return (float(P[0] / s), float(P[1] / s))


@pytest.mark.parametrize("d", [1, 2, 3, 5, 9, 20])
def test_vf_leave_one_out_matches_exact(d):
    rng = np.random.default_rng(d)
    for trial in range(100):
        m = random_msgs(rng, d, specials=trial % 2 == 0)
        for src in range(d):
            got = oracle.vf_message([a for a, _ in m], [b for _, b in m], src)
            want = exact_vf(m, src)
            assert close(got, want, rel=1e-12), (m, src, got, want)


# -------------------------------------------------------------------- trees

def random_tree_list(rng, n_vars_max=12):
    """A causal factor graph whose factor graph is a tree."""
    facs = [(0, 0, [], float(rng.uniform(0.2, 0.95)))]
    n = 1
    while n < n_vars_max - 1:
        anchor = int(rng.integers(0, n))
        k = int(rng.integers(1, 3))
        new = list(range(n, n + k))
        n += k
        kind = int(rng.integers(1, 3))
        if rng.random() < 0.5:
            head, body = anchor, new
        else:
            head, body = new[0], [anchor] + new[1:]
        facs.append((kind, head, body, float(rng.uniform(0.6, 0.99)), float(rng.uniform(0.0, 0.2))))
        for v in new:
            if rng.random() < 0.6:
                facs.append((0, v, [], float(rng.uniform(0.1, 0.9))))
    return fl.FactorList.build(facs)


def exact_marginals(flist):
    labels = sorted(set(int(x) for x in flist.head) | set(int(x) for x in flist.body))
    idx = {l: i for i, l in enumerate(labels)}
    V = len(labels)
    Z = Fraction(0)
    P1 = [Fraction(0)] * V
    for x in itertools.product((0, 1), repeat=V):
        w = Fraction(1)
        for I in range(flist.n_factors):
            vs = sorted(set([int(flist.head[I])] + [int(b) for b in flist.body[flist.body_off[I]:flist.body_off[I + 1]]]))
            xs = tuple(x[idx[v]] for v in vs)
            h = vs.index(int(flist.head[I]))
            w *= cpd(int(flist.type[I]), xs, h, flist.p[I], flist.q[I], bool(flist.head_clamped[I]),
                     (flist.mask0[I], flist.mask1[I]))
            if w == 0:
                break
        Z += w
        for i in range(V):
            if x[i]:
                P1[i] += w
    return np.array([float(v / Z) for v in P1])


@pytest.mark.parametrize("seed", range(8))
def test_bp_exact_on_trees(seed):
    rng = np.random.default_rng(seed)
    flist = random_tree_list(rng)
    if seed % 2:      # a head clamp keeps the graph a tree
        flist = fl.clamp(flist, [int(flist.head[1])], [seed % 4 // 2])
    o = oracle.OracleGraph(flist)
    o.init()
    it, md = o.run(1e-15, 300)
    assert md <= 1e-15
    np.testing.assert_allclose(o.beliefs(), exact_marginals(flist), rtol=0, atol=1e-10)


def test_closed_form_and_over_two_leaves():
    """P(h=1) = p*a*b + q*(1 - a*b) for h = AND(x, y) with priors a, b."""
    a, b, p, q = 0.7, 0.4, 0.9, 0.05
    flist = fl.FactorList.build([(0, 1, [], a), (0, 2, [], b), (1, 3, [1, 2], p, q)])
    o = oracle.OracleGraph(flist)
    o.init()
    o.run(1e-15, 100)
    assert abs(o.beliefs()[2] - (p * a * b + q * (1 - a * b))) < 1e-15


def test_damping_fixed_point_and_first_sweep():
    """Damping changes the path, not the fixed point (trees); the first
    sweep's max diff equals max b because old beliefs start at 0."""
    rng = np.random.default_rng(5)
    flist = random_tree_list(rng)
    o = oracle.OracleGraph(flist)
    o.init()
    d0 = o.sweep(0.0)
    assert d0 == np.max(o.beliefs())
    o2 = oracle.OracleGraph(flist)
    o2.init()
    o2.run(1e-15, 2000, damping=0.5)
    np.testing.assert_allclose(o2.beliefs(), exact_marginals(flist), atol=1e-10)


# ------------------------------------------------------------------ damping
#
# Damping is build-defined (the reference has none): after the factor pass
# has normalized an f->v message, m_fv <- (1 - d) * m_new + d * m_fv_old, on
# both states, with the old state starting at init's unnormalized (1, 1)
# (csrc/ext/causal_bp.cu:824-856).  The pins below use d = 0.3 and 0.8,
# where d != 1 - d, so swapped weights, damping applied before the
# normalization, or an old state other than the previous sweep's damped
# message each fail one of them.

def _fv_by_enumeration(flist, dump, vf):
    """Every var-major f->v message recomputed from the factor-major v->f
    messages vf[E, 2] by exact enumeration (undamped, normalized)."""
    fac_off = dump["fac_off"].astype(np.int64)
    var_src = dump["var_src"].astype(np.int64)
    head_slot = dump["head_slot"].astype(np.int64)
    out = []
    for s in var_src:
        I = int(np.searchsorted(fac_off, s, side="right") - 1)
        lo, hi = int(fac_off[I]), int(fac_off[I + 1])
        m = [(float(vf[j, 0]), float(vf[j, 1])) for j in range(lo, hi)]
        kind = int(flist.type[I])
        clamped = bool(flist.head_clamped[I])
        mask = (float(flist.mask0[I]), float(flist.mask1[I]))
        if kind == 0:
            out.append((1.0 - float(flist.p[I]), float(flist.p[I])))
            continue
        out.append(exact_fv(kind, m, int(s) - lo, int(head_slot[I]) - lo, float(flist.p[I]),
                            float(flist.q[I]), clamped, mask))
    return np.array(out)


@pytest.mark.parametrize("d", [0.3, 0.8])
def test_damping_weights_first_and_second_sweep(d):
    """Sweep 1 from init: each f->v pair is (1-d) * normalize(new) + d * (1, 1),
    with normalize(new) from exact enumeration over the factor (incoming v->f
    messages all (1, 1)).  Sweep 2: (1-d) * normalize(new) + d * (sweep 1's
    damped message), new again by enumeration from the exported v->f state."""
    gen = fl.generate(40, 11, k_max=3, b_max=3)
    flist = gen.factors
    o = oracle.OracleGraph(flist)
    dump = o.dump()
    E = o.n_edges
    o.init()
    ones = np.ones((E, 2))
    new1 = _fv_by_enumeration(flist, dump, ones)
    o.sweep(d)
    fv1, vf1 = o.messages()
    want1 = (1 - d) * new1 + d * 1.0
    np.testing.assert_allclose(fv1, want1, rtol=1e-13, atol=1e-15)
    swapped = d * new1 + (1 - d) * 1.0
    assert np.max(np.abs(swapped - want1)) > 1e-3          # the pin discriminates
    new2 = _fv_by_enumeration(flist, dump, vf1)
    o.sweep(d)
    fv2, _ = o.messages()
    want2 = (1 - d) * new2 + d * fv1
    np.testing.assert_allclose(fv2, want2, rtol=1e-13, atol=1e-15)
    assert np.max(np.abs((1 - d) * new2 + d * want1 - want2)) < 1e-12
    assert np.max(np.abs((1 - d) * new2 + d * new1 - want2)) > 1e-4   # old = damped, not undamped


@pytest.mark.parametrize("d", [Fraction(3, 10), Fraction(4, 5)])
def test_damping_two_sweep_trajectory_by_hand(d):
    """x has a prior a; h = AND(x) with p, q.  Worked by hand from the damping
    rule, in rational arithmetic, from init's (1, 1) messages:

    sweep 1:  F0->x = (1-d)(1-a, a) + d            (sum 1 + d)
              F1->x = (1-d)(1/2, 1/2) + d          (head message (1, 1): flat)
              F1->h = (1-d)((2-p-q)/2, (p+q)/2) + d
              v->f:  x->F1 = F0->x / (1 + d),  h->F1 = (1/2, 1/2)
    sweep 2:  F0->x = (1-d)(1-a, a) + d * F0->x(1)
              F1->x = (1-d)(1/2, 1/2) + d * F1->x(1)   (head message flat)
              F1->h = (1-d)((1-p)u1 + (1-q)u0, p u1 + q u0) + d * F1->h(1),
                      (u0, u1) = x->F1, u0 + u1 = 1
    beliefs:  b_x = F0->x[1] / sum (F1->x is flat), b_h = F1->h[1] / sum."""
    a, p, q = Fraction(7, 10), Fraction(19, 20), Fraction(1, 10)
    flist = fl.FactorList.build([(0, 1, [], float(a)), (1, 2, [1], float(p), float(q))])
    o = oracle.OracleGraph(flist)
    o.init()
    f0x = ((1 - d) * (1 - a) + d, (1 - d) * a + d)
    f1x = ((1 - d) / 2 + d, (1 - d) / 2 + d)
    f1h = ((1 - d) * (2 - p - q) / 2 + d, (1 - d) * (p + q) / 2 + d)
    o.sweep(float(d))
    b1 = o.beliefs()
    assert abs(b1[0] - float(f0x[1] / (f0x[0] + f0x[1]))) < 1e-15
    assert abs(b1[1] - float(f1h[1] / (f1h[0] + f1h[1]))) < 1e-15
    s = f0x[0] + f0x[1]
    u0, u1 = f0x[0] / s, f0x[1] / s
    f0x2 = ((1 - d) * (1 - a) + d * f0x[0], (1 - d) * a + d * f0x[1])
    f1x2 = ((1 - d) / 2 + d * f1x[0], (1 - d) / 2 + d * f1x[1])
    new_h = ((1 - p) * u1 + (1 - q) * u0, p * u1 + q * u0)
    f1h2 = ((1 - d) * new_h[0] + d * f1h[0], (1 - d) * new_h[1] + d * f1h[1])
    o.sweep(float(d))
    b2 = o.beliefs()
    fv2, _ = o.messages()
    assert abs(b2[0] - float(f0x2[1] / (f0x2[0] + f0x2[1]))) < 1e-15
    assert abs(b2[1] - float(f1h2[1] / (f1h2[0] + f1h2[1]))) < 1e-15
    # var-major order: x's edges (F0, F1), then h's (F1)
    want = np.array([[float(f0x2[0]), float(f0x2[1])], [float(f1x2[0]), float(f1x2[1])],
                     [float(f1h2[0]), float(f1h2[1])]])
    np.testing.assert_allclose(fv2, want, rtol=1e-15, atol=1e-16)
    # the same trajectory with the weights swapped, or damping applied to the
    # unnormalized F1->h of sweep 1, lands elsewhere
    sw = ((d) * (2 - p - q) / 2 + (1 - d), d * (p + q) / 2 + (1 - d))
    assert abs(float(sw[1] / (sw[0] + sw[1])) - b1[1]) > 1e-3
    un = ((1 - d) * (2 - p - q) + d, (1 - d) * (p + q) + d)
    assert abs(float(un[1] / (un[0] + un[1])) - b1[1]) > 1e-3
