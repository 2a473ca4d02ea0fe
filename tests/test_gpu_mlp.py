"""Learned proposal on the device: per-vertex features against a plain
restatement, the MFMA forward and backward against a torch float32
reference (with the same bfloat16 rounding points), training, ranking, and
what the trained network gains over uniform proposals."""
import numpy as np
import pytest
import torch

from conftest import make_pair, quirky_pair

pytestmark = pytest.mark.gpu

# Forward tolerance: MFMA and torch sum in different orders, which can move a
# hidden activation across a bfloat16 rounding boundary (one bf16 ulp = 2^-8).
FWD_ATOL, FWD_RTOL = 2e-2, 2e-2
# Gradient tolerance, relative to the largest entry of each parameter block.
GRAD_REL = 2e-2


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def unpack(p):
    o, out = 0, []
    for shape in ((16, 128), (128,), (128, 128), (128,), (128,), (1,)):
        size = int(np.prod(shape))
        out.append(torch.tensor(p[o:o + size].reshape(shape), dtype=torch.float32))
        o += size
    return out


def ref_forward(X, p):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = bf(X) @ bf(W1) + b1
    h1 = bf(torch.relu(z1))
    z2 = h1 @ bf(W2) + b2
    h2 = torch.relu(z2)
    y = h2 @ w3 + b3
    return y, (X, z1, h1, z2, h2, W2, w3)


def ref_gradient(X, t, p):
    """The backward the device kernels implement, restated in float32."""
    y, (X, z1, h1, z2, h2, W2, w3) = ref_forward(X, p)
    k = X.shape[0]
    dy = 2.0 * (y - t) / k
    dw3 = h2.T @ dy
    db3 = dy.sum().reshape(1)
    g2f = dy[:, None] * w3[None, :] * (z2 > 0)
    g2 = bf(g2f)
    dW2 = h1.T @ g2
    db2 = g2f.sum(0)
    g1f = (g2 @ bf(W2).T) * (z1 > 0)
    g1 = bf(g1f)
    dW1 = bf(X).T @ g1
    db1 = g1f.sum(0)
    return [dW1, db1, dW2, db2, dw3, db3], float(((y - t) ** 2).mean())


def pair_inputs(vf, uv):
    return torch.tensor(np.concatenate([vf[uv[:, 0]], vf[uv[:, 1]]], axis=1), dtype=torch.float32)


def features_reference(n, male, female, pi):
    """The 8 per-vertex features from their definition (float64)."""
    fw = {}
    for s, d, w in zip(female.src.tolist(), female.dst.tolist(), female.w.tolist()):
        fw[(s, d)] = w
    mw = {}
    for s, d, w in zip(male.src.tolist(), male.dst.tolist(), male.w.tolist()):
        mw[(s, d)] = w      # last line wins
    fout = np.zeros(n + 1)
    fin = np.zeros(n + 1)
    for (s, d) in fw:
        fout[s] += 1
        fin[d] += 1
    dout = np.zeros(n + 1)
    din = np.zeros(n + 1)
    ws = np.zeros(n + 1)
    L = np.zeros(n + 1)
    miss = np.zeros(n + 1)
    for (s, d), w in mw.items():
        f = fw.get((pi[s], pi[d])) if pi[s] and pi[d] else None
        c = 0 if f is None else min(w, f)
        dout[s] += 1
        din[d] += 1
        for x in ((s,) if s == d else (s, d)):
            ws[x] += w
            L[x] += c
            miss[x] += c == 0
    out = np.zeros((n + 1, 8))
    for x in range(1, n + 1):
        fx = pi[x]
        wpos = max(ws[x], 0.0)
        lpos = max(L[x], 0.0)
        out[x] = [np.log1p(dout[x]), np.log1p(din[x]), np.log1p(wpos), lpos / (wpos + 1.0),
                  miss[x] / (dout[x] + din[x] + 1.0), np.log1p(fout[fx]) if fx else 0.0,
                  np.log1p(fin[fx]) if fx else 0.0, np.log1p(lpos)]
    return out


@pytest.fixture(scope="module")
def c1(bm):
    male, female, truth, pi0 = make_pair(bm, n=2000, deg=20.0, seed=3)
    ctx = bm.Context(2000, male, female)
    ctx.set_mapping(pi0)
    yield dict(ctx=ctx, male=male, female=female, pi0=pi0, n=2000)
    ctx.close()


def test_vertex_features_match_definition(bm):
    for male, female, pi, n in [quirky_pair(bm, seed=7) + (60,), make_pair(bm, n=300, deg=8.0, seed=2)[:2] + (None, 300)]:
        with bm.Context(n, male, female) as ctx:
            if pi is None:
                pi = np.concatenate([[0], np.random.default_rng(1).permutation(np.arange(1, n + 1))]).astype(np.int32)
            ctx.set_mapping(pi)
            with bm.Mlp(ctx, seed=1) as m:
                m.refresh_features()
                got = m.vertex_features()
                ref = features_reference(n, male, female, pi)
                assert np.allclose(got, ref, rtol=1e-5, atol=1e-5)


def test_forward_matches_reference(bm, c1):
    ctx = c1["ctx"]
    with bm.Mlp(ctx, seed=5) as m:
        p = m.params()
        rng = np.random.default_rng(0)
        p[16 * 128:16 * 128 + 128] = rng.normal(0, 0.1, 128)          # nonzero biases
        p[-129 - 128:-129] = rng.normal(0, 0.1, 128)
        p[-1] = 0.3
        m.set_params(p)
        m.refresh_features()
        vf = m.vertex_features()
        uv = rng.integers(1, 2001, (5000, 2)).astype(np.int32)
        uv = uv[uv[:, 0] != uv[:, 1]]
        got = m.score_pairs(uv)
        ref, _ = ref_forward(pair_inputs(vf, uv), p)
        ref = ref.numpy()
        assert np.allclose(got, ref, atol=FWD_ATOL, rtol=FWD_RTOL)
        assert np.median(np.abs(got - ref)) < 1e-3
        # ragged count (not a multiple of the 16-pair tile)
        assert np.allclose(m.score_pairs(uv[:37]), got[:37], atol=1e-6)


def test_gradient_matches_reference(bm, c1, orc):
    ctx = c1["ctx"]
    with bm.Mlp(ctx, seed=6) as m:
        m.refresh_features()
        vf = m.vertex_features()
        p = m.params()
        uv = orc.uniform_pairs(12, 3001, 2000)              # ragged batch
        d = ctx.swap_deltas(uv).astype(np.float64)
        t = (np.sign(d) * np.log1p(np.abs(d))).astype(np.float32)
        g, mse = m.gradient(uv, t)
        X = pair_inputs(vf, uv)
        ref, ref_mse = ref_gradient(X, torch.tensor(t), p)
        assert abs(mse - ref_mse) <= 1e-3 * max(1.0, ref_mse)
        o = 0
        for r in ref:
            block = g[o:o + r.numel()].reshape(r.shape)
            o += r.numel()
            scale = float(r.abs().max()) + 1e-12
            assert float((torch.tensor(block) - r).abs().max()) <= GRAD_REL * scale
        # the same gradient against autograd of the unrounded float32 model
        P = [x.clone().requires_grad_(True) for x in unpack(p)]
        W1, b1, W2, b2, w3, b3 = P
        y = torch.relu(torch.relu(X @ W1 + b1) @ W2 + b2) @ w3 + b3
        ((y - torch.tensor(t)) ** 2).mean().backward()
        o = 0
        for q in P:
            block = torch.tensor(g[o:o + q.numel()].reshape(q.shape))
            o += q.numel()
            cos = torch.nn.functional.cosine_similarity(block.flatten(), q.grad.flatten(), dim=0)
            assert float(cos) > 0.98


def test_training_learns_and_ranking_gains(bm, c1, orc):
    """Trained on the context's own exact deltas, the network's top-ranked
    pairs are far better than uniform pairs and the ranking kernel returns
    exactly the best-scored pairs of its pool."""
    ctx, n = c1["ctx"], c1["n"]
    ctx.set_mapping(c1["pi0"])
    with bm.Mlp(ctx, seed=7) as m:
        m.refresh_features()
        losses = m.train(seed=3, steps=300, batch=1 << 13, lr=2e-3)
        assert losses[-20:].mean() < 0.8 * losses[:20].mean()
        pool, top = 1 << 16, 1024
        b = ctx.batch(pool)
        b.gen_learned(m, 99, pool, top)
        got = b.pairs()
        uv = orc.uniform_pairs(99, pool, n)
        s = m.score_pairs(uv)
        order = np.argsort(-s, kind="stable")
        assert np.isclose(np.sort(s[order[:top]]), np.sort(m.score_pairs(got))).all()
        assert set(map(tuple, got.tolist())) <= set(map(tuple, uv.tolist()))
        d_learned = ctx.swap_deltas(got)
        d_uniform = ctx.swap_deltas(uv[:top])
        assert d_learned.mean() > d_uniform.mean() and (d_learned > 0).sum() > 2 * max(1, (d_uniform > 0).sum())
        # predicted vs exact on held-out pairs: positive rank correlation
        hold = orc.uniform_pairs(1234, 20000, n)
        pr = m.score_pairs(hold)
        ex = ctx.swap_deltas(hold)
        rp = np.argsort(np.argsort(pr))
        re = np.argsort(np.argsort(ex))
        assert np.corrcoef(rp, re)[0, 1] > 0.3
        b.close()


def test_climb_learned_beats_uniform_at_equal_rounds(bm, c1):
    ctx = c1["ctx"]
    with bm.Mlp(ctx, seed=8) as m:
        ctx.set_mapping(c1["pi0"])
        m.refresh_features()
        m.train(seed=4, steps=300, batch=1 << 13, lr=2e-3)
        ctx.set_mapping(c1["pi0"])
        tu, _, ru = ctx.climb(bm.PROPOSAL_UNIFORM, seed=5, max_rounds=40, per_round=1024, verify=True)
        ctx.set_mapping(c1["pi0"])
        tl, _, rl = ctx.climb(bm.PROPOSAL_LEARNED, seed=5, max_rounds=40, per_round=1024, pool=1 << 16, mlp=m,
                              refresh_every=1, verify=True)
        ctx.set_mapping(c1["pi0"])
        t2, _, r2 = ctx.climb(bm.PROPOSAL_TWOHOP, seed=5, max_rounds=40, per_round=1024, per_anchor=64, verify=True)
        for r in (ru, rl, r2):
            assert r["verify_failures"] == 0
        assert rl["final_score"] > ru["final_score"]
        assert tl[-1] == rl["final_score"] and tu[-1] == ru["final_score"] and t2[-1] == r2["final_score"]


def _rank_keys(s):
    """Order-preserving integer image of float32 scores (-0 below +0)."""
    u = np.ascontiguousarray(s, np.float32).view(np.uint32).astype(np.uint64)
    return np.where(u & 0x80000000, (~u) & 0xFFFFFFFF, u | 0x80000000)


@pytest.mark.parametrize("top", [1, 3, 1000, 4096, 8192, 8193])
def test_learned_ranking_order_exact(bm, c1, orc, top):
    """The ranked batch is exactly the stable descending order of the pool's
    scores (ties to the smaller pool index): the radix-select path (top up to
    8,192) and the full-sort path agree with a host sort of the same scores;
    with all-zero weights every score ties and the first pool pairs win."""
    ctx, n = c1["ctx"], c1["n"]
    ctx.set_mapping(c1["pi0"])
    pool = 1 << 15
    with bm.Mlp(ctx, seed=9) as m:
        m.refresh_features()
        m.train(seed=2, steps=20, batch=1 << 12, lr=2e-3)
        uv = orc.uniform_pairs(123, pool, n)
        s = m.score_pairs(uv)
        order = np.lexsort((np.arange(pool), -_rank_keys(s).astype(np.float64)))
        # lexsort on float64 of 32-bit keys is exact (< 2^53)
        b = ctx.batch(top)
        b.gen_learned(m, 123, pool, top)
        assert np.array_equal(b.pairs(), uv[order[:top]])
        m.set_params(np.zeros(m.N_PARAMS, np.float32))
        b.gen_learned(m, 123, pool, top)
        assert np.array_equal(b.pairs(), uv[:top])
        b.close()
