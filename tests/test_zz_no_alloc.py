"""bm_batch_gen_learned_noalloc (fixed device scratch, no allocation) gives
the same ranked pairs as bm_batch_gen_learned, on two calls in a row, the
second with a smaller pool and top, and refuses sizes beyond its limits."""
import numpy as np
import pytest

from conftest import make_pair

pytestmark = pytest.mark.gpu


def test_gen_learned_noalloc_matches(bm, orc):
    n = 2000
    male, female, truth, pi0 = make_pair(bm, n=n, deg=20.0, seed=3)
    print("[no_alloc] n=2000 pair", flush=True)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        with bm.Mlp(ctx, seed=5) as m:
            m.refresh_features()
            m.train(seed=6, steps=30, batch=1 << 12, lr=2e-3)
            ref = ctx.batch(8192)
            got = ctx.batch(8192)
            for seed, pool, top in ((41, 1 << 20, 4096), (42, 1 << 14, 777)):
                print(f"[no_alloc] pool={pool} top={top}", flush=True)
                ref.gen_learned(m, seed, pool, top)
                got.gen_learned_noalloc(m, seed, pool, top)
                assert got.size == top
                assert np.array_equal(got.pairs(), ref.pairs())
                uv = orc.uniform_pairs(seed, pool, n)
                assert set(map(tuple, got.pairs().tolist())) <= set(map(tuple, uv.tolist()))
            with pytest.raises(bm.BMError):
                got.gen_learned_noalloc(m, 1, (1 << 20) + 1, 16)
            with pytest.raises(bm.BMError):
                got.gen_learned_noalloc(m, 1, 1 << 16, 8193)
            ref.close()
            got.close()
