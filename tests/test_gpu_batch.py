"""Batched form: one hipGraph replay covers many uneven C_i = A_i * B_i. Each member must
equal the oracle (and the per-matrix plain pipeline); replays interleaved with other device
work must keep giving the same result."""
import numpy as np
import pytest

import dbd_amd
from dbd_amd import batch
from helpers import compare, reference

pytestmark = pytest.mark.gpu


def make_batch(n_mats, seed, rows_lo=200, rows_hi=6000, dtype=np.float32):
    rng = np.random.default_rng(seed)
    As, Bs = [], []
    for i in range(n_mats):
        n = int(np.exp(rng.uniform(np.log(rows_lo), np.log(rows_hi))))
        m = float(rng.uniform(4, 64))
        As.append(dbd_amd.gen_skewed(n, m, 1.8, seed=1000 * seed + 2 * i, dtype=dtype))
        Bs.append(dbd_amd.gen_skewed(n, m, 1.8, seed=1000 * seed + 2 * i + 1, dtype=dtype))
    return As, Bs


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_graph_matches_oracle(gpu_ok, dtype):
    import torch
    As, Bs = make_batch(24, 2, dtype=dtype)
    A_cat, B_cat, lay = batch.concat(As, Bs)
    runner = batch.BatchRunner(A_cat, B_cat, block_rows=512)
    first = None
    for rep in range(3):
        runner.launch()
        # unrelated eager device work and a synchronisation between replays
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()
        torch.cuda.synchronize()
        Cp, Ci, Cv = runner.result()
        if first is None:
            first = (Cp, Ci)
        else:
            assert np.array_equal(first[0], Cp) and np.array_equal(first[1], Ci), f"replay {rep} structure"
        for k, (A, B, Cm) in enumerate(zip(As, Bs, batch.split(Cp, Ci, Cv, lay))):
            compare(Cm, reference(A, B), dtype, f"replay {rep} member {k}")
    runner.close()


def test_batch_equals_plain_pipeline(gpu_ok):
    As, Bs = make_batch(10, 3)
    A_cat, B_cat, lay = batch.concat(As, Bs)
    runner = batch.BatchRunner(A_cat, B_cat, block_rows=512, tune=True)
    runner.launch()
    Cp, Ci, Cv = runner.result()
    ctx, run = batch.plain_each(As, Bs, block_rows=512, keep=True)
    outs = run()
    for (p, i, v), Cm in zip(outs, batch.split(Cp, Ci, Cv, lay)):
        assert np.array_equal(p.cpu().numpy(), Cm.ptr)
        n = int(Cm.ptr[-1])
        assert np.array_equal(i[:n].cpu().numpy(), Cm.idx)
        np.testing.assert_allclose(v[:n].cpu().numpy(), Cm.val, rtol=1e-5, atol=1e-6)
    ctx.close()
    runner.close()
