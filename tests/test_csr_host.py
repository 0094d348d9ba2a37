"""CPU check of csr_from_w, the drop-in block's conversion of a [4, N, N] weight
tensor to the presynaptic-major CSR: a random connection table laid out the
way generation/read_block.py:64-74 builds w_uij (index rows channel, post,
pre; each synapse in its channel pair 2 kind + {0, 1} with its two weights)
comes back synapse for synapse, from the sparse and the dense tensor."""
import numpy as np
import pytest
import torch

from digital_twin_brain_amd.block import csr_from_w


def _table(n, m, rng, equal_pairs):
    key = rng.choice(n * n * 2, size=m, replace=False)  # unique (pre, post, kind)
    pre, rest = key // (2 * n), key % (2 * n)
    post, kind = rest // 2, rest % 2
    w0 = rng.integers(1, 1000, size=m) / 64.0
    w1 = w0.copy() if equal_pairs else rng.integers(1, 1000, size=m) / 64.0
    return pre, post, kind, w0, w1


def _w_uij(n, pre, post, kind, w0, w1, sparse):
    # read_block.py: idx = [channel offset, output (post), input (pre)], duplicated with channel + 1
    ch = np.stack([2 * kind, 2 * kind + 1], axis=-1).reshape(-1)
    idx = np.stack([ch, np.repeat(post, 2), np.repeat(pre, 2)])
    val = np.stack([w0, w1], axis=-1).reshape(-1)
    w = torch.sparse_coo_tensor(torch.from_numpy(idx), torch.from_numpy(val).float(), (4, n, n))
    return w.coalesce() if sparse else w.to_dense()


@pytest.mark.parametrize("sparse", [True, False])
@pytest.mark.parametrize("equal_pairs", [True, False])
def test_csr_from_w_recovers_the_connection_table(sparse, equal_pairs):
    rng = np.random.default_rng(11)
    n, m = 300, 4000
    pre, post, kind, w0, w1 = _table(n, m, rng, equal_pairs)
    c = csr_from_w(_w_uij(n, pre, post, kind, w0, w1, sparse))
    assert c["n"] == n and c["rowptr"][0] == 0 and c["rowptr"][-1] == m
    assert np.all(np.diff(c["rowptr"]) >= 0)
    got_pre = np.repeat(np.arange(n), np.diff(c["rowptr"]))
    got = sorted(zip(got_pre.tolist(), c["post"].tolist(), c["kind"].tolist(), c["w0"].tolist(),
                     (c["w0"] if c["w1"] is None else c["w1"]).tolist()))
    want = sorted(zip(pre.tolist(), post.tolist(), kind.tolist(), w0.tolist(), w1.tolist()))
    assert got == want
    assert (c["w1"] is None) == equal_pairs
    # rows are presynaptic; inside a row, (post, kind) ascending
    for j in range(0, n, 37):
        s, e = c["rowptr"][j], c["rowptr"][j + 1]
        key = c["post"][s:e].astype(np.int64) * 2 + c["kind"][s:e]
        assert np.all(np.diff(key) > 0)


def test_csr_from_w_half_pairs_and_shape_checks():
    """A synapse present in one channel of its pair only gets weight 0 on the
    other; a tensor that is not [4, N, N] is refused."""
    n = 5
    w = torch.zeros(4, n, n)
    w[1, 2, 3] = 0.5   # NMDA only: pre 3 -> post 2, kind 0, weights (0, 0.5)
    w[2, 4, 0] = 0.25  # GABAa only: pre 0 -> post 4, kind 1, weights (0.25, 0)
    c = csr_from_w(w)
    assert list(np.diff(c["rowptr"])) == [1, 0, 0, 1, 0]
    assert c["post"].tolist() == [4, 2] and c["kind"].tolist() == [1, 0]
    assert c["w0"].tolist() == [0.25, 0.0] and c["w1"].tolist() == [0.0, 0.5]
    with pytest.raises(ValueError):
        csr_from_w(torch.zeros(2, n, n))
    with pytest.raises(ValueError):
        csr_from_w(torch.zeros(4, n, n + 1))
