"""Pin the CPU restatement of the surrogate's network (oracle/gnn_oracle.py):
its backward pass against central finite differences of its own forward
pass (smooth, without bf16 rounding), and its bf16 rounding against the
definition (nearest, ties to even)."""
import numpy as np

from oracle import gnn_oracle as G


def small_net(seed=0, hidden=8, layers=2, out=6):
    offs, dims, n = G.layout(hidden, layers)
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.4, n)
    return G.Net(p, offs, dims, layers, out, bf16_inputs=False)


def small_graph(seed=1, n=7):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3))
    src, dst = [], []
    for i in range(n):
        for j in range(n):
            if i != j and np.linalg.norm(x[i] - x[j]) < 0.7:
                src.append(j)
                dst.append(i)
    src, dst = np.array(src), np.array(dst)
    order = np.argsort(dst, kind="stable")
    return rng.normal(size=(n, 8)), src[order], dst[order], rng.normal(size=(len(src), 7))


def test_backward_matches_finite_differences():
    net = small_net()
    fn, src, dst, fe = small_graph()
    assert len(src) > 4
    rng = np.random.default_rng(3)
    dout = rng.normal(size=(fn.shape[0], 6))
    net.forward(fn, src, dst, fe)
    g = net.backward(dout)
    idx = rng.choice(len(net.p), 60, replace=False)
    eps = 1e-6
    for k in idx:
        p0 = net.p[k]
        net.p[k] = p0 + eps
        fp = np.sum(net.forward(fn, src, dst, fe) * dout)
        net.p[k] = p0 - eps
        fm = np.sum(net.forward(fn, src, dst, fe) * dout)
        net.p[k] = p0
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * (1.0 + abs(fd)), (k, fd, g[k])


def test_message_passing_reaches_neighbours_only():
    """A node's output depends on nodes within `layers` hops only."""
    net = small_net(layers=1)
    fn, src, dst, fe = small_graph(n=9)
    base = net.forward(fn, src, dst, fe)
    fn2 = fn.copy()
    fn2[0] += 1.0
    moved = np.abs(net.forward(fn2, src, dst, fe) - base).max(1) > 0
    reach = {0} | set(dst[src == 0].tolist())
    assert set(np.nonzero(moved)[0].tolist()) <= reach


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-5, 65504.0])
    r = G.bf16(x)
    # 8 significant bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-5) / 3.0e-5 < 2 ** -8
    assert r[5] == 65536.0
