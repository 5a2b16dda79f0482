"""The surrogate's kernels on the GPU against numpy and the CPU restatement
of the network (oracle/gnn_oracle.py, pinned by finite differences)."""
import numpy as np
import pytest

from oracle import gnn_oracle as G
from sdem_amd import gnn, native, scenarios

pytestmark = pytest.mark.gpu


def bits(x):
    """bf16 bit patterns of float values that bf16 holds exactly"""
    return (np.asarray(x, np.float32).view(np.uint32) >> 16).astype(np.uint16)


def test_mfma_lane_layout():
    """v_mfma_f32_32x32x16_bf16: A (32x16), B (16x32) given lane by lane as the
    kernels load them; D read back with the layout the GEMM epilogue uses."""
    print("mfma probe", flush=True)
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    B = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    a = np.zeros((64, 8), np.float32)
    b = np.zeros((64, 8), np.float32)
    for l in range(64):
        for j in range(8):
            a[l, j] = A[l % 32, 8 * (l // 32) + j]
            b[l, j] = B[8 * (l // 32) + j, l % 32]
    d = gnn.mfma_probe(bits(a), bits(b))
    D = A @ B
    for l in range(64):
        for r in range(16):
            assert d[l, r] == D[8 * (r // 4) + 4 * (l // 32) + r % 4, l % 32], (l, r)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("shape", [(70, 45, 37), (128, 128, 384), (5, 200, 16)])
def test_gemm_against_numpy(ta, tb, shape):
    M, N, K = shape
    print(f"gemm ta={ta} tb={tb} {shape}", flush=True)
    rng = np.random.default_rng(M + N + K)
    A = G.bf16(rng.normal(size=(M, K)))
    B = G.bf16(rng.normal(size=(K, N)))
    bias = rng.normal(size=N).astype(np.float32)
    a_st = A.T.copy() if ta else A
    b_st = B.T.copy() if tb else B
    c, pre = gnn.gemm(a_st, b_st, ta=ta, tb=tb, bias=bias, flags=1 | 2 | 4)
    ref_pre = A @ B + bias
    # bf16 inputs are exact; only the fp32 accumulation order differs
    tol = 1e-5 * np.sqrt(K) * (np.abs(A) @ np.abs(B) + 1)
    assert np.all(np.abs(pre - ref_pre) <= tol)
    assert np.allclose(c, G.silu(ref_pre), rtol=1e-5, atol=1e-4 * np.sqrt(K))


def test_gemm_split_k():
    print("gemm split-K", flush=True)
    rng = np.random.default_rng(4)
    A = G.bf16(rng.normal(size=(300, 50000)))  # stored as A^T: 50000 x 300 below
    B = G.bf16(rng.normal(size=(50000, 64)))
    c, _ = gnn.gemm(A.T.copy(), B, ta=True, splits=12)
    ref = A @ B
    assert np.allclose(c, ref, rtol=1e-4, atol=2e-3 * np.sqrt(50000))


def random_graph(n=60, seed=2):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3))
    src, dst = [], []
    for i in range(n):
        d = np.linalg.norm(x - x[i], axis=1)
        for j in np.nonzero((d < 0.3) & (d > 0))[0]:
            src.append(j)
            dst.append(i)
    return (rng.normal(size=(n, 8)).astype(np.float32), np.array(src, np.int32), np.array(dst, np.int32),
            rng.normal(size=(len(src), 7)).astype(np.float32))


def model_and_oracle(hidden=32, layers=2, seed=3):
    cfg = gnn.default_config(hidden=hidden, layers=layers)
    m = gnn.Model(cfg, seed=seed)
    offs, dims = m.layout()
    ref = G.Net(m.params().astype(np.float64), offs, dims, layers, 6, bf16_inputs=True)
    return m, ref


def test_forward_matches_oracle():
    print("gnn forward vs oracle", flush=True)
    m, ref = model_and_oracle()
    fn, src, dst, fe = random_graph()
    m.set_graph(fn, src, dst, fe)
    out = m.forward()
    r = ref.forward(G.bf16(fn), src, dst, G.bf16(fe))
    scale = np.abs(r).max()
    assert np.abs(out - r).max() <= 1e-2 * scale, (np.abs(out - r).max(), scale)


def test_gradients_match_oracle():
    """Gradients of sum(out * dout) from the hand-written backward kernels
    against the CPU restatement with the same bf16 rounding points, MLP by
    MLP (relative L2 error and direction)."""
    print("gnn gradients vs oracle", flush=True)
    m, ref = model_and_oracle()
    fn, src, dst, fe = random_graph()
    m.set_graph(fn, src, dst, fe)
    m.forward()
    rng = np.random.default_rng(7)
    dout = rng.normal(size=(fn.shape[0], 6)).astype(np.float32)
    m.backward(dout)
    g = m.grads().astype(np.float64)
    ref.forward(G.bf16(fn), src, dst, G.bf16(fe))
    gr = ref.backward(dout.astype(np.float64))
    offs, dims = m.layout()
    for k in range(len(offs)):
        lo, hi = offs[k][0], offs[k][5] + dims[k][2]
        a, b = g[lo:hi], gr[lo:hi]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        assert rel < 2e-2 and cos > 0.9998, (k, rel, cos)


def test_graph_from_dem_matches_brute_force(hip_lib):
    """Neighbour graph and features built on the GPU from a DEM state."""
    print("gnn graph from a DEM state", flush=True)
    sc = scenarios.settling(n=20000, seed=5)
    eng = native.build_engine(hip_lib, sc)
    eng.iterate(3)
    st = eng.get_state()
    cfg = gnn.default_config(hidden=32, layers=1, radius=4.8e-3, cutoff_factor=1.2, len_scale=2e-3,
                             vel_scale=0.1, floor_z=0.0)
    m = gnn.Model(cfg)
    m.graph_from_dem(eng)
    fn, src, dst, fe = m.get_graph()
    # every pair with |dx| <= 1.2 * 2 max(r_i, r_j) (nearest periodic image),
    # both directions, by particle id
    from scipy.spatial import cKDTree
    p = st["posr"][:, :3] - np.array(sc["options"]["domain_min"])
    r = st["posr"][:, 3]
    L = np.array(sc["options"]["domain_max"]) - np.array(sc["options"]["domain_min"])
    per = np.array(sc["options"]["is_periodic"], bool)
    box = np.where(per, L, 1e6)
    cand = cKDTree(np.mod(p, box), boxsize=box).query_pairs(4.8e-3, output_type="ndarray")
    d = p[cand[:, 0]] - p[cand[:, 1]]
    d[:, per] -= L[per] * np.round(d[:, per] / L[per])
    keep = np.sqrt((d * d).sum(1)) <= 1.2 * 2 * np.maximum(r[cand[:, 0]], r[cand[:, 1]])
    ids = st["id"]
    want = {(int(ids[a]), int(ids[b])) for a, b in cand[keep]}
    want |= {(b, a) for a, b in want}
    node_id = m.node_ids()
    got = {(int(node_id[s_]), int(node_id[d_])) for s_, d_ in zip(src, dst)}
    assert got == want
    assert np.all(np.diff(dst) >= 0)
    # edge features: |dx| / len_scale consistent with dx; velocities by id
    assert np.allclose(np.sqrt((fe[:, :3].astype(np.float64) ** 2).sum(1)), fe[:, 3], rtol=2e-2, atol=1e-3)
    order = np.argsort(ids)
    slot = order[np.searchsorted(ids[order], node_id)]
    assert np.allclose(fn[:, :3], st["lin_vel"][slot, 0, :] / 0.1, rtol=1e-2, atol=1e-3)
    assert np.allclose(fn[:, 6], st["posr"][slot, 3] / 2e-3, rtol=1e-2)


def test_training_reduces_the_loss():
    """Hand-written training on the GPU fits a smooth per-node target."""
    print("gnn training", flush=True)
    cfg = gnn.default_config(hidden=32, layers=2)
    m = gnn.Model(cfg, seed=11)
    fn, src, dst, fe = random_graph(n=200, seed=9)
    target = np.tanh(fn[:, :6] * 0.5).astype(np.float32)
    m.set_graph(fn, src, dst, fe)
    losses = [m.train_step(target, lr=3e-3) for _ in range(150)]
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])


def test_trained_surrogate_beats_ballistic(hip_lib):
    """Online training on DEM snapshots (20k-sphere settling boxes), then
    on snapshots of an unseen box the network's corrections bring the
    prediction of the DEM's next 10 steps closer than the ballistic step
    alone (the method without the network)."""
    print("surrogate: train 120 snapshots, evaluate on an unseen box", flush=True)
    from sdem_amd import surrogate
    m = surrogate.make_model(hidden=32, layers=2, seed=5, lib=hip_lib)
    losses = surrogate.train(m, n=20000, seeds=(11, 12), warm=400, snapshots=120, lr=2e-3, lr_end=5e-4,
                             spacing=10, log=None, lib=hip_lib)
    # losses are relative to the no-network loss of the same snapshot
    assert np.mean(losses[-20:]) < 0.8 * np.mean(losses[:5]), (losses[:5], losses[-20:])
    sc = scenarios.settling(n=20000, seed=99)
    eng = native.build_engine(hip_lib, sc)
    eng.iterate(500)
    o = sc["options"]
    L = np.array(o["domain_max"]) - np.array(o["domain_min"])
    err_net, err_bal = [], []
    for _ in range(5):
        m.graph_from_dem(eng, for_training=False)
        ids = m.node_ids()
        s0 = eng.get_state()
        pred = m.forward()
        eng.iterate(m.cfg.nsub)
        s1 = eng.get_state()
        dx, dv = surrogate.residuals(s0, s1, ids, o["dt"] * m.cfg.nsub, o["gravity"], L, o["is_periodic"])
        t = np.concatenate([dx / m.cfg.dx_scale, dv / m.cfg.dv_scale], 1)
        err_net.append(np.mean((pred - t) ** 2))
        err_bal.append(np.mean(t ** 2))
        eng.iterate(10)
    assert np.mean(err_net) < 0.9 * np.mean(err_bal), (err_net, err_bal)


def test_surrogate_advances_a_dem_context(hip_lib):
    """Surrogate calls move the DEM state on the device; the DEM then goes
    on stepping from it (iterNumber advanced, history cleared)."""
    print("surrogate: advance a DEM context", flush=True)
    from sdem_amd import surrogate
    m = surrogate.make_model(hidden=32, layers=2, seed=6, lib=hip_lib)
    sc = scenarios.settling(n=20000, seed=3)
    eng = native.build_engine(hip_lib, sc)
    eng.iterate(20)
    it0 = eng.get_iter()[0]
    z0 = eng.get_state()["posr"][:, 2].copy()
    m.advance(eng, 3)
    assert eng.get_iter()[0] == it0 + 30
    z1 = eng.get_state()["posr"][:, 2]
    assert np.all(np.isfinite(z1)) and np.mean(z1) < np.mean(z0)  # falling
    eng.iterate(20)
    assert eng.get_iter()[0] == it0 + 50
    assert np.all(np.isfinite(eng.get_state()["posr"]))


def test_surrogate_keeps_rotvel2_of_the_last_dem_step(hip_lib):
    """With AB2 rotation the force kernel leaves rotVel(2) out of the state
    it writes (it equals that state's linVel(1)).  The surrogate replaces
    linVel(1) and leaves the rotation alone, so before its first call the
    library stores rotVel(2): afterwards it is still the linVel(1) of the last
    DEM step, as if every step had stored it."""
    print("surrogate: rotVel(2) across surrogate calls", flush=True)
    from sdem_amd import surrogate
    m = surrogate.make_model(hidden=32, layers=2, seed=9, lib=hip_lib)
    sc = scenarios.settling(n=20000, seed=5)
    eng = native.build_engine(hip_lib, sc)
    eng.iterate(7)
    s0 = eng.get_state()
    assert s0["rot_vel"].shape[1] == 2
    assert np.array_equal(s0["rot_vel"][:, 1], s0["lin_vel"][:, 0])  # AB2 quirk, read back
    m.advance(eng, 2)
    s1 = eng.get_state()
    v0 = dict(zip(s0["id"].tolist(), s0["lin_vel"][:, 0].tolist()))
    assert not np.array_equal(s1["lin_vel"][:, 0], s0["lin_vel"][:, 0])  # the surrogate moved it
    assert [v0[i] for i in s1["id"].tolist()] == s1["rot_vel"][:, 1].tolist()
    eng.iterate(3)
    s2 = eng.get_state()
    assert np.array_equal(s2["rot_vel"][:, 1], s2["lin_vel"][:, 0])


def test_fused_inference_matches_training_path(hip_lib):
    """Inference runs each message-passing MLP as one fused kernel (input
    gather, three matrix-core products, activations on chip); it gives the
    bits of the training path's separate GEMMs on the same graph."""
    print("surrogate: fused inference vs the three-GEMM path", flush=True)
    from sdem_amd import surrogate
    sc = scenarios.settling(n=20000, seed=8)
    eng = native.build_engine(hip_lib, sc)
    eng.iterate(5)
    m = surrogate.make_model(seed=4, lib=hip_lib)  # 8 layers, 128 hidden units
    rng = np.random.default_rng(2)
    m.set_params(m.params() * (1.0 + 0.5 * rng.standard_normal(m.num_params())).astype(np.float32))
    m.graph_from_dem(eng, for_training=True)
    ids_t, out_t = m.node_ids(), m.forward()
    m.graph_from_dem(eng, for_training=False)
    ids_i, out_i = m.node_ids(), m.forward()
    assert np.array_equal(ids_t, ids_i)
    assert np.abs(out_t).max() > 0
    assert np.array_equal(out_t, out_i)
