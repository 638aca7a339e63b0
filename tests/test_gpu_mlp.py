"""The learned limiter's network on the GPU (BF16 MFMA kernels) against the
numpy oracle: forward pass, loss and gradients, the stencil features of a
mesh state, training, and the coefficients a context computes each step.

Tolerances: the kernels round the same quantities to BF16 as the oracle, but
they accumulate in a different order, so a hidden activation can round to the
neighbouring BF16 value; outputs and gradients are compared in relative L2
norm (1e-2 per tensor), features (pure FP32 arithmetic) to 1e-6."""
import os
import sys

import numpy as np
import pytest

import helpers as H  # noqa: E402
from mahyco_amd import abi, load_gpu

sys.path.insert(0, os.path.join(H.ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu
LEARNED = {"projection_limiteur_id": abi.LIM_LEARNED}


@pytest.fixture(scope="module")
def gpu():
    return load_gpu()


def init_params(seed):
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2 / 125), (64, 128)), "b1": rng.normal(0, 0.05, 64),
         "W2": rng.normal(0, np.sqrt(2 / 64), (64, 64)), "b2": rng.normal(0, 0.05, 64),
         "W3": rng.normal(0, np.sqrt(1 / 64), (2, 64)), "b3": np.array([0.8, 0.8])}
    return M.pack(d).astype(np.float32)


def data(seed, m):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (m, 128)).astype(np.float32)
    X[:, 125:] = 0
    T = rng.normal(0, 1, (m, 16)).astype(np.float32)
    return X, T


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle(gpu):
    p = init_params(1)
    net = abi.Mlp(gpu, p)
    for m in (1, 31, 32, 33, 1000):
        X, _ = data(m, m)
        y = net.forward(X)
        yo, _ = M.forward(p, X)
        assert rel_l2(y, yo) < 1e-2, (m, rel_l2(y, yo))
    net.close()


def test_gradient_matches_oracle(gpu):
    p = init_params(2)
    net = abi.Mlp(gpu, p)
    X, T = data(5, 777)
    net.set_data(X, T)
    for row0, m in [(0, 777), (100, 64), (3, 5)]:
        loss, g = net.gradient(row0, m)
        lo, go = M.loss_grad(p, X[row0:row0 + m], T[row0:row0 + m])
        assert abs(loss - lo) <= 1e-3 * abs(lo), (loss, lo)
        gu, gou = M.unpack(g), M.unpack(go)
        for k in gu:
            print(row0, m, k, f"{rel_l2(gu[k], gou[k]):.2e}")
            assert rel_l2(gu[k], gou[k]) < 1e-2, k
        # and against the exact (float64) gradient, within the BF16 budget
        _, g64 = M.loss_grad(p.astype(np.float64), X[row0:row0 + m], T[row0:row0 + m], bf=False)
        assert rel_l2(g, g64) < 5e-2
    net.close()


def test_training_fits_a_teacher(gpu):
    """Targets made by a teacher network (st = alpha_teacher sc): Adam on the
    GPU drives the loss down by 10x or more, deterministically."""
    teacher = init_params(7)
    X, T = data(8, 8192)
    at, _ = M.forward(teacher, X)
    T = T.reshape(-1, 2, 4, 2)
    T[..., 1] = at[:, :, None] * T[..., 0]
    T = T.reshape(-1, 16)
    losses = []
    for _ in range(2):
        net = abi.Mlp(gpu, init_params(9))
        net.set_data(X, T)
        losses.append(net.train(epochs=60, batch=512, lr=2e-3, seed=4))
        params = net.params()
        net.close()
    print("epoch losses", losses[0][[0, 5, 10, 30, 59]])
    assert np.array_equal(losses[0], losses[1])
    assert losses[0][-1] < 0.1 * losses[0][0]
    assert np.all(np.isfinite(params))


def test_features_match_oracle(gpu):
    s = H.tp_setup(gpu, 56, 24)
    s.run(40)
    X, S = s.learned_features()
    st = s.download(["density", "internal_energy", "velocity", "env_mask", "fracvol_env"])
    multi = np.array([bin(int(v)).count("1") > 1 for v in st["env_mask"]])
    vf = np.where(multi, np.max(np.where((st["env_mask"][None, :] >> np.arange(3)[:, None]) & 1,
                                         st["fracvol_env"], 0.0), axis=0), 1.0)
    Xo, So = M.features(56, 24, st["density"], st["internal_energy"], st["velocity"][0], st["velocity"][1], vf)
    print("max |dX|", np.max(np.abs(X - Xo)), "max rel dS", np.max(np.abs(S - So) / So))
    assert np.max(np.abs(X - Xo)) <= 1e-6
    assert np.max(np.abs(S - So) / So) <= 1e-6


def test_context_inference_and_step_match_oracle(gpu):
    """A learned triple point with the network attached: the coefficients it
    computes match the oracle's network on the oracle's features, and one
    step with them matches the oracle step given the same coefficients."""
    p = init_params(11)
    net = abi.Mlp(gpu, p)
    g = H.tp_setup(gpu, 56, 24, **LEARNED)
    o = H.tp_setup(H.oracle(), 56, 24, **LEARNED)
    o.set_learned_alpha(np.ones((2, o.nc)))
    o.run(20)
    g.upload(o.download())
    g.set_limiter_mlp(net)
    a = g.learned_alpha()
    X, _ = g.learned_features()
    yo, _ = M.forward(p, X)
    mask = g.download(["env_mask"])["env_mask"].reshape(24, 56)
    pad = np.pad(mask, 1, mode="constant", constant_values=-1)
    uni = np.ones((24, 56), dtype=bool)
    for dj in range(3):
        for di in range(3):
            nb = pad[dj:dj + 24, di:di + 56]
            uni &= (nb == mask) | (nb == -1)
    uni &= np.vectorize(lambda v: bin(int(v)).count("1") == 1)(mask)
    ao = np.where(uni.ravel()[None, :], np.clip(yo, 0, 2).T, -1.0)
    assert np.any(ao < 0) and np.any(ao >= 0)
    print("alpha range", a.min(), a.max(), "mean |d|", np.mean(np.abs(a - ao)))
    assert np.mean(np.abs(a - ao)) < 1e-2
    # a fixed coefficient at material interfaces instead of the classical limiter
    g.set_learned_interface_alpha(0.7)
    a07 = g.learned_alpha()
    assert np.all(a07[ao < 0] == 0.7) and np.array_equal(a07[ao >= 0], a[ao >= 0])
    with pytest.raises(abi.MahycoError):
        g.set_learned_interface_alpha(2.5)
    g.set_learned_interface_alpha(-1)
    assert np.array_equal(g.learned_alpha(), a)
    o.set_learned_alpha(a)
    o.step()
    g.step()
    d = H.compare_states(g.download(), o.download(), 3, 56, 24)
    assert max(d.values()) < 1e-12, d
    net.close()


def test_trained_network_gains_over_plain(gpu):
    """The committed network (mahyco_amd/data/limiter_mlp.npz) on a 448x192
    triple point to t = 0.3, against a plain 1792x768 run: applied everywhere
    it lowers the density L1 error of the plain method at the same resolution
    by 10 % or more; with material interfaces on minmod it is no worse."""
    from mahyco_amd import learned
    p = np.load(os.path.join(H.ROOT, "arcane-benchs_amd", "mahyco_amd", "data", "limiter_mlp.npz"))["params"]
    ev = learned.evaluate_tp(gpu, p, 0.3, base=(448, 192), factors=(1.0, 1.5))
    plain = ev["plain"]["448x192"]["l1"]
    everywhere = ev["learned"]["everywhere"]["l1"]
    safe = ev["learned"]["interfaces_minmod"]["l1"]
    print("plain", plain, "learned everywhere", everywhere, "interfaces minmod", safe)
    assert everywhere < 0.9 * plain
    assert safe <= 1.01 * plain
