"""Pins the oracle's FP64 radiance-cache model (oracle/mlp.c): its analytic
gradients against central finite differences of its own loss (which pins
the backward pass against the forward pass), the loss against a closed form
and the hash-grid level table against the instant-NGP definition."""
import numpy as np
import pytest


def small_cfg(orc, log2_table=10, levels=4):
    return orc.cache_cfg(levels, log2_table, 4, 2.0, (0, 0, 0), (1, 1, 1))


def random_problem(n, levels, table, seed=0):
    rng = np.random.default_rng(seed)
    mlp = np.zeros(17408, np.float32)
    for l in range(5):
        off = l * 4096
        nout = 16 if l == 4 else 64
        used = 3 if l == 4 else 64
        a = np.sqrt(6 / (64 + used))
        Wl = rng.uniform(-a, a, (64, used)).astype(np.float32)
        blk = np.zeros((64, nout), np.float32)
        blk[:, :used] = Wl
        mlp[off:off + 64 * nout] = blk.reshape(-1)
    grid = rng.uniform(-0.5, 0.5, levels * table * 2).astype(np.float32)
    inp = np.zeros((n, 12), np.float32)
    inp[:, :3] = rng.uniform(0.02, 0.98, (n, 3))
    nr = rng.normal(size=(n, 3))
    inp[:, 3:6] = nr / np.linalg.norm(nr, axis=1, keepdims=True)
    vd = rng.normal(size=(n, 3))
    inp[:, 6:9] = vd / np.linalg.norm(vd, axis=1, keepdims=True)
    inp[:, 9] = rng.uniform(0, 1, n)
    tgt = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    return mlp, grid, inp, tgt


def test_level_table(orc):
    c = orc.cache_cfg(16, 19, 16, 1.38, (0, 0, 0), (1, 1, 1))
    res = [c.res[l] for l in range(16)]
    assert res[0] == 16 and all(b >= a for a, b in zip(res, res[1:]))
    assert res[15] == int(np.floor(16 * 1.38 ** 15))
    # dense iff (res + 1)^3 fits the table
    for l in range(16):
        assert c.dense[l] == (1 if (res[l] + 1) ** 3 <= 2 ** 19 else 0)


def test_loss_closed_form_zero_network(orc):
    """With every weight zero the output is 0 and the relative L2 loss is
    mean(t^2 / 0.01)."""
    cfg = small_cfg(orc)
    mlp = np.zeros(17408, np.float32)
    grid = np.zeros(cfg.levels * cfg.table_size * 2, np.float32)
    _, _, inp, tgt = random_problem(16, cfg.levels, cfg.table_size)
    loss, _, _, y = orc.cache_loss_grad(cfg, mlp, grid, inp, tgt, emulate_bf16=False)
    assert np.all(y == 0)
    assert abs(loss - np.mean(tgt.astype(np.float64) ** 2 / 0.01)) < 1e-9


@pytest.mark.parametrize("which", ["mlp", "grid"])
def test_gradients_match_finite_differences(orc, which):
    cfg = small_cfg(orc)
    mlp, grid, inp, tgt = random_problem(24, cfg.levels, cfg.table_size, seed=3)
    loss, mg, gg, y = orc.cache_loss_grad(cfg, mlp, grid, inp, tgt, emulate_bf16=False)
    # the loss's denominators are stop-gradient values: hold them fixed
    den = y * y + 0.01
    rng = np.random.default_rng(7)
    if which == "mlp":
        # used entries of every layer (the output layer has 3 live columns)
        cand = [l * 4096 + k * 64 + j for l in range(4) for k in range(64) for j in range(64)]
        cand += [16384 + k * 16 + j for k in range(64) for j in range(3)]
        picks = rng.choice(cand, 40, replace=False)
        base, grad = mlp, mg
    else:
        nz = np.nonzero(gg)[0]
        picks = rng.choice(nz, min(40, nz.size), replace=False)
        base, grad = grid, gg
    h = 1e-3
    for i in picks:
        p = base.astype(np.float64).copy()
        plus, minus = p.copy(), p.copy()
        plus[i] += h
        minus[i] -= h
        if which == "mlp":
            lp = orc.cache_loss_grad(cfg, plus.astype(np.float32), grid, inp, tgt, False, False, den)[0]
            lm = orc.cache_loss_grad(cfg, minus.astype(np.float32), grid, inp, tgt, False, False, den)[0]
        else:
            lp = orc.cache_loss_grad(cfg, mlp, plus.astype(np.float32), inp, tgt, False, False, den)[0]
            lm = orc.cache_loss_grad(cfg, mlp, minus.astype(np.float32), inp, tgt, False, False, den)[0]
        # the FP32 parameter perturbation is not exactly h: use the realised step
        hp = np.float64(np.float32(p[i] + h)) - p[i]
        hm = p[i] - np.float64(np.float32(p[i] - h))
        fd = (lp - lm) / (hp + hm)
        assert abs(fd - grad[i]) <= 2e-4 * max(1.0, abs(grad[i])) + 1e-6, (which, i, fd, grad[i])


def test_bf16_emulation_is_close_to_fp64(orc):
    cfg = small_cfg(orc)
    mlp, grid, inp, tgt = random_problem(64, cfg.levels, cfg.table_size, seed=5)
    l0, m0, g0, y0 = orc.cache_loss_grad(cfg, mlp, grid, inp, tgt, emulate_bf16=False)
    l1, m1, g1, y1 = orc.cache_loss_grad(cfg, mlp, grid, inp, tgt, emulate_bf16=True)
    assert abs(l1 - l0) < 2e-2 * l0
    cos = m0 @ m1 / np.linalg.norm(m0) / np.linalg.norm(m1)
    assert cos > 0.999
