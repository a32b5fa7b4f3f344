"""The surrogate's CPU oracle (oracle/mlp_oracle.py) pinned against library
routines and mathematics, and the host side of the surrogate (feature
encoding, parameter layout, CSV parsing, flags) against the oracle. No
kernel is launched here.

Pins of the oracle:
  * bf16 rounding == torch's float32 -> bfloat16 conversion (round to nearest
    even), including ties, overflow, subnormals and NaN;
  * float64 forward == torch.nn.functional (linear, gelu approximate='none');
  * float64 gradient == torch autograd, and == central finite differences on
    random coordinates of every tensor (independent of torch);
  * Adam == torch.optim.Adam over several steps.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

EXE = os.path.join(ROOT, "ftlm_extended_hubbard_development_amd", "ftlm_n_vs_mu")


def rand_inputs(n, seed):
    rng = np.random.default_rng(seed)
    lo = np.array([4.0, 0.0, -0.3, 0.0, 0.0, 2.0, -5.0])
    hi = np.array([10.0, 2.0, 0.3, 1.0, 1.0, 20.0, 15.0])
    return lo + (hi - lo) * rng.random((n, 7))


def rand_params(seed, scale=1.0):
    from ftlm_extended_hubbard_development_amd import surrogate

    p = surrogate.init_params(seed).astype(np.float64)
    return p * scale


def torch_net(p, feat):
    layers = [(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True)) for W, b in mo.unpack(p)]
    h = torch.tensor(np.asarray(feat, np.float64))
    for j in range(mo.NHID):
        h = torch.nn.functional.gelu(torch.nn.functional.linear(h, *layers[j]), approximate="none")
    y = torch.nn.functional.linear(h, *layers[mo.NHID])
    return y, layers


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-40, 38, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.005859375, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8,
                                  3.3895313892515355e38, 1e-40, -1e-42, np.inf, -np.inf])]).astype(np.float32)
    ours = mo.bf16(x)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, ref)
    assert np.isnan(mo.bf16(np.array([np.nan], np.float32)))[0]


def test_layout_matches_engine():
    from ftlm_extended_hubbard_development_amd import surrogate

    wo, bo, no, ni = surrogate.layout()
    assert list(wo) == mo.W_OFF and list(bo) == mo.B_OFF
    assert [(int(a), int(b)) for a, b in zip(no, ni)] == mo.SHAPES
    assert surrogate.param_count() == mo.N_PARAMS


def test_encoding_matches_engine():
    from ftlm_extended_hubbard_development_amd import surrogate

    x = rand_inputs(4000, 1)
    ours = surrogate.encode(x)
    ref = mo.encode(x)
    assert ours.dtype == np.float32 and ours.shape == (4000, 32)
    assert np.abs(ours - ref).max() <= 1e-7
    assert not ours[:, 16:].any()


def test_init_params_padding_zero():
    p = rand_params(5)
    layers = mo.unpack(p)
    assert not layers[0][0][:, 16:].any()
    assert not layers[4][0][2:, :].any() and not layers[4][1][2:].any()
    assert np.abs(layers[1][0]).max() <= 1 / 16 + 1e-12


def test_forward_matches_torch():
    p = rand_params(7)
    feat = mo.encode(rand_inputs(50, 2))
    y, _ = mo.forward(p, feat, emulate=False)
    yt, _ = torch_net(p, feat)
    assert np.allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)


def test_gradient_matches_torch_autograd():
    p = rand_params(8)
    x = rand_inputs(40, 3)
    feat = mo.encode(x)
    tgt = np.random.default_rng(4).standard_normal((40, 2))
    loss, g = mo.loss_and_grad(p, feat, tgt, emulate=False)
    yt, layers = torch_net(p, feat)
    lt = 0.5 * ((yt[:, :2] - torch.tensor(tgt)) ** 2).sum() / 40
    lt.backward()
    gt = mo.pack([(W.grad.numpy(), b.grad.numpy()) for W, b in layers])
    assert abs(loss - lt.item()) <= 1e-12 * max(1.0, abs(loss))
    assert np.allclose(g, gt, rtol=1e-10, atol=1e-14)


def test_gradient_matches_finite_differences():
    p = rand_params(9)
    feat = mo.encode(rand_inputs(16, 5))
    tgt = np.random.default_rng(6).standard_normal((16, 2))
    _, g = mo.loss_and_grad(p, feat, tgt, emulate=False)
    rng = np.random.default_rng(7)
    idx = []
    for l, (n, k) in enumerate(mo.SHAPES):  # coordinates of every weight and bias tensor (used part)
        kk = 16 if l == 0 else k
        nn = 2 if l == 4 else n
        idx += [mo.W_OFF[l] + int(rng.integers(nn)) * k + int(rng.integers(kk)) for _ in range(4)]
        idx += [mo.B_OFF[l] + int(rng.integers(nn)) for _ in range(2)]
    h = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += h
        lp, _ = mo.loss_and_grad(pp, feat, tgt, emulate=False)
        pp[i] -= 2 * h
        lm, _ = mo.loss_and_grad(pp, feat, tgt, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_emulated_gradient_close_to_exact():
    """bf16 storage perturbs the gradient at the bf16 level, not more."""
    p = rand_params(10)
    feat = mo.encode(rand_inputs(256, 8))
    tgt = np.random.default_rng(9).standard_normal((256, 2))
    _, g = mo.loss_and_grad(p, feat, tgt, emulate=False)
    _, ge = mo.loss_and_grad(p, feat, tgt, emulate=True)
    for l in range(5):
        a = slice(mo.W_OFF[l], mo.B_OFF[l])
        rel = np.linalg.norm(ge[a] - g[a]) / np.linalg.norm(g[a])
        assert rel < 3e-2, (l, rel)


def test_adam_matches_torch():
    rng = np.random.default_rng(11)
    p0 = rng.standard_normal(300)
    tp = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(300), np.zeros(300)
    for t in range(1, 6):
        g = rng.standard_normal(300)
        g[:10] = 0.0  # untouched coordinates stay put
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = mo.adam(p, g, m, v, t, 3e-3)
        assert np.allclose(p, tp.detach().numpy(), rtol=1e-13, atol=1e-15)
    assert np.array_equal(p[:10], p0[:10])


def test_read_multi_beta_csv():
    from ftlm_extended_hubbard_development_amd import surrogate

    b, m, n, k = surrogate._read_multi_beta_csv(os.path.join(ROOT, "tests", "golden", "csv",
                                                             "4x2_multibeta_r8_s60.csv"))
    assert b[0] == 20.0 and m[0] == -5.0 and abs(n[0] - 5.15288403485531e-10) < 1e-24
    assert k[0] == 1.03057680272284e-08
    assert b.size == m.size == n.size == k.size and set(b) == {20.0, 5.0}


@pytest.mark.parametrize("args,msg", [
    (["--twist-list", "0.1:0.2,0.5"], "Error: --twist-list entries must be PX:PY, got: 0.5"),
    (["--twist-list", "0.1:0.2", "--twist-grid", "2"], "Error: --twist-grid and --twist-list are exclusive."),
    (["--twist-list", "0.1:0.2", "--checkpoint", "c.bin"],
     "Error: --twist-grid/--twist-list are not compatible with --checkpoint or trace/debug block output."),
])
def test_twist_list_flag_errors(tmp_path, args, msg):
    p = subprocess.run([EXE] + args, cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert p.returncode == 1
    assert p.stderr.strip() == msg


def test_draws_are_seeded_and_in_range():
    from ftlm_extended_hubbard_development_amd import surrogate

    a = surrogate.draw_models(20, 3)
    assert a == surrogate.draw_models(20, 3) and a != surrogate.draw_models(20, 4)
    for u, v, tp in a:
        assert 4 <= u <= 10 and 0 <= v <= 2 and -0.3 <= tp <= 0.3


def test_pretrained_parameters_ship_with_the_package():
    """The shipped 4x2 network loads without unpickling and fits the engine's
    layout; its normalisation is finite and its provenance recorded."""
    from ftlm_extended_hubbard_development_amd import surrogate

    d = np.load(surrogate.PRETRAINED["4x2"], allow_pickle=False)
    assert d["params"].dtype == np.float32 and d["params"].size == surrogate.param_count()
    assert np.isfinite(d["params"]).all() and (d["y_scale"] > 0).all()
    meta = json.loads(str(d["meta"]))
    assert meta["lattice"] == "4x2" and "generate" in meta["data"]
    with pytest.raises(ValueError, match="no pretrained surrogate"):
        surrogate.pretrained("5x5")


def test_gain_low_samples_needs_exact():
    """The fewer-sample FTLM scans are only reported against the exact answer;
    asking for them without it is an argument error raised before any work."""
    from ftlm_extended_hubbard_development_amd import surrogate

    with pytest.raises(ValueError, match="exact=True"):
        surrogate.gain(None, "4x2", 8.0, 0.5, 0.1, 8, [2.0], low_samples=[16])
