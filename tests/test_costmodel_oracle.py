"""CPU checks of the cost-model oracle (oracle/mlp_oracle.py) and of the
library's host-side cost-model calls (create, save, load, get, set: no GPU
work).  The GPU kernels are compared with this oracle in
tests/test_gpu_costmodel.py."""

import os

import numpy as np
import pytest
import torch

import kclistpp_amd as kcl
from oracle import mlp_oracle as mo


def _rand_model(seed):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed).astype(np.float64)
    p[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.3, mo.H)
    p[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.3, mo.H)
    p[mo.OFF_B3:] = rng.normal(0, 0.3, mo.O)
    norm = mo.identity_norm()
    norm[:mo.F] = rng.normal(0, 1, mo.F)
    norm[mo.F:2 * mo.F] = rng.uniform(0.5, 2, mo.F)
    norm[2 * mo.F:2 * mo.F + mo.O] = rng.normal(0, 1, mo.O)
    norm[2 * mo.F + mo.O:] = rng.uniform(0.5, 2, mo.O)
    x = rng.normal(0, 1.5, (37, mo.F)) + norm[:mo.F]
    y = rng.normal(0, 1, (37, mo.O))
    return p, norm, x, y


def _fd_grad(p, norm, x, y, idx, eps=1e-6):
    g = np.zeros(len(idx))
    for j, i in enumerate(idx):
        a = p.copy()
        b = p.copy()
        a[i] += eps
        b[i] -= eps
        g[j] = (mo.loss_and_grad(a, norm, x, y, bf=False)[0] -
                mo.loss_and_grad(b, norm, x, y, bf=False)[0]) / (2 * eps)
    return g


@pytest.mark.parametrize("seed", [1, 2])
def test_gradient_matches_finite_differences(seed):
    p, norm, x, y = _rand_model(seed)
    _, g = mo.loss_and_grad(p, norm, x, y, bf=False)
    rng = np.random.default_rng(seed + 10)
    # every bias and a sample of every weight matrix
    idx = np.concatenate([
        rng.choice(np.arange(mo.OFF_W1, mo.OFF_B1), 60, replace=False),
        np.arange(mo.OFF_B1, mo.OFF_W2),
        rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 120, replace=False),
        np.arange(mo.OFF_B2, mo.OFF_W3),
        np.arange(mo.OFF_W3, mo.P)])
    fd = _fd_grad(p, norm, x, y, idx)
    scale = np.max(np.abs(fd))
    assert scale > 1e-3
    np.testing.assert_allclose(g[idx], fd, rtol=0, atol=1e-6 * scale + 1e-9)


@pytest.mark.parametrize("plant", ["transpose_w2", "drop_mask1"])
def test_planted_backprop_mistakes_are_caught(plant):
    p, norm, x, y = _rand_model(5)
    _, g = mo.loss_and_grad(p, norm, x, y, bf=False, plant=plant)
    idx = np.arange(mo.OFF_W1, mo.OFF_W2)  # first-layer weights and biases
    fd = _fd_grad(p, norm, x, y, idx)
    assert np.max(np.abs(g[idx] - fd)) > 1e-3 * np.max(np.abs(fd))


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([
        rng.normal(0, 3, 5000).astype(np.float32),
        # exact ties: bit 15 set, lower bits clear, both parities of bit 16
        (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
        np.array([0.0, -0.0, 1.0, -1.0, 1e-30, 3e38], np.float32)])
    want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16(v), want)


def test_bf16_forward_is_close_to_float64():
    p, norm, x, _ = _rand_model(3)
    a = mo.predict(p, norm, x, bf=True)
    b = mo.predict(p, norm, x, bf=False)
    assert np.max(np.abs(a - b)) < 0.05 * (1 + np.max(np.abs(b)))


def test_closed_form_single_linear_path():
    # Only x0 -> unit 0 -> unit 0 -> output 0, all weights 1, biases 0:
    # out = relu(relu(x0)) for both precisions when x0 is bf16-exact.
    p = np.zeros(mo.P)
    p[mo.OFF_W1 + 0 * mo.H + 0] = 1.0
    p[mo.OFF_W2 + 0 * mo.H + 0] = 1.0
    p[mo.OFF_W3 + 0 * mo.O + 0] = 1.0
    x = np.zeros((4, mo.F))
    x[:, 0] = [-2.0, 0.5, 3.0, 1.25]
    out = mo.predict(p, mo.identity_norm(), x)
    np.testing.assert_array_equal(out[:, 0], [0.0, 0.5, 3.0, 1.25])
    np.testing.assert_array_equal(out[:, 1], 0.0)
    # dL/db3 = mean(out - y) per output
    y = np.ones((4, mo.O))
    _, g = mo.loss_and_grad(p, mo.identity_norm(), x, y, bf=False)
    np.testing.assert_allclose(g[mo.OFF_B3:], [np.mean(out[:, 0] - 1), -1.0])


def test_library_init_matches_oracle():
    with kcl.CostModel(seed=1234) as m:
        params, norm = m.get()
    np.testing.assert_array_equal(params, mo.init_params(1234))
    np.testing.assert_array_equal(norm, mo.identity_norm())
    assert np.all(params[mo.OFF_B1:mo.OFF_W2] == 0)


def test_model_file_round_trip(tmp_path):
    p, norm, _, _ = _rand_model(7)
    path = str(tmp_path / "m.kclcm")
    with kcl.CostModel(seed=0) as m:
        m.set(p.astype(np.float32), norm)
        m.save(path)
    assert os.path.getsize(path) == 8 + 4 * (mo.NORM + mo.P)
    with kcl.CostModel(path=path) as m2:
        p2, n2 = m2.get()
    np.testing.assert_array_equal(p2, p.astype(np.float32))
    np.testing.assert_array_equal(n2, norm)


def test_model_file_and_values_are_checked(tmp_path):
    bad = tmp_path / "bad.kclcm"
    bad.write_bytes(b"KCLCM001" + b"\0" * 10)
    with pytest.raises(kcl.KclError) as e:
        kcl.CostModel(path=str(bad))
    assert e.value.name == "KCL_E_PARSE"
    with pytest.raises(kcl.KclError):
        kcl.CostModel(path=str(tmp_path / "missing.kclcm"))
    with kcl.CostModel(seed=0) as m:
        norm = mo.identity_norm()
        norm[mo.F] = 0.0  # zero feature std
        with pytest.raises(kcl.KclError):
            m.set(None, norm)
        p = np.zeros(mo.P, np.float32)
        p[3] = np.nan
        with pytest.raises(kcl.KclError):
            m.set(p, None)


def _committed():
    here = os.path.dirname(os.path.abspath(__file__))
    d = np.load(os.path.join(here, "golden", "cost_samples_k5to7.npz"))
    with kcl.CostModel(path=os.path.join(kcl.PKG, "cost_model.kclcm")) as m:
        p, norm = m.get()
    return d["x"], d["y"], p, norm


def _spearman(a, b):
    ra = np.argsort(np.argsort(a))
    rb = np.argsort(np.argsort(b))
    return np.corrcoef(ra, rb)[0, 1]


def test_committed_model_predicts_measured_costs():
    """The shipped model (trained on the GPU from MI355X per-root timings)
    fits the counting cycles of k >= 5 roots, and ranks roots by total cost
    better than the plain schedule's d+^2 does."""
    x, y, p, norm = _committed()
    assert np.all(np.isfinite(p)) and np.all(norm[mo.F:2 * mo.F] > 0)
    pr = mo.predict(p, norm, x)
    r2 = 1 - np.mean((pr[:, 1] - y[:, 1]) ** 2) / np.var(y[:, 1])
    assert r2 > 0.85, r2
    true_cost = np.exp2(y[:, 0]) + np.exp2(y[:, 1])
    pred_cost = np.exp2(pr[:, 0]) + np.exp2(pr[:, 1])
    d = np.exp2(x[:, 0]) - 1
    learned, plain = _spearman(pred_cost, true_cost), _spearman(d * d, true_cost)
    assert learned > plain + 0.1, (learned, plain)
