"""Pins of the FP64 restatement of the learned T2 guess (oracle/mlp.py).

- The backpropagated gradient against central finite differences.
- GELU (tanh form) against the exact erf form and its derivative against
  finite differences.
- Adam's first step is -lr * sign(g) for |g| >> eps (m_hat = g, v_hat = g^2).
- The features of the psi3 H2O run, rounded to FP32 and inverted, give the
  reference's saved E_MP2 (psi3/output.c1_fzc, iteration 0).
- The guess keeps the pair symmetry t2_ijab = t2_jiba and reduces to MP2 for
  a network whose output is zero."""
import json
import math
import os

import numpy as np
import pytest

from oracle import mlp

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


def _random_params(rng, scale=1.0):
    d = {}
    for name, shape in mlp.layout():
        fan = shape[1] if name.startswith("W") and len(shape) == 2 else mlp.WIDTH
        if name in ("b1",):
            fan = mlp.NFEAT
        d[name] = rng.uniform(-1, 1, size=shape) * scale / math.sqrt(fan)
    return mlp.pack(d)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _random_params(rng, scale=1.5)
    X = rng.normal(size=(24, mlp.NFEAT))
    y = rng.normal(size=24)
    shift, scale = rng.normal(size=mlp.NFEAT) * 0.1, rng.uniform(0.5, 2.0, size=mlp.NFEAT)
    loss, g = mlp.loss_grad(p, X, y, shift, scale)
    # every parameter of the small layers, and a sample of the hidden ones
    idx = list(range(0, 1152)) + list(rng.choice(np.arange(1152, len(p)), size=600, replace=False)) + [len(p) - 1]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_grad(pp, X, y, shift, scale)[0] - mlp.loss_grad(pm, X, y, shift, scale)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])
    assert loss == pytest.approx(np.mean((mlp.forward(p, X, shift, scale) - y) ** 2), rel=1e-14)


def test_gelu_against_erf_form_and_derivative():
    z = np.linspace(-8, 8, 4001)
    exact = 0.5 * z * (1.0 + np.array([math.erf(t / math.sqrt(2.0)) for t in z]))
    # the tanh form approximates the erf form to about 5e-4 everywhere;
    # dropping the cubic term would give 2e-2
    assert np.max(np.abs(mlp.gelu(z) - exact)) < 1e-3
    assert mlp.gelu(np.array([0.0]))[0] == 0.0
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp.gelu_grad(z), fd, atol=1e-8)


def test_adam_first_step_is_signed_learning_rate():
    rng = np.random.default_rng(0)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    p1, m, v = mlp.adam_step(p, g, np.zeros(100), np.zeros(100), 1, lr=1e-3)
    np.testing.assert_allclose(p1 - p, -1e-3 * np.sign(g), rtol=1e-6)
    # on f = |p|^2 / 2 Adam walks to the minimum
    q, m, v = np.full(5, 3.0), np.zeros(5), np.zeros(5)
    for t in range(1, 3001):
        q, m, v = mlp.adam_step(q, q, m, v, t, lr=1e-2)
    assert np.max(np.abs(q)) < 1e-2


@pytest.fixture(scope="module")
def h2o_fc():
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    return Reference(molecules.PSI3_H2O, "cc-pvdz", nfzc=1)


def test_features_reproduce_reference_mp2_energy(h2o_fc):
    ref = h2o_fc
    no, nv = ref.no, ref.nv
    eps = np.diag(ref.fock)
    X = mlp.features(no, nv, eps, ref.blocks["oovv"]).astype(np.float32)
    t, I, K, D = mlp.invert_features(X)
    emp2 = float(np.sum((2.0 * I - K) * t))
    assert abs(emp2 - KA["h2o_ccpvdz_fc1"]["ccsd_trace"][0]) < 1e-7
    # the denominators and orbital energies agree with the Fock diagonal
    e_o, e_v = eps[:no], eps[no:]
    D_ref = (e_o[:, None, None, None] + e_o[None, :, None, None] - e_v[None, None, :, None]
             - e_v[None, None, None, :]).ravel()
    np.testing.assert_allclose(D, D_ref, rtol=1e-6)
    np.testing.assert_allclose(X[:, 4], np.repeat(e_o, no * nv * nv), rtol=1e-6)


def test_guess_symmetry_and_mp2_limit(h2o_fc):
    ref = h2o_fc
    no, nv = ref.no, ref.nv
    eps = np.diag(ref.fock)
    rng = np.random.default_rng(1)
    p = _random_params(rng)
    shift, scale = np.zeros(8), np.ones(8)
    t2 = mlp.guess(p, no, nv, eps, ref.blocks["oovv"], shift, scale, 1e-3)
    np.testing.assert_allclose(t2, t2.transpose(1, 0, 3, 2), atol=1e-15)
    d = mlp.unpack(p)
    d["W4"][:] = 0.0
    d["b4"][:] = 0.0
    t2_0 = mlp.guess(mlp.pack(d), no, nv, eps, ref.blocks["oovv"], shift, scale, 1e-3)
    e_o, e_v = eps[:no], eps[no:]
    D = e_o[:, None, None, None] + e_o[None, :, None, None] - e_v[None, None, :, None] - e_v[None, None, None, :]
    np.testing.assert_allclose(t2_0, ref.blocks["oovv"] / D, rtol=1e-15)
