"""Pins of the path-guiding oracle (oracle/guiding.py): finite differences,
quadrature, sampling histograms and a closed-form Adam step.  CPU only."""
import numpy as np
import pytest

from oracle import guiding as G


def _batch(n=32, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 2, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    X = G.features(pos, nrm, wo, [-1, -1, -1], [2, 2, 2])
    y = rng.uniform(0.1, 3.0, n)
    q = rng.uniform(0.05, 1.0, n)
    return X, d, y, q


def test_features_layout():
    X = G.features(np.array([[0.5, 0.25, 1.0]]), np.array([[0, 0, 1.0]]),
                   np.array([[1.0, 0, 0]]), [0, 0, 0], [1, 1, 1])
    assert X.shape == (1, 64)
    assert np.isclose(X[0, 0], np.sin(np.pi * 0.5)) and np.isclose(X[0, 1], np.cos(np.pi * 0.5))
    assert np.isclose(X[0, 10 + 2 * 3], np.sin(8 * np.pi * 0.25))
    assert np.array_equal(X[0, 32:38], [0, 0, 1, 1, 0, 0])
    assert not X[0, 38:].any() and not X[0, 30:32].any()


@pytest.mark.parametrize("width", [64, 128])
def test_backward_matches_finite_differences(width):
    P = G.init_params(7, width=width)
    assert [W.shape for W, _ in P] == [(64, width), (width, width), (width, width),
                                       (width, width), (width, 48)]
    X, d, y, q = _batch()
    out, acts = G.forward(P, X)
    _, Gout = G.loss_and_grad_out(out, d, y, q)
    grads = G.backward(P, acts, Gout)
    rng = np.random.default_rng(3)
    h = 1e-6
    for li in range(len(P)):
        for j in range(2):
            arr = P[li][j]
            flat = arr.reshape(-1)
            for idx in rng.choice(flat.size, 6, replace=False):
                old = flat[idx]
                flat[idx] = old + h
                lp = G.loss_fn(P, X, d, y, q)
                flat[idx] = old - h
                lm = G.loss_fn(P, X, d, y, q)
                flat[idx] = old
                num = (lp - lm) / (2 * h)
                ana = grads[li][j].reshape(-1)[idx]
                assert abs(num - ana) <= 1e-5 * max(1.0, abs(num)), (li, j, idx, num, ana)


def test_output_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    out = rng.normal(size=(16, 48))
    _, d, y, q = _batch(16, 1)
    _, Gout = G.loss_and_grad_out(out, d, y, q)
    h = 1e-6
    for i in range(16):
        for k in rng.choice(40, 8, replace=False):
            o1, o2 = out.copy(), out.copy()
            o1[i, k] += h
            o2[i, k] -= h
            num = (G.loss_and_grad_out(o1, d, y, q)[0] - G.loss_and_grad_out(o2, d, y, q)[0]) / (2 * h)
            assert abs(num - Gout[i, k]) <= 1e-5 * max(1.0, abs(num))
    assert not Gout[:, 40:].any()


def _sphere_quadrature(n=400):
    th = (np.arange(n) + 0.5) / n * np.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / n) * (np.pi / n)).reshape(-1)
    return d, dA


def test_vmf_normalised_and_mixture_normalised():
    d, dA = _sphere_quadrature(600)
    for kap in (0.05, 1.0, 10.0, 200.0):
        mu = np.array([0.3, -0.5, 0.81])
        mu /= np.linalg.norm(mu)
        tot = (G.vmf_pdf(mu[None], np.array([kap]), d) * dA).sum()
        assert abs(tot - 1) < 2e-3 * max(1, kap / 50), (kap, tot)
    rng = np.random.default_rng(2)
    out = rng.normal(size=(1, 48))
    tot = (G.mixture_pdf(np.repeat(out, len(d), 0), d) * dA).sum()
    assert abs(tot - 1) < 5e-3


def test_vmf_sampling_histogram():
    """Samples of Jakob's method match the density (chi-square over bins of
    the cosine to the mean), and the mean cosine matches coth(k) - 1/k."""
    rng = np.random.default_rng(9)
    mu = np.array([0.0, 0.6, 0.8])
    for kap in (0.5, 5.0, 50.0):
        n = 200000
        s = G.sample_vmf(np.tile(mu, (n, 1)), np.full(n, kap), rng.uniform(size=n),
                         rng.uniform(size=n))
        assert np.allclose(np.linalg.norm(s, axis=1), 1, atol=1e-9)
        c = s @ mu
        expect = 1 / np.tanh(kap) - 1 / kap
        assert abs(c.mean() - expect) < 5 * c.std() / np.sqrt(n)
        # cos is distributed with density k e^{k(c-1)} / (1 - e^{-2k}) on [-1, 1]
        edges = np.linspace(-1, 1, 21)
        cdf = lambda x: (np.exp(kap * (x - 1)) - np.exp(-2 * kap)) / (1 - np.exp(-2 * kap))
        p = np.diff(cdf(edges))
        h = np.histogram(c, edges)[0]
        m = p * n > 20
        chi2 = (((h - p * n) ** 2) / (p * n))[m].sum()
        assert chi2 < 3 * m.sum(), (kap, chi2)


def test_adam_single_step_closed_form():
    """First Adam step moves every parameter by -lr * sign(g) (bias-corrected
    m / sqrt(v) = g / |g|)."""
    P = [[np.array([[1.0, -2.0]]), np.array([0.5, 0.0])]]
    grads = [(np.array([[0.3, -4.0]]), np.array([1e-3, -2.0]))]
    opt = G.Adam(P, lr=1e-2)
    P = opt.step(P, grads)
    assert np.allclose(P[0][0], [[1.0 - 1e-2, -2.0 + 1e-2]], atol=1e-9)
    assert np.allclose(P[0][1], [0.5 - 1e-2 * 1e-3 / (1e-3 + 1e-8), 0.0 + 1e-2], atol=1e-9)


def test_training_reduces_kl_on_a_fixed_target():
    """Fitting a known single-lobe target distribution: the KL objective
    decreases and the fitted mixture puts its mass near the target mean."""
    rng = np.random.default_rng(4)
    P = G.init_params(11)
    opt = G.Adam(P, lr=1e-2)
    target_mu = np.array([0.0, 0.0, 1.0])
    X, _, _, _ = _batch(256, 8)
    losses = []
    for it in range(150):
        d = rng.normal(size=(256, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        q_s = np.full(256, 1 / (4 * np.pi))
        y = G.vmf_pdf(target_mu[None], np.array([8.0]), d)
        out, acts = G.forward(P, X)
        loss, Gout = G.loss_and_grad_out(out, d, y, q_s)
        P = opt.step(P, G.backward(P, acts, Gout))
        losses.append(loss)
    assert np.mean(losses[-10:]) < 0.8 * np.mean(losses[:10])
    out, _ = G.forward(P, X[:1])
    pdf_at_mean = G.mixture_pdf(out, target_mu[None])[0]
    assert pdf_at_mean > 5 / (4 * np.pi)
