"""CPU checks of the MLP oracle (oracle/mlp_ref.py): its analytic gradient
against central finite differences, the Adam update against its closed
form, and the feature normalisation's invariances."""
import numpy as np

from oracle import mlp_ref as R


def test_param_count_matches_header():
    assert R.NP == 6 * 32 + 32 + 2 * (32 * 32 + 32) + 32 + 1 == 2369


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = R.init_params(3)
    p[R.NP - 1] = 0.3  # non-zero output bias
    X = rng.standard_normal((40, 6))
    t = rng.standard_normal(40)
    mask = (rng.random(40) < 0.7).astype(float)
    loss, g = R.loss_grad(p, X, t, mask)
    idx = rng.choice(R.NP, 60, replace=False)
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_grad(pp, X, t, mask)[0] - R.loss_grad(pm, X, t, mask)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_masked_rows_do_not_contribute():
    rng = np.random.default_rng(1)
    p = R.init_params(1)
    X = rng.standard_normal((30, 6))
    t = rng.standard_normal(30)
    m = np.zeros(30)
    m[:10] = 1
    l1, g1 = R.loss_grad(p, X, t, m)
    l2, g2 = R.loss_grad(p, X[:10], t[:10])
    assert np.isclose(l1, l2) and np.allclose(g1, g2)


def test_adam_first_step_is_sign_times_lr():
    g = np.array([0.5, -2.0, 1e-3])
    p, m, v = R.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    # bias-corrected first step: m_hat/sqrt(v_hat) = g/|g| (up to eps)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-4)


def test_features_invariances():
    """p*/p_ref is invariant under rho -> a rho, p -> b p, u -> sqrt(b/a) u + U
    (scaling and Galilean invariance of the Riemann problem): the features
    must be too."""
    rng = np.random.default_rng(2)
    F = np.stack([np.exp(rng.uniform(-2, 2, 50)), np.exp(rng.uniform(-2, 2, 50)),
                  rng.uniform(-1, 1, 50), np.zeros(50), np.exp(rng.uniform(-2, 2, 50)),
                  np.exp(rng.uniform(-2, 2, 50)), rng.uniform(-1, 1, 50), np.zeros(50)])
    X0, _, _ = R.features(F, 5 / 3)
    a, b, U = 3.0, 0.2, 0.7
    G = F.copy()
    G[[0, 4]] *= a
    G[[1, 5]] *= b
    G[[2, 6]] = G[[2, 6]] * np.sqrt(b / a) + U
    X1, _, _ = R.features(G, 5 / 3)
    assert np.allclose(X0, X1, atol=1e-12)
