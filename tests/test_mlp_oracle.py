"""Pins the float64 network oracle: its gradient against central finite differences, its bfloat16
rounding against hand-checked values (no GPU needed)."""
import numpy as np

from oracle import mlp_ref as R


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1.0; 1 + 3 * 2^-9 rounds up;
    # 3e-5 = 1.96608 * 2^-16 -> 252/128 * 2^-16 (7 mantissa bits); 65504 = 1.99902 * 2^15 -> 2^16
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e-5, 65504.0], np.float32)
    got = R.bf16(x)
    want = [1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, 252.0 / 128.0 * 2.0 ** -16, 65536.0]
    assert np.allclose(got, want, rtol=0, atol=0), got


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = R.init_params(5)
    X = rng.normal(0, 1, (7, R.N_IN))
    T = rng.normal(0, 0.3, (7, R.N_OUT))
    mu = np.zeros(R.N_IN)
    inv = np.ones(R.N_IN)
    scale = np.array([0.5, 0.2, 0.2])
    # float32 features go through the same normalisation; keep them exactly representable
    X = X.astype(np.float32).astype(np.float64)
    loss, g = R.loss_and_grad(p, X, T, mu, inv, scale, bf=False)
    idx = np.concatenate([rng.choice(np.arange(R.P_W1, R.P_B1), 12, replace=False),
                          rng.choice(np.arange(R.P_B1, R.P_W2), 4, replace=False),
                          rng.choice(np.arange(R.P_W2, R.P_B2), 12, replace=False),
                          rng.choice(np.arange(R.P_B2, R.P_W3), 4, replace=False),
                          rng.choice(np.arange(R.P_W3, R.P_B3), 8, replace=False),
                          np.arange(R.P_B3, R.N_PARAMS)])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_and_grad(pp, X, T, mu, inv, scale, bf=False)[0] -
              R.loss_and_grad(pm, X, T, mu, inv, scale, bf=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_bf16_oracle_close_to_exact():
    rng = np.random.default_rng(4)
    p = R.init_params(6)
    X = rng.normal(0, 1, (64, R.N_IN))
    T = rng.normal(0, 0.3, (64, R.N_OUT))
    args = (np.zeros(R.N_IN), np.ones(R.N_IN), np.ones(3))
    y_bf, _ = R.forward(p, X, *args, bf=True)
    y_ex, _ = R.forward(p, X, *args, bf=False)
    assert np.max(np.abs(y_bf - y_ex)) < 0.05 * np.max(np.abs(y_ex)) + 1e-3
    _, g_bf = R.loss_and_grad(p, X, T, *args, bf=True)
    _, g_ex = R.loss_and_grad(p, X, T, *args, bf=False)
    assert np.linalg.norm(g_bf - g_ex) < 0.05 * np.linalg.norm(g_ex)
