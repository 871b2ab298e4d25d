"""Pins the numpy oracle of the learned-guess network (CPU only): the backward
pass against central finite differences of the FP64 loss, the convolution
against a direct loop, and the bf16 rounding against known bit patterns."""
import numpy as np

from oracle import cnn_oracle as N


def sample(shape=(5, 6, 7), seed=0):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.linspace(0, 1, s) for s in shape], indexing="ij")
    Tnm1 = np.sin(3 * x) * np.cos(2 * y) + z
    Tn = Tnm1 + 0.1 * np.exp(-((x - 0.4) ** 2 + (y - 0.5) ** 2) / 0.05) + 0.01 * rng.standard_normal(shape)
    Tnp1 = Tn + 0.12 * np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.05)
    q = np.where(np.arange(shape[0])[:, None, None] == shape[0] - 1, np.exp(-((x - 0.5) ** 2) / 0.02), 0.0)
    # stands in for the scaled residual of the linear extrapolation, which is close to the
    # correction T_{n+1} - (2 T_n - T_{n-1}) for a mass-dominated operator: that plus noise
    zres = 0.8 * (Tnp1 - (2 * Tn - Tnm1)) + 0.002 * rng.standard_normal(shape)
    return Tnm1, Tn, q, zres, Tnp1


def init(seed=1):
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in zip(N.CIN, N.COUT):
        parts.append(rng.standard_normal(co * ci * 27) * np.sqrt(2.0 / (27 * ci)))
        parts.append(0.1 * rng.standard_normal(co))
    return np.concatenate(parts)


def test_bf16_rounding():
    assert N.bf16(1.0) == 1.0
    assert N.bf16(1.0 + 2 ** -8) == 1.0          # tie rounds to even
    assert N.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert N.bf16(-3.14159) == -3.140625


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 4, 5, 6))
    W = rng.standard_normal((2, 3, 27))
    b = rng.standard_normal(2)
    out = N.conv(x, W, b)
    ref = np.zeros_like(out)
    C, nz, ny, nx = x.shape
    for co in range(2):
        for z in range(nz):
            for yy in range(ny):
                for xx in range(nx):
                    s = b[co]
                    for t in range(27):
                        dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
                        if 0 <= xx + dx < nx and 0 <= yy + dy < ny and 0 <= z + dz < nz:
                            s += W[co, :, t] @ x[:, z + dz, yy + dy, xx + dx]
                    ref[co, z, yy, xx] = s
    assert np.allclose(out, ref, atol=1e-12)


def test_backward_matches_finite_differences():
    Tnm1, Tn, q, z, Tnp1 = sample()
    w = init()
    loss, g = N.loss_and_gradients(w, Tnm1, Tn, q, z, Tnp1, bf=False)
    rng = np.random.default_rng(3)
    idx = rng.choice(w.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fp = N.loss_and_gradients(wp, Tnm1, Tn, q, z, Tnp1, bf=False)[0]
        fm = N.loss_and_gradients(wm, Tnm1, Tn, q, z, Tnp1, bf=False)[0]
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max())


def test_prediction_is_linear_extrapolation_plus_output():
    Tnm1, Tn, q, z, _ = sample()
    w = init()
    w[-1] = 0.0  # last bias
    nlast = N.COUT[3] * N.CIN[3] * 27
    w[-1 - nlast:-1] = 0.0  # last weights: output 0
    p = N.predict(w, Tnm1, Tn, q, z, bf=False)
    assert np.allclose(p, 2 * Tn - Tnm1, atol=1e-12)
    # a bias b on the output adds b rms(z) everywhere
    w[-1] = 0.5
    p = N.predict(w, Tnm1, Tn, q, z, bf=False)
    assert np.allclose(p - (2 * Tn - Tnm1), 0.5 * np.sqrt(np.mean(z * z)), atol=1e-12)


def test_features_channels():
    """Each channel is its normalised input; without z the output scale is rms(T_n - T_{n-1})."""
    Tnm1, Tn, q, z, _ = sample()
    c, sz = N.features(Tnm1, Tn, q, z, 0.5)
    d = Tn - Tnm1
    s = np.sqrt(np.mean(d * d))
    assert np.isclose(sz, np.sqrt(np.mean(z * z)))
    assert np.allclose(c[0] * s, d) and np.allclose(c[1] * s, 0.5 * q) and np.allclose(c[3] * sz, z)
    # c2 is T_n standardised; its scale carries a 1e-6 |mean| guard against constant fields
    assert abs(np.mean(c[2])) < 1e-12
    assert np.isclose(np.std(c[2]) * (np.std(Tn) + 1e-12 + 1e-6 * abs(np.mean(Tn))), np.std(Tn), rtol=1e-12)
    c, s2 = N.features(Tnm1, Tn, None, None, 0.5)
    assert np.isclose(s2, s) and not c[1].any() and not c[3].any()


def test_prolongation_is_exact_for_coarse_polynomials():
    """P carries the coarse nodal values of a polynomial of degree <= k to its fine nodal values."""
    for k in (1, 2, 3):
        for rho in (2, 4):
            P, W = N.transfer_1d(k, 8, rho)
            nodes = N.gl_nodes(k)
            xf = np.array([(min(g // k, 7) + nodes[g - k * min(g // k, 7)]) / 8 for g in range(8 * k + 1)])
            ncc = 8 // rho
            xc = np.array([(min(g // k, ncc - 1) + nodes[g - k * min(g // k, ncc - 1)]) / ncc
                           for g in range(ncc * k + 1)])
            for p in (lambda x: 1.0 + 0 * x, lambda x: 2 * x - 1, lambda x: (x - 0.3) ** k):
                assert np.allclose(P @ p(xc), p(xf), atol=1e-13)
            assert np.allclose(W @ np.ones(len(xf)), 1.0, atol=1e-14)
            assert np.allclose(W.sum(axis=1), 1.0)


def test_backward_with_coarsening_matches_finite_differences():
    """Q2 on 4^3 cells, network on 2^3 cells (5^3 nodes)."""
    Tnm1, Tn, q, z, Tnp1 = sample((9, 9, 9), seed=5)
    tr = N.transfers(2, (4, 4, 4), 2)
    w = init(2)
    loss, g = N.loss_and_gradients(w, Tnm1, Tn, q, z, Tnp1, bf=False, tr=tr)
    rng = np.random.default_rng(4)
    h = 1e-6
    for i in rng.choice(w.size, 25, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (N.loss_and_gradients(wp, Tnm1, Tn, q, z, Tnp1, bf=False, tr=tr)[0] -
              N.loss_and_gradients(wm, Tnm1, Tn, q, z, Tnp1, bf=False, tr=tr)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max())
    # with a zero network the guess is the linear extrapolation on the fine grid
    w0 = np.zeros_like(w)
    assert np.allclose(N.predict(w0, Tnm1, Tn, q, z, bf=False, tr=tr), 2 * Tn - Tnm1, atol=1e-12)
