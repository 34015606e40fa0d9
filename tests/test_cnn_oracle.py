"""Pins the CPU oracle of the learned-guess network (oracle/cnn_oracle.py)
by what the mathematics fixes: the convolution against a brute-force loop,
the backward pass against central finite differences (float64), gelu' against
a difference quotient, bfloat16 rounding against torch's conversion."""
import numpy as np
import pytest

from oracle import cnn_oracle as N


def test_conv_matches_bruteforce():
    rng = np.random.default_rng(1)
    nz, ny, nx, ci, co = 3, 4, 5, 2, 3
    x = rng.standard_normal((nz, ny, nx, ci))
    W = rng.standard_normal((27, ci, co))
    b = rng.standard_normal(co)
    y = N.conv(x, W, b)
    ref = np.zeros_like(y)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                s = b.copy()
                for t, (dk, dj, di) in enumerate(N.TAPS):
                    kk, jj, ii = k + dk, (j + dj) % ny, i + di
                    if 0 <= kk < nz and 0 <= ii < nx:     # zero padding in theta and r, periodic in phi
                        s += x[kk, jj, ii] @ W[t]
                ref[k, j, i] = s
    np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_conv_t_is_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 6, 5, 3))
    g = rng.standard_normal((4, 6, 5, 2))
    W = rng.standard_normal((27, 3, 2))
    lhs = np.sum(g * (N.conv(x, W, np.zeros(2))))
    rhs = np.sum(x * N.conv_t(g, W))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gelu_derivative():
    u = np.linspace(-6, 6, 1001)
    h = 1e-6
    fd = (N.gelu(u + h) - N.gelu(u - h)) / (2 * h)
    np.testing.assert_allclose(N.gelu_d(u), fd, rtol=1e-7, atol=1e-9)
    # closed-form anchors: gelu(0) = 0, gelu(u) -> u for large u, -> 0 for very negative u
    assert N.gelu(np.array([0.0]))[0] == 0.0
    assert abs(N.gelu(np.array([10.0]))[0] - 10.0) < 1e-12
    assert abs(N.gelu(np.array([-10.0]))[0]) < 1e-12


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-8, 8, 10000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 65504.0, 1e-30]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(N.bf16(v), ref)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    nz, ny, nx = 3, 4, 5
    x0 = rng.standard_normal((nz, ny, nx, 5))
    target = rng.standard_normal((nz, ny, nx))
    w, b = N.init_params(7)
    w = w.astype(np.float64) * 3.0        # larger last layer so every layer matters
    b = rng.standard_normal(b.size) * 0.1
    loss, gw, gb, _ = N.loss_grad(x0, target, w, b, bf=False)
    h = 1e-6
    idx = np.concatenate([rng.choice(w.size, 80, replace=False),
                          np.arange(w.size - 27 * 16, w.size - 27 * 16 + 10)])
    for q in idx:
        wp, wm = w.copy(), w.copy()
        wp[q] += h
        wm[q] -= h
        fd = (N.loss_grad(x0, target, wp, b, bf=False)[0] - N.loss_grad(x0, target, wm, b, bf=False)[0]) / (2 * h)
        assert abs(fd - gw[q]) <= 1e-6 * max(1.0, abs(gw[q])) + 1e-8, (q, fd, gw[q])
    for q in range(b.size):
        bp, bm = b.copy(), b.copy()
        bp[q] += h
        bm[q] -= h
        fd = (N.loss_grad(x0, target, w, bp, bf=False)[0] - N.loss_grad(x0, target, w, bm, bf=False)[0]) / (2 * h)
        assert abs(fd - gb[q]) <= 1e-6 * max(1.0, abs(gb[q])) + 1e-8, (q, fd, gb[q])


def test_parameter_counts():
    # 5->16, three 16->16, 16->1 with 3x3x3 kernels: 23,328 weights, 65 biases
    assert N.n_weights() == 27 * (5 * 16 + 3 * 16 * 16 + 16) == 23328
    assert N.n_biases() == 65
