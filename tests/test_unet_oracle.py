"""Pins of the network oracle (oracle/unet_oracle.py), CPU only: the analytic gradient against FP64 finite
differences, the convolution's tap convention against a delta input, adjoint identities of the transposed passes,
BF16 rounding against torch, and the parameter layout against the library's own count."""
import numpy as np
import pytest

import unet_oracle as U


def small_domain(seed=1, n=16):
    rng = np.random.default_rng(seed)
    ctype = np.zeros((n, n, n), np.uint32)
    ctype[: n // 3] = 1
    for sl in (np.s_[0], np.s_[-1], np.s_[:, 0], np.s_[:, -1], np.s_[:, :, 0], np.s_[:, :, -1]):
        ctype[sl] = 2
    alloc = np.ones((n, n, n), bool)
    alloc[8:, 8:, 8:] = False
    diag = rng.uniform(4.0, 6.0, (n, n, n)).astype(np.float32)
    dom = U.Domain(ctype, alloc, diag)
    r = (rng.standard_normal((n, n, n)) * dom.fluid).astype(np.float32)
    return dom, r


def test_gradient_matches_finite_differences():
    dom, r = small_domain()
    p = U.init_params(3).astype(np.float64)
    offs, n = U.layer_offsets()
    p[offs[10][1]] = 0.1
    _, g = U.loss_and_grad(p, dom, r, exact=True)
    idx = [offs[k][0] + 7 for k in range(len(offs))] + [offs[k][1] for k in (0, 5, 9, 10)]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (U.loss_and_grad(pp, dom, r, exact=True)[0] - U.loss_and_grad(pm, dom, r, exact=True)[0]) / (2 * eps)
        assert abs(g[i] - fd) <= 1e-3 * max(abs(fd), 1e-6), (i, g[i], fd)


def test_conv_tap_convention_on_a_delta():
    rng = np.random.default_rng(0)
    w = rng.standard_normal((27, 2, 3))
    x = np.zeros((3, 6, 6, 6))
    x[1, 3, 3, 3] = 1.0
    out = U.conv3(x, w, np.zeros(2), np.ones((6, 6, 6)))
    for t, (dx, dy, dz) in enumerate(U.TAPS):
        # out[v] = sum_t W[t] x[v + d_t]: the delta at v0 shows up at v0 - d_t
        assert np.allclose(out[:, 3 - dz, 3 - dy, 3 - dx], w[t][:, 1])


def test_adjoint_identities():
    rng = np.random.default_rng(2)
    m = (rng.uniform(size=(8, 8, 8)) > 0.3).astype(np.float64)
    x = rng.standard_normal((4, 8, 8, 8)) * m
    g = rng.standard_normal((5, 8, 8, 8)) * m
    w = rng.standard_normal((27, 5, 4))
    lhs = (U.conv3(x, w, np.zeros(5), m) * g).sum()
    rhs = (x * U.conv3_t(g, w, m)).sum()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    dw, _ = U.conv3_dw(x, g)
    assert abs((dw * w).sum() - lhs) <= 1e-10 * abs(lhs)
    c = rng.standard_normal((3, 4, 4, 4))
    f = rng.standard_normal((3, 8, 8, 8))
    assert abs((U.up(c) * f).sum() - (c * U.down_sum(f)).sum()) <= 1e-10 * abs((U.up(c) * f).sum())
    assert np.allclose(U.pool(f), U.down_sum(f) / 8.0)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.standard_normal(100000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 100000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0], np.float32)])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(U.bf16(a).view(np.uint32), ref.view(np.uint32))


def test_parameter_layout_matches_library():
    from slosh_amd import load_library, unet

    assert unet.param_count() == U.layer_offsets()[1] == load_library().slosh_unet_param_count()
    assert [l[1:3] for l in unet.LAYERS] == [l[1:3] for l in U.LAYERS]
