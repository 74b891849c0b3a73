"""The guiding distribution without a GPU: the FP64 oracle pinned by finite
differences and closed forms, and the device math of guide.h (compiled for
the host) against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import guide_oracle as go


def random_params(rng, scale=1.0):
    p = np.zeros(go.N_PARAMS)
    for l in range(go.LAYERS):
        k = go.layer_in(l)
        w = rng.uniform(-1, 1, (k, 64)) * np.sqrt(6 / k) * scale
        if l == go.LAYERS - 1:
            w[:, go.N_OUT:] = 0
        p[go.w_offset(l):go.w_offset(l) + k * 64] = w.ravel()
        p[go.b_offset(l):go.b_offset(l) + 64] = rng.uniform(-0.1, 0.1, 64)
    p[go.b_offset(go.LAYERS - 1) + go.N_OUT:go.b_offset(go.LAYERS - 1) + 64] = 0
    return p


def random_dirs(rng, n):
    w = rng.normal(size=(n, 3))
    return w / np.linalg.norm(w, axis=1, keepdims=True)


def random_samples(rng, n):
    s = np.zeros((n, 10))
    s[:, :6] = rng.random((n, 6))
    s[:, 6:9] = random_dirs(rng, n)
    s[:, 9] = rng.uniform(0.1, 2.0, n)
    return s


def test_oracle_gradient_matches_finite_differences():
    """Central differences of the float64 loss against the analytic
    backpropagation, on parameters of every layer (weights and biases)."""
    rng = np.random.default_rng(3)
    p = random_params(rng)
    s = random_samples(rng, 24)
    g, _ = go.gradient(p, s)
    idx = []
    for l in range(go.LAYERS):
        k = go.layer_in(l)
        cols = go.N_OUT if l == go.LAYERS - 1 else 64
        for _ in range(12):
            idx.append(go.w_offset(l) + rng.integers(k) * 64 + rng.integers(cols))
        for _ in range(4):
            idx.append(go.b_offset(l) + rng.integers(cols))
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = go.loss(q, s)
        q[i] -= 2 * h
        lm = go.loss(q, s)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_oracle_encoding_matches_per_element_definition():
    """Input 16 s + 2 b is sin(2^b pi x_s), input 16 s + 2 b + 1 its cosine
    (the encoding the renderer's network reads), element by element with
    the math module."""
    import math
    f = np.random.default_rng(21).random((7, 6))
    got = go.encode(f)
    for r in range(7):
        for sc in range(6):
            for b in range(8):
                x = f[r, sc] * 2.0 ** b * math.pi
                assert got[r, 16 * sc + 2 * b] == pytest.approx(math.sin(x), abs=1e-14)
                assert got[r, 16 * sc + 2 * b + 1] == pytest.approx(math.cos(x), abs=1e-14)


def test_oracle_bf16_rounding_matches_torch():
    """Round to nearest even at bfloat16, ties, subnormals, infinities and
    values that round up into the next binade, against torch's conversion."""
    import torch
    rng = np.random.default_rng(22)
    x = np.concatenate([rng.normal(0, 10, 4000), rng.normal(0, 1e-30, 200), [0.0, -0.0, 1.0, -1.0, 3e38, -3e38,
                                                                             np.inf, -np.inf, 1e-40, 1.00390625,
                                                                             1.01171875, 255.5, 256.5]])
    x32 = x.astype(np.float32)
    # exact ties between two bfloat16 values: low 16 bits = 0x8000
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x32 = np.concatenate([x32, ties, -ties])
    want = torch.from_numpy(x32).to(torch.bfloat16).to(torch.float64).numpy()
    got = go.bf16(x32)
    assert np.array_equal(got, want)


def torch_net(p):
    """The network as a torch float64 Linear/ReLU stack with the oracle's
    parameter layout (W_l [K, 64] row-major then b_l)."""
    import torch
    layers = []
    for l in range(go.LAYERS):
        k = go.layer_in(l)
        lin = torch.nn.Linear(k, 64, dtype=torch.float64)
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(p[go.w_offset(l):go.w_offset(l) + k * 64].reshape(k, 64).T.copy()))
            lin.bias.copy_(torch.from_numpy(p[go.b_offset(l):go.b_offset(l) + 64].copy()))
        layers.append(lin)
        if l < go.LAYERS - 1:
            layers.append(torch.nn.ReLU())
    return torch.nn.Sequential(*layers)


def torch_encode(f):
    import torch
    f = torch.as_tensor(f, dtype=torch.float64)
    cols = []
    for sc in range(6):
        for b in range(8):
            x = f[:, sc] * (2.0 ** b) * np.pi
            cols += [torch.sin(x), torch.cos(x)]
    return torch.stack(cols, 1)


def test_oracle_forward_matches_torch_linear_stack():
    """go.forward against an independent torch float64 evaluation with the
    same weights: 1e-12 relative."""
    import torch
    rng = np.random.default_rng(23)
    p = random_params(rng)
    f = rng.random((300, 6))
    got, acts = go.forward(p, f)
    with torch.no_grad():
        want = torch_net(p)(torch_encode(f)).numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.abs(want).max()
    assert len(acts) == go.LAYERS and acts[0].shape == (300, 96)


def test_oracle_bf16_forward_matches_torch_bfloat16_storage():
    """With bf16=True the oracle rounds the input, the weights and the hidden
    activations to bfloat16 and accumulates in float64: the same with torch
    tensors cast through torch.bfloat16 at those points."""
    import torch
    rng = np.random.default_rng(24)
    p = random_params(rng)
    f = rng.random((200, 6))

    def rnd(t):
        return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)

    with torch.no_grad():
        h = rnd(torch_encode(f))
        for l in range(go.LAYERS):
            k = go.layer_in(l)
            w = rnd(torch.from_numpy(p[go.w_offset(l):go.w_offset(l) + k * 64].reshape(k, 64).copy()))
            z = h @ w + torch.from_numpy(p[go.b_offset(l):go.b_offset(l) + 64].copy())
            h = rnd(torch.relu(z)) if l < go.LAYERS - 1 else z
    got, _ = go.forward(p, f, bf=True)
    assert np.max(np.abs(got - h.numpy())) <= 1e-12 * np.abs(h.numpy()).max()


def torch_loss(p_t, samples):
    """Summed -c log q(w) with q the vMF mixture written in the textbook form
    C(kappa) = kappa / (4 pi sinh kappa), through torch autograd."""
    import torch
    s = torch.as_tensor(samples, dtype=torch.float64)
    h = torch_encode(s[:, :6].numpy())
    for l in range(go.LAYERS):
        k = go.layer_in(l)
        w = p_t[go.w_offset(l):go.w_offset(l) + k * 64].reshape(k, 64)
        h = h @ w + p_t[go.b_offset(l):go.b_offset(l) + 64]
        if l < go.LAYERS - 1:
            h = torch.relu(h)
    q = h[:, :go.N_OUT].reshape(-1, go.LOBES, 5)
    m = q[:, :, :3]
    mu = m / torch.clamp(torch.linalg.norm(m, dim=2, keepdim=True), min=1e-6)
    kappa = torch.clamp(torch.exp(torch.clamp(q[:, :, 3], max=80.0)), go.KAPPA_MIN, go.KAPPA_MAX)
    wgt = torch.softmax(q[:, :, 4], dim=1)
    d = s[:, 6:9]
    comp = kappa / (4 * np.pi * torch.sinh(kappa)) * torch.exp(kappa * (mu * d[:, None, :]).sum(2))
    return -(s[:, 9] * torch.log((wgt * comp).sum(1))).sum()


def test_oracle_gradient_matches_torch_autograd():
    """go.gradient (hand-derived loss gradient and backpropagation) against
    torch autograd of an independently written loss, 1e-9 relative per
    layer; and the summed loss."""
    import torch
    rng = np.random.default_rng(25)
    p = random_params(rng)
    s = random_samples(rng, 64)
    g, lval = go.gradient(p, s)
    p_t = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    lt = torch_loss(p_t, s)
    lt.backward()
    gt = p_t.grad.numpy()
    assert lval == pytest.approx(lt.item(), rel=1e-10)
    for l in range(go.LAYERS):
        for off, size in ((go.w_offset(l), go.layer_in(l) * 64), (go.b_offset(l), 64)):
            a, b = g[off:off + size], gt[off:off + size]
            assert np.max(np.abs(a - b)) <= 1e-9 * max(np.abs(b).max(), 1e-12), (l, off)


def test_oracle_adam_first_step_closed_form():
    """At step 1 the bias-corrected moments are g and g^2, so the update is
    exactly -lr g / (|g| + eps) for every element (0 for g = 0, -lr/11 for
    g = 1e-9 and eps = 1e-8)."""
    g = np.array([2.0, -3.0, 0.0, 1e-9, -1e-9, 1e3])
    p0 = np.array([0.5, 0.25, -1.0, 0.0, 0.0, 7.0])
    lr = 1e-3
    p1, m1, v1 = go.adam(p0, np.zeros(6), np.zeros(6), g, 1, lr=lr)
    np.testing.assert_allclose(p1 - p0, -lr * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-18)
    assert p1[3] - p0[3] == pytest.approx(-lr / 11, rel=1e-12)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-15)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=1e-15)


def test_oracle_adam_second_step_by_hand():
    """g1 = 2 then g2 = -1 (beta1 0.9, beta2 0.999): m2 = 0.09*2 - 0.1 =
    0.08, v2 = 0.000999*4 + 0.001 = 0.004996; bias corrections 1 - 0.81 =
    0.19 and 1 - 0.998001 = 0.001999, so the second update is
    -lr (8/19) / (sqrt(4996/1999) + eps)."""
    lr = 1e-3
    p1, m1, v1 = go.adam(np.array([1.0]), np.zeros(1), np.zeros(1), np.array([2.0]), 1, lr=lr)
    p2, m2, v2 = go.adam(p1, m1, v1, np.array([-1.0]), 2, lr=lr)
    assert m2[0] == pytest.approx(0.08, rel=1e-13)
    assert v2[0] == pytest.approx(0.004996, rel=1e-13)
    assert p2[0] - p1[0] == pytest.approx(-lr * (8 / 19) / (np.sqrt(4996 / 1999) + 1e-8), rel=1e-12)
    assert p1[0] == pytest.approx(1.0 - lr * 2 / (2 + 1e-8), rel=1e-15)


def test_padding_outputs_have_zero_gradient():
    rng = np.random.default_rng(4)
    p = random_params(rng)
    g, _ = go.gradient(p, random_samples(rng, 16))
    w4 = g[go.w_offset(4):go.w_offset(4) + 64 * 64].reshape(64, 64)
    assert np.all(w4[:, go.N_OUT:] == 0)
    assert np.all(g[go.b_offset(4) + go.N_OUT:go.b_offset(4) + 64] == 0)


def sphere_grid(n_theta=720, n_phi=1440):
    """Midpoint quadrature nodes and weights on the unit sphere."""
    th = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    return d.reshape(-1, 3), w.ravel()


@pytest.mark.parametrize("kscale", [-8.0, 0.0, 3.0, 6.0])
def test_oracle_mixture_integrates_to_one(kscale):
    """The vMF mixture is a density: its integral over the sphere is 1 for
    small (kappa -> 1e-3), moderate and sharp (kappa ~ 400) lobes."""
    rng = np.random.default_rng(5)
    o = rng.normal(size=(1, go.N_OUT))
    o[0, 3::5] = kscale + rng.normal(scale=0.2, size=8)
    d, w = sphere_grid()
    q = go.pdf(np.repeat(o, len(d), 0), d)
    assert abs((q * w).sum() - 1) < 2e-3


def test_oracle_small_kappa_is_uniform():
    o = np.zeros((3, go.N_OUT))
    o[:, 3::5] = -30  # exp -> below the clamp: kappa = 1e-3
    o[:, 0::5] = 1
    q = go.pdf(o, random_dirs(np.random.default_rng(0), 3))
    np.testing.assert_allclose(q, 1 / (4 * np.pi), rtol=2e-3)


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("guide") / "guide_host"
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "lambda_amd", "csrc"),
           os.path.join(ROOT, "tests", "host_math", "guide_host.hip"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return str(out)


def run(driver, tmp_path, args, arrays, out_shape):
    paths = []
    for k, a in enumerate(arrays):
        p = tmp_path / ("in%d.f32" % k)
        np.ascontiguousarray(a, np.float32).tofile(p)
        paths.append(str(p))
    op = tmp_path / "out.f32"
    subprocess.run([driver, args[0], *paths, *args[1:], str(op)], check=True)
    return np.fromfile(op, np.float32).reshape(out_shape)


def test_device_encoding_matches_oracle(driver, tmp_path):
    f = np.random.default_rng(1).random((500, 6)).astype(np.float32)
    got = run(driver, tmp_path, ["encode"], [f], (500, 96))
    # one sincos and 7 double-angle steps carry ~1e-5 of rounding; the
    # network stores its input in bfloat16 (8 significant bits) anyway
    np.testing.assert_allclose(got, go.encode(f), atol=1e-4)
    # by halves (the inference kernel): bands 4..7 start from their own sincos
    got2 = run(driver, tmp_path, ["encode2"], [f], (500, 96))
    np.testing.assert_allclose(got2, go.encode(f), atol=1e-4)
    assert np.array_equal(got2[:, :8], got[:, :8])


def test_device_pdf_and_gradient_match_oracle(driver, tmp_path):
    rng = np.random.default_rng(2)
    n = 400
    o = rng.normal(size=(n, go.N_OUT)).astype(np.float32)
    o[:, 3::5] = rng.uniform(-4, 6, (n, 8))
    w = random_dirs(rng, n).astype(np.float32)
    c = rng.uniform(0.1, 3, n).astype(np.float32)
    q = run(driver, tmp_path, ["pdf"], [o, w], (n,))
    np.testing.assert_allclose(q, go.pdf(o.astype(np.float64), w.astype(np.float64)), rtol=2e-4)
    # the compact form the renderer shades with
    qc = run(driver, tmp_path, ["pdfc"], [o, w], (n,))
    np.testing.assert_allclose(qc, go.pdf(o.astype(np.float64), w.astype(np.float64)), rtol=2e-4)
    got = run(driver, tmp_path, ["grad"], [o, w, c], (n, 41))
    lref, dref = go.loss_grad(o.astype(np.float64), w.astype(np.float64), c.astype(np.float64))
    np.testing.assert_allclose(got[:, 0], lref, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(got[:, 1:], dref, rtol=2e-3, atol=2e-3 * np.abs(dref).max())


def test_compact_pdf_at_high_concentration(driver, tmp_path):
    """kappa up to the 2000 clamp: the compact form's exponent c + kappa mu.w
    cancels as the lobe form's kappa (mu.w - 1) does; directions within a few
    lobe widths of each mean."""
    rng = np.random.default_rng(3)
    n = 400
    o = rng.normal(size=(n, go.N_OUT)).astype(np.float32)
    o[:, 3::5] = rng.uniform(4, 8, (n, 8))          # kappa 55 .. 2000 (clamped)
    k = rng.integers(0, 8, n)
    mu = o[np.arange(n)[:, None], 5 * k[:, None] + np.arange(3)].astype(np.float64)
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    w = mu + rng.normal(scale=0.02, size=(n, 3))
    w = (w / np.linalg.norm(w, axis=1, keepdims=True)).astype(np.float32)
    ref = go.pdf(o.astype(np.float64), w.astype(np.float64))
    ok = ref > 1e-3
    assert ok.sum() > n // 2
    for mode in ("pdf", "pdfc"):
        q = run(driver, tmp_path, [mode], [o, w], (n,))
        np.testing.assert_allclose(q[ok], ref[ok], rtol=2e-3)


@pytest.mark.parametrize("mode", ["sample", "samplec"])
def test_device_sampling_follows_pdf(driver, tmp_path, mode):
    """Histogram of 400k mixture samples over 12 x 24 sphere cells against
    the cells' integrated density (oracle quadrature); the lobe form and the
    compact form the renderer samples with."""
    rng = np.random.default_rng(6)
    o = rng.normal(size=(1, go.N_OUT)).astype(np.float32)
    o[0, 3::5] = rng.uniform(0, 4, 8)
    count = 400000
    d = run(driver, tmp_path, [mode, str(count), "11"], [o], (count, 3)).astype(np.float64)
    assert np.all(np.abs(np.linalg.norm(d, axis=1) - 1) < 1e-5)
    nt, nphi = 12, 24
    th = np.arccos(np.clip(d[:, 1], -1, 1))
    ph = np.mod(np.arctan2(d[:, 2], d[:, 0]), 2 * np.pi)
    hist = np.histogram2d(th, ph, bins=[nt, nphi], range=[[0, np.pi], [0, 2 * np.pi]])[0] / count
    g, wq = sphere_grid(nt * 40, nphi * 40)
    q = go.pdf(np.repeat(o.astype(np.float64), len(g), 0), g) * wq
    cell = q.reshape(nt, 40, nphi, 40).sum(axis=(1, 3))
    assert abs(cell.sum() - 1) < 1e-3
    sigma = np.sqrt(cell * (1 - cell) / count) + 1e-7
    z = np.abs(hist - cell) / sigma
    assert z.max() < 6, z.max()


def test_table_entries_cells_and_normal_bins(driver, tmp_path):
    """guide_table_entry / guide_table_query: a point's cell is the floor of
    its offset over the cell size, clamped into the grid (NaN to cell 0);
    the normal bin is 2 x the axis of the largest |component| (ties to the
    lower axis) + 1 for a negative component; an entry's centre lies in its
    cell, its normal is the bin's axis direction, and looking the centre up
    with that normal gives the entry back."""
    lo = np.array([-1.0, 0.5, 2.0], np.float32)
    inv_cell, dims = np.float32(4.0), np.array([7, 3, 5])
    geom = np.concatenate([lo, [inv_cell], dims]).astype(np.float32)
    rng = np.random.default_rng(7)
    m = 4000
    p = lo + rng.uniform(-0.3, 1.3, (m, 3)).astype(np.float32) * dims / inv_cell
    nrm = rng.normal(size=(m, 3)).astype(np.float32)
    nrm[:8] = [[1, 1, 0], [0, -1, -1], [1, -1, 1], [0, 0, -2], [-3, 0, 0], [0, 0.5, 0], [1, 1, 1], [-1, -1, 0]]
    p[8] = [np.nan, 0.7, 2.1]
    got = run(driver, tmp_path, ["table"], [geom, np.concatenate([p, nrm], axis=1)], (m, 7))
    with np.errstate(invalid="ignore"):
        f = np.floor((p - lo) * inv_cell)
    cell = np.where(np.isnan(f), 0, np.clip(f, 0, dims - 1)).astype(np.int64)
    a = np.abs(nrm)
    axis = np.argmax(a, axis=1)  # first maximum: ties to the lower axis
    neg = nrm[np.arange(m), axis] < 0
    want = ((cell[:, 2] * dims[1] + cell[:, 1]) * dims[0] + cell[:, 0]) * 6 + 2 * axis + neg
    np.testing.assert_array_equal(got[:, 0].astype(np.int64), want)
    assert list(got[:8, 0].astype(np.int64) % 6) == [0, 3, 0, 5, 1, 2, 0, 1]
    # centres inside their cells, normals on their axes
    centre, axis_dir = got[:, 1:4], got[:, 4:7]
    np.testing.assert_array_equal(np.floor((centre - lo) * inv_cell).astype(np.int64), cell)
    sign = np.where(neg, -1.0, 1.0)
    np.testing.assert_array_equal(axis_dir, np.eye(3)[axis] * sign[:, None])
    # round trip
    back = run(driver, tmp_path, ["table"], [geom, got[:, 1:7]], (m, 7))
    np.testing.assert_array_equal(back[:, 0], got[:, 0])
