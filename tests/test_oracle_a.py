"""Oracle A (the FP64 restatement of the reference's src/train/reference.cpp)
pinned without running the reference:

* analytic gradients against central finite differences, as the reference pins
  its own (tests/test_train.cpp:86-182), with its bars (cosine > 0.999, length
  within 1%);
* the loss adjoint against finite differences of the loss (test_train.cpp:115-142);
* closed forms: compositing of 1-3 camera-facing isotropic Gaussians, whose
  projection and alpha are known exactly; SSIM against an independent
  convolution formulation with the same renormalised border window; SSIM(x, x) = 1;
  the MCMC penalties (test_train.cpp:295-321); SH orthonormality on the sphere;
  rotation matrices orthonormal with determinant 1.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O
from vsplat import fixtures as F

WEIGHTS = np.array([0.00102838, 0.00759876, 0.03600077, 0.10936069, 0.21300554, 0.26601172, 0.21300554, 0.10936069,
                    0.03600077, 0.00759876, 0.00102838])


def check_close(a, b, what, min_cos, max_ratio):
    """tests/test_train.cpp:38-53."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    assert (a @ a) > 0 and (b @ b) > 0, what
    cos = (a @ b) / math.sqrt((a @ a) * (b @ b))
    ratio = math.sqrt((a @ a) / (b @ b))
    assert cos > min_cos, f"{what}: cosine {cos}"
    assert abs(ratio - 1) < max_ratio, f"{what}: ratio {ratio}"


def functional(cloud, cam, adjoint, options):
    img = O.render(cloud, cam, options, threads=1)
    return float((adjoint * img.rgb).sum())


def finite_differences(cloud, member, cam, adjoint, step, options):
    """test_train.cpp:66-82: central differences over one parameter array, in float."""
    c = cloud.copy()
    values = getattr(c, member).reshape(-1)
    out = np.zeros(values.size)
    for i in range(values.size):
        original = values[i]
        values[i] = np.float32(original + np.float32(step))
        up = values[i]
        high = functional(c, cam, adjoint, options)
        values[i] = np.float32(original - np.float32(step))
        down = values[i]
        low = functional(c, cam, adjoint, options)
        values[i] = original
        out[i] = (high - low) / (float(up) - float(down))
    return out


def _adjoint(cam, seed):
    return np.random.default_rng(seed).standard_normal((cam.width * cam.height, 3))


SMOOTH = O.RefOptions(no_cutoffs=True)


@pytest.mark.parametrize("member", ["positions", "scales", "rotations", "opacities"])
def test_gradients_match_finite_differences_degree0(member):
    cam = F.train_camera(40, 32)
    adjoint = _adjoint(cam, 9)
    cloud = F.random_cloud(12, 0, 3)
    g = O.gradients(cloud, cam, adjoint, SMOOTH, threads=1)
    check_close(getattr(g, member), finite_differences(cloud, member, cam, adjoint, 2e-3, SMOOTH), member, 0.999, 0.01)


@pytest.mark.parametrize("degree,seed,member", [(3, 6, "positions"), (3, 6, "sh"), (2, 4, "sh")])
def test_gradients_match_finite_differences_sh(degree, seed, member):
    """The view direction's term in the position gradient and the SH gradients (test_train.cpp:103-112)."""
    cam = F.train_camera(40, 32)
    adjoint = _adjoint(cam, 9)
    cloud = F.random_cloud(12, degree, seed)
    g = O.gradients(cloud, cam, adjoint, SMOOTH, threads=1)
    check_close(getattr(g, member), finite_differences(cloud, member, cam, adjoint, 2e-3, SMOOTH), member, 0.999, 0.01)


@pytest.mark.parametrize("filter3d", [False, True])
def test_mip_filter_gradients_match_finite_differences(filter3d):
    """test_train.cpp:144-182: the 2D Mip filter, and with it the 3D smoothing filter."""
    cam = F.train_camera(40, 32)
    adjoint = _adjoint(cam, 21)
    cloud = F.random_cloud(12, 0, 3)
    f = (0.002 * np.arange(1, 13)).astype(np.float32) if filter3d else None
    opts = O.RefOptions(no_cutoffs=True, mip_filter_2d=True, filter3d=f)
    g = O.gradients(cloud, cam, adjoint, opts, threads=1)
    for member in ("positions", "scales", "rotations", "opacities"):
        check_close(getattr(g, member), finite_differences(cloud, member, cam, adjoint, 2e-3, opts), member, 0.999, 0.01)
    plain = O.RefOptions(no_cutoffs=True, mip_filter_2d=filter3d)
    assert functional(cloud, cam, adjoint, opts) != functional(cloud, cam, adjoint, plain)


@pytest.mark.parametrize("lam", [1.0, 0.2])
def test_loss_adjoint_matches_finite_differences(lam):
    """test_train.cpp:115-142."""
    w, h = 18, 14
    rng = np.random.default_rng(21)
    rgb = rng.uniform(0.1, 0.9, (w * h, 3))
    target = F.smooth_target(w, h, 5)
    base = O.loss(O.RefImage(w, h, rgb, np.zeros(w * h), np.zeros(w * h, np.uint64)), target, lam, threads=1)
    analytic, numeric = [], []
    for i in range(0, w * h, 7):
        for c in range(3):
            step = 1e-5
            up = rgb.copy()
            up[i, c] += step
            down = rgb.copy()
            down[i, c] -= step
            hi = O.loss(O.RefImage(w, h, up, np.zeros(w * h), np.zeros(w * h, np.uint64)), target, lam, threads=1).loss
            lo = O.loss(O.RefImage(w, h, down, np.zeros(w * h), np.zeros(w * h, np.uint64)), target, lam, threads=1).loss
            numeric.append((hi - lo) / (2 * step))
            analytic.append(base.adjoint[i, c])
    check_close(analytic, numeric, f"lambda {lam}", 0.9999, 0.001)


def _independent_ssim(x, y):
    """SSIM per pixel and channel by 2D correlation with the 11x11 window cut at
    the border and renormalised by the correlation of ones (the same window
    the reference uses, written as convolutions rather than loops)."""
    win = np.outer(WEIGHTS, WEIGHTS)
    norm = ndimage.correlate(np.ones(x.shape[:2]), win, mode="constant")
    out = np.zeros(x.shape)
    for c in range(3):
        m = lambda a: ndimage.correlate(a, win, mode="constant") / norm  # noqa: E731
        mx, my = m(x[..., c]), m(y[..., c])
        vx, vy, cxy = m(x[..., c] ** 2) - mx ** 2, m(y[..., c] ** 2) - my ** 2, m(x[..., c] * y[..., c]) - mx * my
        out[..., c] = ((2 * mx * my + 1e-4) * (2 * cxy + 9e-4)) / ((mx ** 2 + my ** 2 + 1e-4) * (vx + vy + 9e-4))
    return out


def test_ssim_and_l1_against_independent_formulation():
    w, h = 37, 29
    rng = np.random.default_rng(4)
    rgb = rng.uniform(0, 1, (w * h, 3))
    target = F.smooth_target(w, h, 8)
    out = O.loss(O.RefImage(w, h, rgb, np.zeros(w * h), np.zeros(w * h, np.uint64)), target, 0.2, threads=3)
    y = target[..., :3].astype(np.float64) / 255.0
    ssim = _independent_ssim(rgb.reshape(h, w, 3), y).mean()
    l1 = np.abs(rgb.reshape(h, w, 3) - y).mean()
    assert abs(out.ssim - ssim) < 1e-12 and abs(out.l1 - l1) < 1e-12
    assert abs(out.loss - (0.8 * l1 + 0.2 * (1 - ssim))) < 1e-12


def test_ssim_of_an_image_with_itself_is_one():
    w, h = 20, 16
    target = F.smooth_target(w, h, 3)
    rgb = target[..., :3].reshape(-1, 3).astype(np.float64) / 255.0
    out = O.loss(O.RefImage(w, h, rgb, np.zeros(w * h), np.zeros(w * h, np.uint64)), target, 0.5, threads=1)
    assert abs(out.ssim - 1) < 1e-12 and out.l1 == 0 and abs(out.loss) < 1e-12


def _facing_gaussian(cloud, i, x, y, z, sigma, logit, rgb):
    cloud.positions[i] = (x, y, z)
    cloud.scales[i] = math.log(sigma)
    cloud.rotations[i] = (1, 0, 0, 0)
    cloud.opacities[i] = logit
    cloud.sh[i, 0] = (np.asarray(rgb) - 0.5) / float(F.SH_C0)


def test_compositing_matches_closed_form():
    """Isotropic Gaussians in front of an identity camera project, by the EWA
    Jacobian J = [[f/z, 0, -f x/z^2], [0, f/z, -f y/z^2]], to the 2D covariance
    s^2 J J^T + 0.3 I; alpha = o exp(-d^T cov^-1 d / 2), cut at the radius
    ceil(3 sqrt(lambda_max)) (lambda_max floored as mid + sqrt(max(0.1, .))),
    at 1/255 and at 0.99; they composite front to back, C = sum T_i a_i c_i + T bg."""
    cam = F.train_camera(48, 40)
    cloud = F.Cloud.zeros(3, 0)
    _facing_gaussian(cloud, 0, 0.00, 0.00, 4.0, 0.08, 0.5, (0.9, 0.2, 0.1))
    _facing_gaussian(cloud, 1, 0.05, -0.03, 3.0, 0.05, 1.2, (0.1, 0.8, 0.3))
    _facing_gaussian(cloud, 2, -0.06, 0.04, 5.0, 0.10, -0.3, (0.2, 0.3, 0.9))
    bg = (0.25, 0.5, 0.75)
    img = O.render(cloud, cam, O.RefOptions(background=bg), threads=1)
    order = np.argsort(-cloud.positions[:, 2].astype(np.float64))[::-1]  # front (small z) first
    yy, xx = np.mgrid[0:cam.height, 0:cam.width]
    px, py = xx + 0.5, yy + 0.5
    T = np.ones(xx.shape)
    C = np.zeros(xx.shape + (3,))
    for i in order:
        x, y, z = cloud.positions[i].astype(np.float64)
        s = math.exp(float(cloud.scales[i, 0]))
        J = np.array([[60.0 / z, 0.0, -60.0 * x / z ** 2], [0.0, 60.0 / z, -60.0 * y / z ** 2]])
        cov = s * s * J @ J.T + 0.3 * np.eye(2)
        mid, det = 0.5 * (cov[0, 0] + cov[1, 1]), np.linalg.det(cov)
        radius = math.ceil(3 * math.sqrt(mid + math.sqrt(max(0.1, mid * mid - det))))
        o = 1 / (1 + math.exp(-float(cloud.opacities[i])))
        cx, cy = 60.0 * x / z + cam.cx, 60.0 * y / z + cam.cy
        inv = np.linalg.inv(cov)
        dx, dy = px - cx, py - cy
        q = inv[0, 0] * dx * dx + 2 * inv[0, 1] * dx * dy + inv[1, 1] * dy * dy
        a = np.minimum(0.99, o * np.exp(-0.5 * q))
        inside = (np.abs(px - cx) <= radius) & (np.abs(py - cy) <= radius) & (a >= 1 / 255)
        a = np.where(inside, a, 0.0)
        colour = 0.5 + 0.28209479177387814 * cloud.sh[i, 0].astype(np.float64)
        C += (T * a)[..., None] * colour
        T *= 1 - a
    C += T[..., None] * np.asarray(bg)
    assert np.abs(img.rgb - C.reshape(-1, 3)).max() < 1e-12
    assert np.abs(img.log_t - np.log(T).reshape(-1)).max() < 1e-12


def test_penalties_reach_every_gaussian():
    """test_train.cpp:295-321: behind the camera, the penalties are the whole gradient."""
    cam = F.train_camera(64, 48)
    cloud = F.random_cloud(20, 1, 9)
    cloud.positions[:, 2] = -5.0
    n = cloud.count
    for reg in (0.0, 0.5):
        opts = O.RefOptions(opacity_reg=reg, scale_reg=reg * 0.5)
        g = O.gradients(cloud, cam, np.ones((cam.width * cam.height, 3)), opts, threads=1)
        o = 1 / (1 + np.exp(-cloud.opacities.astype(np.float64)))
        assert np.allclose(g.opacities, reg / n * o * (1 - o), atol=1e-12)
        assert np.allclose(g.scales, np.float32(reg * 0.5) / n * np.exp(cloud.scales.astype(np.float64)), atol=1e-12)
        assert (g.positions == 0).all()


def test_sh_basis_is_orthonormal_on_the_sphere():
    """Real SH up to degree 3: orthonormal under the uniform measure on the sphere
    (Gauss-Legendre in cos(theta), uniform in phi, exact for these degrees)."""
    import ctypes as C
    nodes, weights = np.polynomial.legendre.leggauss(16)
    phis = np.linspace(0, 2 * math.pi, 32, endpoint=False)
    G = np.zeros((16, 16))
    out = np.zeros(16)
    for ct, wt in zip(nodes, weights):
        st = math.sqrt(1 - ct * ct)
        for ph in phis:
            d = np.array([st * math.cos(ph), st * math.sin(ph), ct])
            O.lib().oa_sh_basis(d.ctypes.data_as(C.c_void_p), 3, out.ctypes.data_as(C.c_void_p))
            G += np.outer(out, out) * wt * (2 * math.pi / len(phis))
    assert np.allclose(G, np.eye(16), atol=1e-12)


def test_quaternion_rotation_is_orthonormal():
    import ctypes as C
    rng = np.random.default_rng(5)
    R = np.zeros(9)
    for _ in range(100):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        O.lib().oa_quat_to_mat(q.ctypes.data_as(C.c_void_p), R.ctypes.data_as(C.c_void_p))
        M = R.reshape(3, 3)
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-12) and abs(np.linalg.det(M) - 1) < 1e-12
        # Rotating by q is the quaternion product q v q*.
        w, x, y, z = q
        v = rng.standard_normal(3)
        t = 2 * np.cross([x, y, z], v)
        assert np.allclose(M @ v, v + w * t + np.cross([x, y, z], t), atol=1e-12)


def test_threads_do_not_change_results_beyond_summation_order():
    cam = F.train_camera(72, 56)
    cloud = F.random_cloud(60, 1, 13)
    a = O.render(cloud, cam, threads=1)
    b = O.render(cloud, cam, threads=7)
    assert np.array_equal(a.rgb, b.rgb) and np.array_equal(a.stop, b.stop)
    target = F.smooth_target(72, 56, 11)
    la, lb = O.loss(a, target, 0.2, threads=1), O.loss(b, target, 0.2, threads=5)
    assert np.array_equal(la.adjoint, lb.adjoint) and abs(la.loss - lb.loss) < 1e-15
    ga = O.gradients(cloud, cam, la.adjoint, threads=1)
    gb = O.gradients(cloud, cam, la.adjoint, threads=6)
    assert np.allclose(ga.positions, gb.positions, rtol=1e-12, atol=1e-18)
