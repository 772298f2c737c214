"""The GPU's hand-written backward through the autograd harness
(tests/gradcheck.py) against FP64 central finite differences (h = 1e-6) of the
FP64 restatement of the reference, on the differentiation-stress crop (200
Gaussians, heavy-tailed scales, high opacities, 64x64, most pixels at the
transmittance floor): relative L2 error at most 1e-3 per parameter group,
with float sums, for the backward that reads the forward pass's transmittance,
for the one that recomputes it, and with the SH stored as bf16. The FP64 loss
is taken of the cloud as the trainer stores it (SH rounded to the storage type).

And the recomputing backward against the stored one, bit for bit with
fixed-point sums (and so against Oracle B), on the crop and on larger cases.
"""
import numpy as np
import pytest

import gradcheck as G
from oracle import oracle as O
from vsplat import fixtures as F
import vsplat

pytestmark = pytest.mark.gpu


class GpuBackend:
    """The trainer through the C ABI: compute_gradients, then the gradients."""

    def __init__(self, cam, target, lam, **options):
        self.cam, self.target, self.lam, self.options = cam, target, lam, options

    def __call__(self, cloud):
        with vsplat.Trainer(cloud, vsplat.TrainerOptions(lambda_ssim=self.lam, **self.options)) as t:
            stats = t.compute_gradients(self.cam, self.target)
            g = t.gradients()
        return stats.loss, [g.positions, g.scales, g.rotations, g.opacities, g.sh]


@pytest.mark.parametrize("recompute,sh_storage", [(False, 0), (True, 0), (False, 1)])
def test_backward_against_fp64_finite_differences(recompute, sh_storage):
    cloud, cam, target, lam = G.crop_problem(sh_storage=sh_storage)
    gaussians = np.random.default_rng(11).choice(cloud.count, 16, replace=False)
    backend = GpuBackend(cam, target, lam, force_fixed_point=False, recompute_transmittance=recompute, sh_storage=sh_storage)
    for r in G.check(backend, cloud, cam, target, lam, gaussians):
        print(r, flush=True)
        assert r.on_a_jump <= 0.02 * (r.compared + r.on_a_jump), r
        assert r.rel_err <= 1e-3, r


def _cases():
    cloud, cam, target, _ = G.crop_problem()
    yield "crop", cloud, cam, target
    w = F.config(1, count=20_000)
    yield "config1_small", w.cloud, w.cameras[0], w.targets[0]
    w = F.config(4, count=50_000)
    yield "config4_small", w.cloud, w.cameras[0], w.targets[0]


@pytest.mark.parametrize("case", list(range(3)))
def test_recomputed_transmittance_gives_the_same_sums(case):
    name, cloud, cam, target = list(_cases())[case]
    raws, images, sizes = [], [], []
    for recompute in (False, True):
        with vsplat.Trainer(cloud, vsplat.TrainerOptions(force_fixed_point=True, recompute_transmittance=recompute)) as t:
            t.compute_gradients(cam, target)
            raws.append(t.inspect("raw"))
            images.append(t.inspect("image"))
            sizes.append(t.gpu_bytes())
            if recompute:  # no per-pixel stops are kept for a backward that recomputes them
                with pytest.raises(vsplat.LogicError, match="stops"):
                    t.inspect("stops")
    np.testing.assert_array_equal(images[1], images[0])
    assert sizes[0] - sizes[1] >= cam.width * cam.height * 4
    print(f"{name}: {int((raws[0] != 0).any(axis=-1).sum()) if raws[0].ndim > 1 else int((raws[0] != 0).sum())} non-zero", flush=True)
    np.testing.assert_array_equal(raws[1], raws[0])
    if name == "crop":
        m = O.Model(cloud, cam)
        r = m.forward()
        m.loss(r, target, 0.2)
        m.backward_fixed(r)
        np.testing.assert_array_equal(raws[1], r.raw_fixed)
