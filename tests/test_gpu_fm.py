"""GPU checks of the force-matching adjoint (mtmd_fm_*).

- model forces (FP64) against minus the central-difference gradient of an
  independent numpy energy (tests/fm_model.py) on a tiny water box;
- FP64 adjoint gradient against central finite differences of the FP64 loss
  (relative step 1e-6), to 1e-6 relative;
- FP32 adjoint against the FP64 gradient;
- the torch.autograd.Function and an Adam fit that recovers perturbed
  parameters.
"""
import numpy as np
import pytest

from moltwister_amd import workloads as W
from moltwister_amd.forcematch import ForceMatch, params_of

import fm_model

pytestmark = pytest.mark.gpu

CUT = 6.0


@pytest.fixture(scope="module")
def water():
    # 100 waters, L = 14.4 A; frames: jittered copies of the generated box
    s = W.water_box(100, seed=9)
    rng = np.random.default_rng(5)
    frames = np.stack([(s.pos + rng.normal(0, 0.05, s.pos.shape)).reshape(-1) for _ in range(3)])
    return s, frames


def perturbed(p0, seed):
    return p0 * np.random.default_rng(seed).uniform(0.9, 1.1, p0.shape)


def test_fm_forces_against_numpy_energy_gradient(water):
    print("case fm forces vs numpy", flush=True)
    s, frames = water
    sub = 8                                                      # 8 molecules: numpy FD stays fast
    from moltwister_amd.system import System
    t = System(pos=s.pos[:3 * sub], mass=s.mass[:3 * sub], charge=s.charge[:3 * sub], types=s.types[:3 * sub],
               box=s.box, nonbonded=s.nonbonded, bond_ff=s.bond_ff, angle_ff=s.angle_ff,
               bonds=s.bonds[:2 * sub], angles=s.angles[:sub])
    fr = frames[:, :9 * sub]
    fm = ForceMatch(t, CUT, fr)
    p = perturbed(params_of(t), 1)
    for f in range(fr.shape[0]):
        pos = fr[f].reshape(-1, 3)
        Fg = fm.forces(p, f, fp64=True)
        Fn = fm_model.forces_fd(t, pos, p, CUT)
        scale = np.abs(Fn).max()
        assert np.abs(Fg - Fn).max() < 1e-6 * scale, np.abs(Fg - Fn).max() / scale
        F32 = fm.forces(p, f, fp64=False)
        assert np.sqrt(np.sum((F32 - Fg) ** 2) / np.sum(Fg ** 2)) < 1e-5
    fm.close()


def test_fm_fp64_gradient_against_finite_differences(water):
    print("case fm gradient vs fd", flush=True)
    s, frames = water
    fm = ForceMatch(s, CUT, frames)
    p0 = params_of(s)
    fm.target_from_params(perturbed(p0, 2))
    L, g = fm.loss_grad(p0, fp64=True)
    assert L > 0
    for k in range(len(p0)):
        h = 1e-6 * abs(p0[k])
        pp, pm = p0.copy(), p0.copy()
        pp[k] += h
        pm[k] -= h
        fd = (fm.loss_grad(pp, fp64=True)[0] - fm.loss_grad(pm, fp64=True)[0]) / (2 * h)
        assert abs(g[k] - fd) <= 1e-6 * max(abs(fd), 1e-12 * L), (k, g[k], fd)
    fm.close()


def test_fm_fp32_gradient_against_fp64(water):
    print("case fm fp32 vs fp64", flush=True)
    s, frames = water
    fm = ForceMatch(s, CUT, frames)
    p0 = params_of(s)
    fm.target_from_params(perturbed(p0, 3))
    L64, g64 = fm.loss_grad(p0, fp64=True)
    L32, g32 = fm.loss_grad(p0, fp64=False)
    assert abs(L32 - L64) <= 1e-4 * L64
    assert np.max(np.abs(g32 - g64) / np.abs(g64)) < 1e-3, np.abs(g32 - g64) / np.abs(g64)
    fm.close()


@pytest.mark.timeout(900)
def test_fm_torch_function_and_adam_recovers_parameters(water):
    print("case fm torch + adam", flush=True)
    import torch
    from moltwister_amd.forcematch import fit, loss_function
    s, frames = water
    fm = ForceMatch(s, CUT, frames)
    p0 = params_of(s)
    truth = perturbed(p0, 4)
    fm.target_from_params(truth)
    # autograd: backward of the Function equals the adjoint gradient
    x = torch.tensor(p0, requires_grad=True)
    loss = loss_function(fm, fp64=True)(x)
    loss.backward()
    _, g = fm.loss_grad(p0, fp64=True)
    assert np.allclose(x.grad.numpy(), g, rtol=0, atol=0)
    theta, hist = fit(fm, p0, iters=500, lr=2e-2)
    assert hist[-1] < 1e-5 * hist[0], (hist[0], hist[-1])
    # bonded parameters are pinned by the data; (eps, sigma) of O-O sit in a
    # shallow valley (LJ forces are small beside the bonded ones and the two
    # trade off over the sampled distances), so they are held to 10 %
    rel = np.abs(theta / truth - 1)
    assert np.max(rel[2:]) < 1e-3 and np.max(rel[:2]) < 0.1, theta / truth
    fm.close()
