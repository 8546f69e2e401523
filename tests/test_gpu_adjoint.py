"""Hand-written adjoint of the velocity chain (gpk_velocity_field /
gpk_velocity_adjoint behind gpk_amd.adjoint.VelocityField).

Loss L = 1/2 sum over the band of (v(theta) - v_t)^2, v_t = v(theta_t),
theta = (eps4, eps6, M0).  The device gradient is checked three ways:
  * against torch autograd through an independent FP64 restatement of the
    chain (law + pair projection written with torch.roll), built from the
    normals and curvature the device reports;
  * against central finite differences of the FP64 forward;
  * FP32 against FP64.
Cases: 2-D 512^2 disc r = 100, sigma(n) = 1 + eps4 cos 4t + eps6 cos 6t;
3-D 128^3 sphere r = 40, cubic sigma (d/d eps6 = 0)."""
import numpy as np
import pytest
import torch

import gpk_amd
from gpk_amd import _abi
from gpk_amd.adjoint import velocity_field

pytestmark = pytest.mark.gpu

THETA = (0.03, 0.01, 1.0)
TARGET = (0.05, 0.0, 1.0)
VEL_SIGMA = 0.85
PROJ_R = 3


def disc(N, r):
    y, x = np.mgrid[0:N, 0:N]
    return (((y - N // 2) ** 2 + (x - N // 2) ** 2) <= r * r).astype(np.int32)[None]


def sphere(N, r):
    z, y, x = np.mgrid[0:N, 0:N, 0:N]
    c = N // 2
    return (((z - c) ** 2 + (y - c) ** 2 + (x - c) ** 2) <= r * r).astype(np.int32)


def make_grid(phi, prec):
    nz, ny, nx = phi.shape
    G = gpk_amd.GPKGrid(nz, ny, nx, 1.0, _abi.anisotropic_growth(1.0, 1.0, THETA[0], THETA[1]),
                        0.0, 1.5, precision=prec)
    G.set("phi", phi)
    G.set("g", np.zeros(phi.shape, np.float32))
    return G


def loss_and_grad(G, theta, vt):
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    v = velocity_field(th, G)
    loss = 0.5 * ((v.double() - vt) ** 2).sum()
    loss.backward()
    return float(loss.detach()), th.grad.numpy().copy()


def forward_loss(G, theta, vt):
    with torch.no_grad():
        v = velocity_field(torch.tensor(theta, dtype=torch.float64), G)
        return float(0.5 * ((v.double() - vt) ** 2).sum())


def target(G):
    with torch.no_grad():
        return velocity_field(torch.tensor(TARGET, dtype=torch.float64), G).double()


def torch_reference(G, theta, vt_np):
    """FP64 chain on the CPU from the device's geometry (labels fixed)."""
    mask = torch.from_numpy(G.get("mask").astype(bool))
    phi = torch.from_numpy(G.get("phi").astype(np.int64))
    comp = torch.from_numpy(G.get("phi_comp").astype(np.int64))
    nx = torch.from_numpy(G.get("nx").astype(np.float64))
    ny = torch.from_numpy(G.get("ny").astype(np.float64))
    kap = torch.from_numpy(G.get("kappa").astype(np.float64))
    # the model descriptor holds float parameters
    th = torch.tensor([float(np.float32(t)) for t in theta], dtype=torch.float64, requires_grad=True)
    ang = torch.atan2(ny, nx)
    sig = 1.0 * (1.0 + th[0] * torch.cos(4 * ang) + th[1] * torch.cos(6 * ang))
    vraw = torch.where(mask, th[2] * (0.0 - sig * kap), torch.zeros_like(kap))
    proj = mask & (phi != comp)
    lo, hi = torch.minimum(phi, comp), torch.maximum(phi, comp)
    o = torch.where(phi == lo, 1.0, -1.0).double()
    num = torch.zeros_like(kap)
    den = torch.zeros_like(kap)
    axes = tuple(range(phi.dim()))
    for dy in range(-PROJ_R, PROJ_R + 1):
        for dx in range(-PROJ_R, PROJ_R + 1):
            sh = (-dy, -dx)
            pj = torch.roll(proj, sh, axes)
            same = (torch.roll(lo, sh, axes) == lo) & (torch.roll(hi, sh, axes) == hi)
            c = (proj & pj & same).double()
            # FP64 mode: weights exp(-r^2 / 2 s^2) in double, s = float(0.85)
            sd = float(np.float32(VEL_SIGMA))
            w = float(np.exp(-float(dx * dx + dy * dy) * (0.5 / (sd * sd))))
            num = num + w * c * torch.roll(o * vraw, sh, axes)
            den = den + w * c
    v = torch.where(proj, o * num / torch.where(den > 0, den, torch.ones_like(den)), vraw)
    loss = 0.5 * ((v - torch.from_numpy(vt_np)) ** 2).sum()
    loss.backward()
    return v.detach().numpy(), float(loss.detach()), th.grad.numpy()


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def test_adjoint_2d_matches_torch_autograd_reference():
    G = make_grid(disc(512, 100), "fp64")
    vt = target(G)
    loss, grad = loss_and_grad(G, THETA, vt)
    with torch.no_grad():
        v = velocity_field(torch.tensor(THETA, dtype=torch.float64), G).cpu().numpy()
    v_ref, loss_ref, grad_ref = torch_reference(G, THETA, vt.cpu().numpy())
    band = G.get("mask").astype(bool)
    print("band", int(band.sum()), "grad", grad, "ref", grad_ref, flush=True)
    assert band.sum() > 1000
    assert np.abs(v - v_ref).max() <= 1e-12 * np.abs(v_ref).max()
    assert abs(loss - loss_ref) <= 1e-10 * loss_ref
    assert np.all(rel(grad, grad_ref) <= 1e-9), (grad, grad_ref)
    G.close()


def central_fd(G, theta, vt, k, h=1e-4):
    tp, tm = list(theta), list(theta)
    tp[k] += h
    tm[k] -= h
    # the model stores float parameters: difference over the values it used
    step = float(np.float32(tp[k])) - float(np.float32(tm[k]))
    return (forward_loss(G, tp, vt) - forward_loss(G, tm, vt)) / step


@pytest.mark.parametrize("shape", ["disc2d", "sphere3d"])
def test_adjoint_fp64_matches_central_differences(shape):
    phi = disc(512, 100) if shape == "disc2d" else sphere(128, 40)
    G = make_grid(phi, "fp64")
    vt = target(G)
    _, grad = loss_and_grad(G, THETA, vt)
    fd = np.array([central_fd(G, THETA, vt, k) for k in range(3)])
    print(shape, "adjoint", grad, "fd", fd, flush=True)
    if shape == "sphere3d":
        assert grad[1] == 0.0
        ks = (0, 2)
    else:
        ks = (0, 1, 2)
    for k in ks:
        assert rel(grad[k], fd[k]) <= 1e-5, (k, grad[k], fd[k])
    G.close()


@pytest.mark.parametrize("shape", ["disc2d", "sphere3d"])
def test_adjoint_fp32_against_fp64(shape):
    phi = disc(512, 100) if shape == "disc2d" else sphere(128, 40)
    G64 = make_grid(phi, "fp64")
    G32 = make_grid(phi, "fp32")
    vt64 = target(G64)
    _, g64 = loss_and_grad(G64, THETA, vt64)
    _, g32 = loss_and_grad(G32, THETA, vt64.float())
    print(shape, "fp32", g32, "fp64", g64, flush=True)
    ks = (0, 2) if shape == "sphere3d" else (0, 1, 2)
    for k in ks:
        assert rel(g32[k], g64[k]) <= 1e-3, (k, g32[k], g64[k])
    G64.close()
    G32.close()
