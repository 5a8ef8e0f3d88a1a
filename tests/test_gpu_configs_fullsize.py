"""Two benchmark configurations asserted at their full sizes.

Precision study (the w-projection layout of the headline workload at 1530
time samples: 49,939,200 points, 4096^2 grid, 32 w-planes, supports 8..64):
FP32 gridding with plain and with compensated (TwoSum) tile flushes, and
FP32 degridding, within 1e-5 relative L2 of the FP64 GPU path on the same
complex64-rounded inputs (the FP64 path is pinned to the reference's CPU
gridder at 1e-12 elsewhere).

Measurement operator (2048^2 grid, 16 w-planes, supports 8..48, 256
antennas x 153 times = 4,993,920 points): the adjoint identity
y^H (A x) = (A^H y)^H x within 1e-5 relative, and one central finite
difference of chi^2 = |A x - d|^2 against the autograd gradient within 1e-3
(chi^2 is quadratic, so the difference has no truncation error; the bar
covers FP32 rounding of chi^2)."""
import numpy as np
import pytest
import torch

from rc_amd import device as D
from rc_amd import operator as M
from rc_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    d = a.to(torch.complex128) - b
    return float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(b))


def test_precision_study_fullsize(hip_device):
    vc, gcf, N = synth.wproj(n_times=1530)
    assert vc.n == 49_939_200
    gcf = gcf.astype(np.complex64).astype(np.complex128)
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg32 = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    dg64 = D.DeviceGCF.from_table(gcf, dtype=np.complex128)
    vis32 = torch.from_numpy(vc.vis.reshape(-1).astype(np.complex64)).cuda()
    del vc
    plan = D.TiledPlan(geom, tile=32, chunk=1024, max_support=64)
    plan.bin(dg32)
    assert plan.stats()["err"] == 0
    pg, _, _ = D.pregrid(geom, dg64)
    g64 = D.grid_plain_f64(geom, dg64, pg, vis32.to(torch.complex128))
    errs = {}
    for comp in (False, True):
        plan.compensated(comp)
        g32 = plan.grid(dg32, vis32)
        torch.cuda.synchronize()
        errs["compensated" if comp else "plain"] = rel(g32, g64)
        del g32
    plan.compensated(False)
    g_in = g64.to(torch.complex64)
    del g64
    v32 = plan.degrid(dg32, g_in)
    v64 = D.degrid_plain_f64(geom, dg64, pg, g_in.to(torch.complex128))
    torch.cuda.synchronize()
    errs["degrid"] = rel(v32, v64)
    plan.close()
    print(errs)
    assert errs["plain"] <= 1e-5
    assert errs["compensated"] <= 1e-5
    assert errs["compensated"] <= errs["plain"]
    assert errs["degrid"] <= 1e-5


def test_operator_fullsize_adjoint_and_fd(hip_device):
    N = 2048
    vc, gcf, _ = synth.wproj(seed=5, n_times=153, grid_size=N, n_planes=16, s_min=8, s_max=48)
    assert vc.n == 4_993_920
    geom = D.DeviceGeom.from_chunks(vc, N)
    dg = D.DeviceGCF.from_table(gcf, dtype=np.complex64)
    op = M.MeasurementOp(geom, dg, tile=32, chunk=1024, max_support=48)
    gen = torch.Generator(device="cuda").manual_seed(0)
    for complex_image in (False, True):
        if complex_image:
            x = torch.randn((N, N), dtype=torch.complex64, device="cuda", generator=gen)
        else:
            x = torch.randn((N, N), dtype=torch.float32, device="cuda", generator=gen)
        y = torch.randn(op.n, dtype=torch.complex64, device="cuda", generator=gen)
        ax = op.forward(x).to(torch.complex128)
        ahy = op.adjoint(y).to(torch.complex128)
        lhs = torch.vdot(y.to(torch.complex128), ax)
        rhs = torch.vdot(ahy.reshape(-1), x.to(torch.complex128).reshape(-1))
        assert abs(complex(lhs - rhs)) <= 1e-5 * abs(complex(lhs)), (complex_image, lhs, rhs)
    # one central finite difference of chi^2 along a random direction
    x = torch.randn((N, N), dtype=torch.float32, device="cuda", generator=gen)
    d = op.forward(x + 0.3 * torch.randn((N, N), dtype=torch.float32, device="cuda",
                                         generator=gen))
    xg = x.clone().requires_grad_(True)
    M.chi2(op, xg, d).backward()
    v = torch.randn((N, N), dtype=torch.float32, device="cuda", generator=gen)
    h = 0.05

    def chi2(z):
        r = (op.forward(z) - d).to(torch.complex128)
        return float((r.real.square() + r.imag.square()).sum())

    fd = (chi2(x + h * v) - chi2(x - h * v)) / (2 * h)
    an = float((xg.grad.double() * v.double()).sum())
    op.close()
    assert abs(fd - an) <= 1e-3 * abs(an), (fd, an)
