"""Host-side pieces of openradioss_amd.autograd (no GPU): cinmas against the deck
builder's NumPy CINMAS (cinmas.F:1294-1296,1322), and its thickness derivative against
the closed form d(MS)/d(t) = sum over the node's shells of rho*A/4."""
import numpy as np
import pytest
import torch

from openradioss_amd import deck as dk
from openradioss_amd.autograd import cinmas


def test_cinmas_matches_deck_builder():
    d = dk.shell_plate(n=7, patch=3, thk_noise=0.1)
    x = torch.from_numpy(d["x"])
    ixc = torch.from_numpy(d["ixc"].astype(np.int64))
    thk = torch.from_numpy(d["thk"])
    ms, inr, area = cinmas(x, ixc, thk, d["mat"]["rho0"])
    ms0, in0, area0 = dk.shell_masses(d["x"], d["ixc"], d["thk"], d["mat"]["rho0"])
    np.testing.assert_allclose(ms.numpy(), ms0, rtol=1e-14)
    np.testing.assert_allclose(inr.numpy(), in0, rtol=1e-14)
    np.testing.assert_allclose(area.numpy(), area0, rtol=1e-14)


def test_cinmas_thickness_derivative_closed_form():
    d = dk.shell_plate(n=5, patch=1, thk_noise=0.1)
    x = torch.from_numpy(d["x"])
    ixc = torch.from_numpy(d["ixc"].astype(np.int64))
    thk = torch.from_numpy(d["thk"]).clone().requires_grad_(True)
    rho = d["mat"]["rho0"]
    ms, _, area = cinmas(x, ixc, thk, rho)
    node = int(d["ixc"][7, 2])
    ms[node].backward()
    expect = np.zeros(len(d["thk"]))
    for e in range(len(d["thk"])):
        if node in d["ixc"][e]:
            expect[e] = rho * area[e].item() / 4.0
    np.testing.assert_allclose(thk.grad.numpy(), expect, rtol=1e-14)


def test_force_weights_match_the_oracle_tables():
    import pyoracle as po
    from openradioss_amd.autograd import force_weights
    for npt in range(1, 6):
        assert np.array_equal(force_weights(npt), po.shell_tables(npt)[1]), npt


def test_plastic_work_closed_form_and_gradient():
    """One element, two points: W = sum V_pt (A p + B p^(n+1)/(n+1)), dW/dp = V_pt (A + B p^n)."""
    from openradioss_amd.autograd import plastic_work
    A = torch.tensor(200.0, dtype=torch.float64, requires_grad=True)
    B = torch.tensor(450.0, dtype=torch.float64, requires_grad=True)
    n = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    thk = torch.tensor([1.2], dtype=torch.float64)
    area = torch.tensor([4.0], dtype=torch.float64)
    pla = torch.tensor([[0.01, 0.0, 0.0, 0.0, 0.0]], dtype=torch.float64, requires_grad=True)
    W = plastic_work(A, B, n, thk, area, pla, 2)
    V = 0.5 * 1.2 * 4.0
    assert float(W) == pytest.approx(V * (200 * 0.01 + 450 * 0.01 ** 1.5 / 1.5), rel=1e-14)
    W.backward()
    assert float(pla.grad[0, 0]) == pytest.approx(V * (200 + 450 * 0.01 ** 0.5), rel=1e-14)
    assert float(pla.grad[0, 1]) == 0.0
    assert float(A.grad) == pytest.approx(V * 0.01, rel=1e-14)
