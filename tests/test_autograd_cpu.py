"""Host-side pieces of openradioss_amd.autograd (no GPU): cinmas against the deck
builder's NumPy CINMAS (cinmas.F:1294-1296,1322), and its thickness derivative against
the closed form d(MS)/d(t) = sum over the node's shells of rho*A/4."""
import numpy as np
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
