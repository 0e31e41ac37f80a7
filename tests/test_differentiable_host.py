"""Host-side pieces of the differentiable form (no GPU): the torch mapping
from initial elements to the initial state must reproduce the generator's
state, and the final-state -> elements mapping must invert it."""
import numpy as np
import torch

from posidonius_amd import differentiable as dF, generators as gen


def test_state_from_elements_reproduces_the_generator():
    spec = gen.config4(8)
    p = dF.parameters(spec, requires_grad=False)
    x, v = dF.state_from_elements(spec, p["a"], p["e"])
    np.testing.assert_allclose(x.numpy(), spec.body["position"], rtol=0, atol=1e-16)
    np.testing.assert_allclose(v.numpy(), spec.body["velocity"], rtol=0, atol=1e-16)


def test_elements_from_state_inverts_the_construction():
    spec = gen.config4(8)
    x = torch.tensor(spec.body["position"])
    v = torch.tensor(spec.body["velocity"])
    a, e = dF.elements_from_state(spec, x, v)
    planet = spec.elements["a"] > 0
    np.testing.assert_allclose(a.numpy()[planet], spec.elements["a"][planet], rtol=1e-12)
    np.testing.assert_allclose(e.numpy()[planet], spec.elements["e"][planet], rtol=1e-7, atol=1e-12)
    assert np.all(a.numpy()[~planet] == 0)


def test_elements_state_round_trip_gradient_is_identity():
    """d a_out / d a_in = 1 and d e_out / d e_in = 1 through state_from_elements
    and elements_from_state (the Kepler-equation step carries the implicit
    derivative)."""
    spec = gen.config4(2)
    p = dF.parameters(spec)
    x, v = dF.state_from_elements(spec, p["a"], p["e"])
    a, e = dF.elements_from_state(spec, x, v)
    planet = torch.tensor(spec.elements["a"] > 0)
    ga, = torch.autograd.grad(a[planet].sum(), p["a"], retain_graph=True)
    ge, = torch.autograd.grad(e[planet].sum(), p["e"])
    # the barycentric shift does not change host-relative elements
    np.testing.assert_allclose(ga.numpy()[planet.numpy()], 1.0, rtol=1e-9)
    np.testing.assert_allclose(ge.numpy()[planet.numpy()], 1.0, rtol=1e-6)
