"""Differentiable form of the ensemble integration (PyTorch autograd).

``WHFastSteps`` is a ``torch.autograd.Function`` whose forward pass runs
``n_steps`` WHFast steps of every system on the GPU (``pos_adjoint_record``:
the regular step kernel plus a state checkpoint every ``segment_steps``
steps) and whose backward pass is the library's hand-written discrete adjoint
(``pos_adjoint``: each segment is recomputed from its checkpoint while its
step tape is recorded, then swept in reverse).  Inputs are the initial
barycentric positions, velocities and spin angular momenta ([n_bodies, 3])
and per-body log dissipation factors, tidal and fluid Love numbers
([n_bodies]); outputs are the final positions, velocities and spin angular
momenta.

Around it, plain torch code maps initial osculating elements to the initial
state (as the case generators do) and the final state to semi-major axes,
eccentricities (src/tools.rs:104-150) and spin rates, so that autograd
composes gradients of final a, e and |Omega| with respect to initial a, e,
log sigma and k2.  Torch is the plumbing here: every step and its adjoint run
in libposidonius_hip.so, and there is no CPU path.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np
import torch

from .case import EnsembleSpec
from .ensemble import Ensemble

AU = 1.49597870700e11
DAY = 24.0 * 3600.0
G = 6.67428e-11 / AU ** 3 * 1.9818e30 * DAY ** 2    # K2 of src/constants.rs


def _np(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy()).copy()


class WHFastSteps(torch.autograd.Function):
    """(x0, v0, L0, log_sigma, k2, k2f) -> (x, v, L) after n_steps steps."""

    @staticmethod
    def forward(ctx, x0, v0, L0, log_sigma, k2, k2f, spec: EnsembleSpec, n_steps: int, segment_steps: int,
                options: Optional[dict] = None):
        s = spec.copy()
        s.body["position"] = _np(x0).reshape(-1, 3)
        s.body["velocity"] = _np(v0).reshape(-1, 3)
        L = _np(L0).reshape(-1, 3)
        inertia = s.body["moment_of_inertia"]
        if inertia is None:
            inertia = s.body["mass"] * s.body["radius_of_gyration_2"] * s.body["radius"] ** 2
        s.body["angular_momentum"] = L
        s.body["spin"] = L / inertia[:, None]
        sigma = np.exp(_np(log_sigma))
        s.body["tides_dissipation_factor"] = sigma
        s.body["tides_dissipation_factor_scale"] = np.ones_like(sigma)
        s.body["tides_love_number"] = _np(k2)
        s.body["flattening_love_number"] = _np(k2f)
        opts = dict(mapping="packed", use_graph=True)
        opts.update(options or {})
        ens = Ensemble(s, **opts)
        try:
            ens.adjoint_record(n_steps, segment_steps)
            st = ens.get_state(fields=("position", "velocity", "angular_momentum", "status"))
        except Exception:
            ens.close()
            raise
        ctx.ens = ens
        ctx.sigma = sigma
        ctx.status = st["status"]
        dev = x0.device
        return tuple(torch.from_numpy(st[k]).to(dev) for k in ("position", "velocity", "angular_momentum"))

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gx, gv, gL):
        ens = ctx.ens
        try:
            g = ens.adjoint(None if gx is None else _np(gx), None if gv is None else _np(gv),
                            None if gL is None else _np(gL))
        finally:
            ens.close()
            ctx.ens = None
        dev = gx.device if gx is not None else (gv.device if gv is not None else gL.device)
        t = lambda a: torch.from_numpy(a).to(dev)
        return (t(g["position"]), t(g["velocity"]), t(g["angular_momentum"]),
                t(g["dissipation_factor"] * ctx.sigma), t(g["love_number"]), t(g["fluid_love_number"]),
                None, None, None, None)


# ---------------------------------------------------------------------------
# Elements <-> state, in torch
# ---------------------------------------------------------------------------
def _system_index(spec: EnsembleSpec) -> np.ndarray:
    return np.repeat(np.arange(spec.n_systems), spec.n_per_system())


def _host_index(spec: EnsembleSpec) -> np.ndarray:
    """Per body, the index of its system's host (most massive body)."""
    m = spec.body["mass"]
    hosts = np.array([spec.body_offset[s] + int(np.argmax(m[spec.body_offset[s]:spec.body_offset[s + 1]]))
                      for s in range(spec.n_systems)])
    return hosts[_system_index(spec)]


def state_from_elements(spec: EnsembleSpec, a: torch.Tensor, e: torch.Tensor):
    """Initial barycentric positions and velocities from per-body semi-major
    axis and eccentricity (host rows ignored); the other elements are the
    spec's (spec.elements).  Same construction as the case generators:
    elliptic Kepler orbit about the host with gm = G (m_host + m_body), then
    the shift to the barycentre of each system."""
    el = spec.elements
    if not el:
        raise ValueError("the spec carries no osculating elements (build it with posidonius_amd.generators)")
    dev = a.device
    T = lambda v: torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)
    m = T(spec.body["mass"])
    hidx = _host_index(spec)
    planet = torch.as_tensor(np.arange(spec.n_bodies) != hidx, device=dev)
    gm = G * (m[torch.as_tensor(hidx, device=dev)] + m)
    inc, varpi, node, M = T(el["inc"]), T(el["varpi"]), T(el["node"]), T(el["mean_anomaly"])
    a_ = torch.where(planet, a, torch.ones_like(a))
    e_ = torch.where(planet, e, torch.zeros_like(e))
    # eccentric anomaly: converged without gradient, then one Newton step that
    # carries the implicit derivative dE = (dM + sin E de) / (1 - e cos E)
    with torch.no_grad():
        E0 = M.clone()
        for _ in range(60):
            E0 = E0 - (E0 - e_ * torch.sin(E0) - M) / (1.0 - e_ * torch.cos(E0))
    E = E0 - (E0 - e_ * torch.sin(E0) - M) / (1.0 - e_ * torch.cos(E0))
    g = varpi - node
    si, ci = torch.sin(inc), torch.cos(inc)
    sg, cg = torch.sin(g), torch.cos(g)
    sn, cn = torch.sin(node), torch.cos(node)
    z1, z2, z3, z4 = cg * cn, cg * sn, sg * cn, sg * sn
    d11, d12, d13 = z1 - z4 * ci, z2 + z3 * ci, sg * si
    d21, d22, d23 = -z3 - z2 * ci, -z4 + z1 * ci, cg * si
    se, ce = torch.sin(E), torch.cos(E)
    romes = torch.sqrt(1.0 - e_ * e_)
    x1 = a_ * (ce - e_)
    x2 = a_ * romes * se
    temp = torch.sqrt(gm / a_) / (1.0 - e_ * ce)
    v1 = -se * temp
    v2 = romes * ce * temp
    pos = torch.stack([d11 * x1 + d21 * x2, d12 * x1 + d22 * x2, d13 * x1 + d23 * x2], dim=-1)
    vel = torch.stack([d11 * v1 + d21 * v2, d12 * v1 + d22 * v2, d13 * v1 + d23 * v2], dim=-1)
    zero = torch.zeros_like(pos)
    pos = torch.where(planet[:, None], pos, zero)
    vel = torch.where(planet[:, None], vel, zero)
    sidx = torch.as_tensor(_system_index(spec), device=dev)
    mtot = torch.zeros(spec.n_systems, dtype=m.dtype, device=dev).index_add_(0, sidx, m)
    cpos = torch.zeros(spec.n_systems, 3, dtype=m.dtype, device=dev).index_add_(0, sidx, m[:, None] * pos)
    cvel = torch.zeros(spec.n_systems, 3, dtype=m.dtype, device=dev).index_add_(0, sidx, m[:, None] * vel)
    return pos - (cpos / mtot[:, None])[sidx], vel - (cvel / mtot[:, None])[sidx]


def elements_from_state(spec: EnsembleSpec, x: torch.Tensor, v: torch.Tensor):
    """Semi-major axis and eccentricity of every body about its host
    (src/tools.rs:104-150 with gm = G (m_host + m_body)); host rows are zero."""
    dev = x.device
    m = torch.as_tensor(spec.body["mass"], dtype=torch.float64, device=dev)
    hidx = torch.as_tensor(_host_index(spec), device=dev)
    planet = torch.arange(spec.n_bodies, device=dev) != hidx
    r_vec = x - x[hidx]
    u_vec = v - v[hidx]
    gm = G * (m[hidx] + m)
    # host rows get a harmless unit orbit so that no derivative is infinite
    unit = torch.tensor([1.0, 0.0, 0.0], dtype=x.dtype, device=dev)
    r_vec = torch.where(planet[:, None], r_vec, unit)
    u_vec = torch.where(planet[:, None], u_vec, torch.sqrt(gm)[:, None] * unit.roll(1))
    r = torch.sqrt((r_vec * r_vec).sum(-1))
    v2 = (u_vec * u_vec).sum(-1)
    h = torch.cross(r_vec, u_vec, dim=-1)
    h2 = (h * h).sum(-1)
    a = gm * r / (2.0 * gm - r * v2)
    s = h2 / gm
    temp = 1.0 + s * (v2 / gm - 2.0 / r)
    e = torch.sqrt(torch.clamp(temp, min=1e-300))
    zero = torch.zeros_like(r)
    return torch.where(planet, a, zero), torch.where(planet, e, zero)


def spin_rate(spec: EnsembleSpec, L: torch.Tensor):
    """|Omega| = |L| / I per body."""
    I = spec.body["moment_of_inertia"]
    if I is None:
        I = spec.body["mass"] * spec.body["radius_of_gyration_2"] * spec.body["radius"] ** 2
    return torch.sqrt((L * L).sum(-1)) / torch.as_tensor(I, dtype=torch.float64, device=L.device)


def parameters(spec: EnsembleSpec, device="cpu", requires_grad=True) -> Dict[str, torch.Tensor]:
    """Leaf tensors for the differentiable inputs of a spec: initial a, e
    (from spec.elements), log sigma (sigma = scale * dissipation factor) and
    the tidal Love number k2 (the fluid Love number is tied to it)."""
    B = spec.body
    sig = B["tides_dissipation_factor"] * B["tides_dissipation_factor_scale"]
    out = {"a": spec.elements["a"], "e": spec.elements["e"], "log_sigma": np.log(sig), "k2": B["tides_love_number"]}
    return {k: torch.tensor(np.asarray(v, dtype=np.float64), device=device, requires_grad=requires_grad)
            for k, v in out.items()}


def integrate(spec: EnsembleSpec, n_steps: int, params: Dict[str, torch.Tensor], segment_steps: int = 1000,
              options: Optional[dict] = None):
    """Final a, e, spin rate (and state) of every body after n_steps steps,
    differentiable with respect to params (see ``parameters``)."""
    x0, v0 = state_from_elements(spec, params["a"], params["e"])
    dev = x0.device
    L0 = torch.as_tensor(spec.body["angular_momentum"], dtype=torch.float64, device=dev)
    k2 = params["k2"]
    x, v, L = WHFastSteps.apply(x0, v0, L0, params["log_sigma"], k2, k2, spec, int(n_steps), int(segment_steps),
                                options)
    a, e = elements_from_state(spec, x, v)
    return {"a": a, "e": e, "spin_rate": spin_rate(spec, L), "x": x, "v": v, "L": L}
