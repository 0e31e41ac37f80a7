"""Ensemble specifications and the reference's JSON case format.

A *case* is what the reference writes with ``Universe.write`` (Python,
posidonius/particles/universe.py:350-364) or ``write_recovery_snapshot``
(Rust, src/integrator/output.rs:28-54): a serialised ``WHFast`` integrator
holding a ``Universe`` of up to MAX_PARTICLES particles (serde layout of
whfast.rs:98-120, universe.rs:50-63, particle.rs:16-52).  ``read_case``
turns one such file into an :class:`EnsembleSpec` of one system;
:meth:`EnsembleSpec.concat` stacks many systems into an ensemble, which is
the unit the device works on.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import abi

# Per-system and per-body field names of an EnsembleSpec, with dtypes and
# widths (3 = vector).  Optional fields may be None.
SYSTEM_FIELDS = {
    "coordinates": np.int32, "effects": np.uint32, "gr_implementation": np.int32,
    "time_step": np.float64, "time_limit": np.float64, "current_time": np.float64,
    "historic_snapshot_period": np.float64, "last_historic_snapshot_time": np.float64,
    "current_iteration": np.int64,
}
BODY_FIELDS = {
    "mass": (np.float64, 1), "radius": (np.float64, 1), "radius_of_gyration_2": (np.float64, 1),
    "moment_of_inertia": (np.float64, 1), "position": (np.float64, 3), "velocity": (np.float64, 3),
    "spin": (np.float64, 3), "angular_momentum": (np.float64, 3), "heliocentric_velocity": (np.float64, 3),
    "roles": (np.uint32, 1), "tides_dissipation_factor": (np.float64, 1),
    "tides_dissipation_factor_scale": (np.float64, 1), "tides_love_number": (np.float64, 1),
    "flattening_love_number": (np.float64, 1), "wind_k_factor": (np.float64, 1),
    "wind_rotation_saturation": (np.float64, 1), "velocity_error": (np.float64, 3),
    "angular_momentum_error": (np.float64, 3), "evolution_type": (np.int32, 1),
    "evolver_left_index": (np.int32, 1), "pair_dissipation_factor": (np.float64, 1),
    "roche_radius": (np.float64, abi.MAX_BODIES),
}
# fill value of a per-body field a spec does not carry (NaN: no map entry)
BODY_FILL = {"pair_dissipation_factor": np.nan, "roche_radius": np.nan}
MAX_PARTICLES = 10   # src/constants.rs:3, the key base of the pair-dependent factor map
TABLE_FIELDS = ("evolver_time", "evolver_radius", "evolver_radius_of_gyration_2", "evolver_love_number",
                "evolver_inverse_tidal_q_factor")
# Per-system fields wider than one value: DiskProperties of the disk host
# (inner_edge_distance, outer_edge_distance, lifetime, alpha,
# surface_density_normalization, mean_molecular_weight; disk.rs:7-14)
SYSTEM_WIDE_FIELDS = {"disk_properties": 6}
DISK_KEYS = ("inner_edge_distance", "outer_edge_distance", "lifetime", "alpha", "surface_density_normalization",
             "mean_molecular_weight")

_EVOLUTION_CODES = {
    "NonEvolving": abi.EVO_NON_EVOLVING, "Leconte2011": abi.EVO_LECONTE2011,
    "Baraffe1998": abi.EVO_BARAFFE1998, "Baraffe2015": abi.EVO_BARAFFE2015,
    "BolmontMathis2016": abi.EVO_BOLMONT_MATHIS2016, "GalletBolmont2017": abi.EVO_GALLET_BOLMONT2017,
}


@dataclass
class EnsembleSpec:
    """Structure-of-arrays description of an ensemble (mirrors pos_ensemble_desc)."""
    body_offset: np.ndarray
    system: Dict[str, Optional[np.ndarray]] = field(default_factory=dict)
    body: Dict[str, Optional[np.ndarray]] = field(default_factory=dict)
    evolver_offset: Optional[np.ndarray] = None
    tables: Dict[str, np.ndarray] = field(default_factory=dict)
    names: List[str] = field(default_factory=list)
    # optional per-body osculating elements the initial state was built from
    # (host rows zero): a, e, inc, varpi (longitude of perihelion), node, mean_anomaly
    elements: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def n_systems(self) -> int:
        return len(self.body_offset) - 1

    @property
    def n_bodies(self) -> int:
        return int(self.body_offset[-1])

    def n_per_system(self) -> np.ndarray:
        return np.diff(self.body_offset)

    # -- composition ------------------------------------------------------
    @staticmethod
    def concat(specs: List["EnsembleSpec"]) -> "EnsembleSpec":
        counts = np.concatenate([s.n_per_system() for s in specs])
        off = np.zeros(len(counts) + 1, dtype=np.int32)
        off[1:] = np.cumsum(counts)
        out = EnsembleSpec(body_offset=off)
        for k, dt in SYSTEM_FIELDS.items():
            parts = [s.system.get(k) for s in specs]
            out.system[k] = None if any(p is None for p in parts) else np.concatenate(parts).astype(dt)
        for k, w in SYSTEM_WIDE_FIELDS.items():
            parts = [s.system.get(k) for s in specs]
            if all(p is None for p in parts):
                out.system[k] = None
            else:
                out.system[k] = np.concatenate([np.zeros((s.n_systems, w)) if p is None else np.asarray(p).reshape(-1, w)
                                                for s, p in zip(specs, parts)]).astype(np.float64)
        for k, (dt, w) in BODY_FIELDS.items():
            parts = [s.body.get(k) for s in specs]
            if all(p is None for p in parts):
                out.body[k] = None
                continue
            filled = []
            for s, p in zip(specs, parts):
                if p is None:
                    p = np.full((s.n_bodies, w) if w > 1 else s.n_bodies, BODY_FILL.get(k, 0), dtype=dt)
                filled.append(p)
            out.body[k] = np.concatenate(filled).astype(dt)
        if any(s.evolver_offset is not None for s in specs):
            eo = [0]
            tabs = {k: [] for k in TABLE_FIELDS}
            for s in specs:
                if s.evolver_offset is None:
                    eo.extend([eo[-1]] * s.n_bodies)
                    continue
                base = eo[-1]
                eo.extend((base + s.evolver_offset[1:]).tolist())
                rows = int(s.evolver_offset[-1])
                for k in TABLE_FIELDS:
                    tabs[k].append(s.tables[k] if k in s.tables else np.zeros(rows))
            out.evolver_offset = np.asarray(eo, dtype=np.int32)
            out.tables = {k: (np.concatenate(v) if v else np.zeros(0)).astype(np.float64) for k, v in tabs.items()}
        out.names = sum((s.names or [""] * s.n_systems for s in specs), [])
        if specs and all(s.elements for s in specs):
            out.elements = {k: np.concatenate([s.elements[k] for s in specs]) for k in specs[0].elements}
        return out

    def subset(self, systems) -> "EnsembleSpec":
        return EnsembleSpec.concat([self._one(int(s)) for s in systems])

    def _one(self, s: int) -> "EnsembleSpec":
        b0, b1 = int(self.body_offset[s]), int(self.body_offset[s + 1])
        out = EnsembleSpec(body_offset=np.array([0, b1 - b0], dtype=np.int32))
        for k, v in self.system.items():
            out.system[k] = None if v is None else v[s:s + 1].copy()
        for k, v in self.body.items():
            out.body[k] = None if v is None else v[b0:b1].copy()
        if self.evolver_offset is not None:
            e0, e1 = int(self.evolver_offset[b0]), int(self.evolver_offset[b1])
            out.evolver_offset = (self.evolver_offset[b0:b1 + 1] - e0).astype(np.int32)
            out.tables = {k: v[e0:e1].copy() for k, v in self.tables.items()}
        out.names = [self.names[s]] if self.names else []
        out.elements = {k: v[b0:b1].copy() for k, v in self.elements.items()}
        return out

    def copy(self) -> "EnsembleSpec":
        return EnsembleSpec.concat([self])

    # -- C view -------------------------------------------------------------
    def desc(self):
        """A ctypes pos_ensemble_desc; the returned holder keeps arrays alive."""
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        d = abi.EnsembleDesc()
        d.n_systems = self.n_systems
        d.body_offset = abi.ptr(arr(self.body_offset, np.int32), abi.C.c_int32)
        ctypes_of = {np.int32: abi.C.c_int32, np.uint32: abi.C.c_uint32, np.float64: abi.C.c_double,
                     np.int64: abi.C.c_int64}
        for k, dt in SYSTEM_FIELDS.items():
            setattr(d, k, abi.ptr(arr(self.system.get(k), dt), ctypes_of[dt]))
        for k in SYSTEM_WIDE_FIELDS:
            setattr(d, k, abi.ptr(arr(self.system.get(k), np.float64), abi.C.c_double))
        for k, (dt, _w) in BODY_FIELDS.items():
            setattr(d, k, abi.ptr(arr(self.body.get(k), dt), ctypes_of[dt]))
        if self.evolver_offset is not None:
            d.evolver_offset = abi.ptr(arr(self.evolver_offset, np.int32), abi.C.c_int32)
            rows = int(self.evolver_offset[-1])
            for k in TABLE_FIELDS:
                setattr(d, k, abi.ptr(arr(self.tables.get(k, np.zeros(rows)), np.float64), abi.C.c_double))
        return _DescHolder(d, keep)


class _DescHolder:
    def __init__(self, desc, keep):
        self.desc = desc
        self._keep = keep

    @property
    def ref(self):
        return abi.C.byref(self.desc)


# ---------------------------------------------------------------------------
# The reference JSON case format
# ---------------------------------------------------------------------------
def _axes(d):
    return [float(d["x"]), float(d["y"]), float(d["z"])]


def _enum(value):
    """serde externally-tagged enum -> (tag, payload)."""
    if isinstance(value, str):
        return value, None
    (tag, payload), = value.items()
    return tag, payload


def _evolution_code(value) -> int:
    tag, payload = _enum(value)
    if tag == "LeconteChabrier2013":
        return abi.EVO_LECONTE_CHABRIER2013_DYN if payload else abi.EVO_LECONTE_CHABRIER2013
    if tag not in _EVOLUTION_CODES:
        raise ValueError(f"unknown evolution type {tag!r}")
    return _EVOLUTION_CODES[tag]


def read_case(path_or_dict) -> EnsembleSpec:
    """Parse a WHFast case.json (or an already-loaded dict) into one system."""
    if isinstance(path_or_dict, dict):
        case = path_or_dict
        name = case.get("_name", "")
    else:
        with open(path_or_dict) as f:
            case = json.load(f)
        name = str(path_or_dict)
    if "alternative_coordinates_type" not in case:
        raise ValueError("not a WHFast case (no alternative_coordinates_type)")
    u = case["universe"]
    n = int(u["n_particles"])
    if n > abi.MAX_BODIES:
        raise ValueError(f"{n} bodies exceed the maximum of {abi.MAX_BODIES}")
    ce = u["consider_effects"]
    effects = 0
    for key, bit in (("tides", abi.EFFECT_TIDES), ("rotational_flattening", abi.EFFECT_ROTATIONAL_FLATTENING),
                     ("general_relativity", abi.EFFECT_GENERAL_RELATIVITY), ("disk", abi.EFFECT_DISK),
                     ("wind", abi.EFFECT_WIND), ("evolution", abi.EFFECT_EVOLUTION)):
        if ce.get(key):
            effects |= bit
    spec = EnsembleSpec(body_offset=np.array([0, n], dtype=np.int32))
    spec.names = [name]
    spec.system = {
        "coordinates": np.array([abi.COORDINATES[case["alternative_coordinates_type"]]], dtype=np.int32),
        "effects": np.array([effects], dtype=np.uint32),
        "gr_implementation": np.array([abi.GR_IMPLEMENTATIONS[u["general_relativity_implementation"]]], dtype=np.int32),
        "time_step": np.array([case["time_step"]], dtype=np.float64),
        "time_limit": np.array([u["time_limit"]], dtype=np.float64),
        "current_time": np.array([case["current_time"]], dtype=np.float64),
        "historic_snapshot_period": np.array([case["historic_snapshot_period"]], dtype=np.float64),
        "last_historic_snapshot_time": np.array([case["last_historic_snapshot_time"]], dtype=np.float64),
        "current_iteration": np.array([case["current_iteration"]], dtype=np.int64),
    }
    body = {k: [] for k in BODY_FIELDS}
    evolvers = u.get("particles_evolvers") or []
    eo = [0]
    tabs = {k: [] for k in TABLE_FIELDS}
    for i, p in enumerate(u["particles"][:n]):
        body["mass"].append(p["mass"])
        body["radius"].append(p["radius"])
        body["radius_of_gyration_2"].append(p["radius_of_gyration_2"])
        body["moment_of_inertia"].append(p["moment_of_inertia"])
        body["position"].append(_axes(p["inertial_position"]))
        body["velocity"].append(_axes(p["inertial_velocity"]))
        body["spin"].append(_axes(p["spin"]))
        body["angular_momentum"].append(_axes(p["angular_momentum"]))
        body["heliocentric_velocity"].append(_axes(p["heliocentric_velocity"]))
        roles = 0
        sigma = scale = k2 = 0.0
        tag, payload = _enum(p["tides"]["effect"])
        if tag in ("CentralBody", "OrbitingBody"):
            model, params = _enum(payload)
            if model != "ConstantTimeLag":
                raise ValueError(f"tidal model {model} is not supported")
            roles |= abi.ROLE_TIDES_CENTRAL if tag == "CentralBody" else abi.ROLE_TIDES_ORBITING
            sigma, scale, k2 = params["dissipation_factor"], params["dissipation_factor_scale"], params["love_number"]
        k2f = 0.0
        tag, payload = _enum(p["rotational_flattening"]["effect"])
        if tag in ("CentralBody", "OrbitingBody"):
            model, params = _enum(payload)
            if model != "OblateSpheroid":
                raise ValueError(f"flattening model {model} is not supported")
            roles |= abi.ROLE_FLAT_CENTRAL if tag == "CentralBody" else abi.ROLE_FLAT_ORBITING
            k2f = params["love_number"]
        tag, payload = _enum(p["general_relativity"]["effect"])
        if tag == "CentralBody" and payload != "Disabled":
            roles |= abi.ROLE_GR_CENTRAL
        elif tag == "OrbitingBody":
            roles |= abi.ROLE_GR_ORBITING
        disk = p.get("disk")
        if disk is not None:
            tag, payload = _enum(disk["effect"])
            if tag == "CentralBody":
                roles |= abi.ROLE_DISK_CENTRAL
                spec.system["disk_properties"] = np.array([[float(payload[k]) for k in DISK_KEYS]])
            elif tag == "OrbitingBody":
                roles |= abi.ROLE_DISK_ORBITING
        wk = ws = 0.0
        if p["wind"]["effect"] == "Interaction":
            roles |= abi.ROLE_WIND
            wk = p["wind"]["parameters"]["input"]["k_factor"]
            ws = p["wind"]["parameters"]["input"]["rotation_saturation"]
        body["roles"].append(roles)
        body["tides_dissipation_factor"].append(sigma)
        body["tides_dissipation_factor_scale"].append(scale)
        body["tides_love_number"].append(k2)
        body["flattening_love_number"].append(k2f)
        body["wind_k_factor"].append(wk)
        body["wind_rotation_saturation"].append(ws)
        body["velocity_error"].append(_axes(case["inertial_velocity_errors"][i]))
        body["angular_momentum_error"].append(_axes(case["particle_angular_momentum_errors"][i]))
        body["evolution_type"].append(_evolution_code(p["evolution"]))
        ev = evolvers[i] if i < len(evolvers) else None
        if ev is not None and _evolution_code(ev["evolution"]) != abi.EVO_NON_EVOLVING:
            t = np.asarray(ev["time"], dtype=np.float64)
            tabs["evolver_time"].append(t)
            tabs["evolver_radius"].append(np.asarray(ev["radius"], dtype=np.float64))
            rg = ev.get("radius_of_gyration_2") or []
            lv = ev.get("love_number") or []
            iq = ev.get("inverse_tidal_q_factor") or []
            tabs["evolver_radius_of_gyration_2"].append(
                np.asarray(rg, dtype=np.float64) if len(rg) == len(t) else np.zeros(len(t)))
            tabs["evolver_love_number"].append(
                np.asarray(lv, dtype=np.float64) if len(lv) == len(t) else np.zeros(len(t)))
            tabs["evolver_inverse_tidal_q_factor"].append(
                np.asarray(iq, dtype=np.float64) if len(iq) == len(t) else np.zeros(len(t)))
            body["evolver_left_index"].append(int(ev.get("left_index", 0)))
            eo.append(eo[-1] + len(t))
        else:
            body["evolver_left_index"].append(0)
            eo.append(eo[-1])
    # pair-dependent dissipation map (constant_time_lag.rs:188-194): the
    # entries id * MAX_PARTICLES + host_id of orbiting bodies
    pairs = np.full(n, np.nan)
    ids = [int(p.get("id", i)) for i, p in enumerate(u["particles"][:n])]
    host = int(np.argmax(body["mass"])) if n else 0
    for key, value in (u.get("pair_dependent_scaled_dissipation_factor") or {}).items():
        owner, dep = divmod(int(key), MAX_PARTICLES)
        if dep == ids[host] and owner in ids and ids.index(owner) != host:
            pairs[ids.index(owner)] = float(value)
    body["pair_dissipation_factor"] = pairs
    # Roche radii (universe.rs:177-196): row i of the n x n block holds body
    # i's pairs; a resumed case keeps the values computed at t = 0
    roche = np.zeros((n, abi.MAX_BODIES))
    flat = u.get("roche_radiuses") or []
    if len(flat) >= n * n:
        roche[:, :n] = np.asarray(flat[:n * n], dtype=np.float64).reshape(n, n)
    body["roche_radius"] = roche
    for k, (dt, w) in BODY_FIELDS.items():
        a = np.asarray(body[k], dtype=dt)
        spec.body[k] = a.reshape(n, w) if w > 1 else a
    spec.evolver_offset = np.asarray(eo, dtype=np.int32)
    spec.tables = {k: (np.concatenate(v) if v else np.zeros(0)).astype(np.float64) for k, v in tabs.items()}
    return spec
