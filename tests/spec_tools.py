"""Ensemble transformations shared by the GPU tests."""
import numpy as np


def move_host(spec, rng, first=0):
    """The same systems with each system's body 0 (the star, or a binary's
    primary) moved to a random index in [first, n) of its system, the other
    bodies keeping their order.  Returns the spec and the map from new to old
    body index."""
    out = spec.copy()
    perm = np.arange(spec.n_bodies)
    for s in range(spec.n_systems):
        b0, b1 = int(spec.body_offset[s]), int(spec.body_offset[s + 1])
        local = list(range(1, b1 - b0))
        local.insert(int(rng.integers(first, b1 - b0)), 0)
        perm[b0:b1] = b0 + np.asarray(local)
    for k, v in out.body.items():
        if v is not None:
            out.body[k] = np.ascontiguousarray(v[perm])
    out.elements = {k: np.ascontiguousarray(v[perm]) for k, v in spec.elements.items()}
    return out, perm
