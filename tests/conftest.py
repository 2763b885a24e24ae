import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs on the MI355X box)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def golden():
    import json
    return json.loads((GOLDEN / "digests.json").read_text())


def load_fixture(name):
    import numpy as np
    return dict(np.load(GOLDEN / f"{name}.npz"))


def tiny_scene(n, rng):
    """n random triangles in [-1, 1]^3 and 4,000 rays, every other one aimed at a triangle's
    centroid."""
    import numpy as np
    from phantasyenginetestbed_amd import phx
    tris = np.zeros(n, phx.TRI)
    tris["v"][:, :, :3] = rng.uniform(-1, 1, (n, 3, 3)).astype(np.float32)
    rays = np.zeros(4000, phx.RAY)
    rays["o"] = rng.uniform(-2, 2, (len(rays), 3)).astype(np.float32)
    # every other ray aimed at the centroid of a random triangle, so many of them hit
    target = tris["v"][rng.integers(0, n, len(rays)), :, :3].mean(axis=1)
    d = rng.normal(size=(len(rays), 3))
    d[::2] = target[::2] - rays["o"][::2]
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"] = np.float32(1e-4)
    rays["tmax"] = np.float32(phx.FLT_MAX)
    return tris, rays


def degenerate_rays(fx, rng):
    """Rays the API accepts but no camera makes: zero, NaN and infinite directions or origins,
    subnormal and negative-zero direction components, empty and inverted [tMin, tMax], tMax 0,
    negative or NaN, NaN tMin. Each case is applied to 64 rays of fixture `fx` that hit as they
    are; the result is shuffled so the cases mix within waves."""
    import numpy as np
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    cases = [  # (field, component or None for all, value)
        [("d", None, 0.0)], [("d", 1, nan)], [("o", 0, nan)], [("d", 2, inf)], [("o", 1, -inf)],
        [("d", 0, np.float32(1e-41))], [("d", 1, np.float32(-0.0))],
        [("tmin", None, 2.0), ("tmax", None, 1.0)], [("tmax", None, 0.0)],
        [("tmax", None, -1.0)], [("tmin", None, nan)], [("tmax", None, nan)],
        [("tmax", None, inf)], [("o", None, np.float32(3e38))]]
    base = np.ascontiguousarray(fx["rays"][fx["hits"]["tri"] != 0xFFFFFFFF][:64])
    out = []
    for case in cases:
        r = base.copy()
        for field, comp, value in case:
            if comp is None:
                r[field] = value
            else:
                r[field][:, comp] = value
        out.append(r)
    rays = np.concatenate(out)
    return rays[rng.permutation(len(rays))]
