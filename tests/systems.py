"""Small seeded systems shared by the tests (sizes the oracle finishes in seconds)."""
import numpy as np

from rbmd_amd import make_salt, make_water, salt_config, water_config

_cache = {}


def salt(n=4000, seed=3):
    key = ("salt", n, seed)
    if key not in _cache:
        _cache[key] = make_salt(n_ions=n, seed=seed)
    return _cache[key]


def water(nmol=600, seed=5):
    key = ("water", nmol, seed)
    if key not in _cache:
        _cache[key] = make_water(n_mol=nmol, seed=seed)
    return _cache[key]


def salt_run(**kw):
    return salt_config(**kw)


def water_run(cut=10.0, **kw):
    # the 600-molecule box is 26.2 A: rc = 10 A stays below L/2, which the
    # reference's single-image minimum image needs, and dims = 2 exercises
    # stencils that visit one cell twice through different images
    kw.setdefault("cut_off", cut)
    kw.setdefault("r_core", cut)
    return water_config(**kw)


def rel_err(a, b):
    """max_i |a_i - b_i| / max(|b_i|, rms(b)) over atoms (vector norms)."""
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    nb = np.linalg.norm(b, axis=1)
    rms = np.sqrt(np.mean(nb ** 2))
    if rms == 0:
        return float(np.abs(a - b).max())
    den = np.maximum(nb, rms)
    return float((np.linalg.norm(a - b, axis=1) / den).max())
