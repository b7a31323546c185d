"""Helpers shared by the tests: matching oracle / device setups and seeded
physical states.  Test infrastructure only."""
import numpy as np

import freeshearlayer_amd as fsl
from oracle_lib import Oracle

FIELD_NAMES = ("rho", "rho_u", "rho_v", "rho_w", "E")


def random_state(L, H, D, seed=0, smooth=True, with_ghosts=True):
    """A physical conserved state: rho ~ 1 +- 10 %, u ~ 150 +- 60, v, w ~ +-25,
    p ~ 1e5 +- 10 %.  smooth adds long-wave structure so the limiter sees both
    smooth and sharp regions."""
    rng = np.random.default_rng(seed)
    shp = (L + 2, H + 2, D + 2)
    x = np.linspace(0, 2 * np.pi, shp[0])[:, None, None]
    y = np.linspace(0, 2 * np.pi, shp[1])[None, :, None]
    z = np.linspace(0, 2 * np.pi, shp[2])[None, None, :]
    base = np.sin(x) * np.cos(y) * np.sin(2 * z) if smooth else 0.0
    rho = 1.0 + 0.05 * base + 0.05 * rng.uniform(-1, 1, shp)
    u = 150.0 + 40.0 * base + 20.0 * rng.uniform(-1, 1, shp)
    v = 15.0 * np.cos(x + y) + 10.0 * rng.uniform(-1, 1, shp)
    w = 15.0 * np.sin(y + z) + 10.0 * rng.uniform(-1, 1, shp)
    p = 1e5 * (1.0 + 0.05 * base + 0.05 * rng.uniform(-1, 1, shp))
    E = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    U = np.stack([rho, rho * u, rho * v, rho * w, E])
    if not with_ghosts:
        U[:, 0, :, :] = 0; U[:, -1, :, :] = 0
        U[:, :, 0, :] = 0; U[:, :, -1, :] = 0
        U[:, :, :, 0] = 0; U[:, :, :, -1] = 0
    return np.ascontiguousarray(U)


def make_pair(L, H, D, dynamic=1, nstages=3, seed=None, **cfg_over):
    """An oracle and a device context set up identically (reference case
    items, reference inflow).  When seed is given both start from the same
    random state, else from the reference two-stream state."""
    orc = Oracle(L, H, D, dynamic=dynamic)
    orc.set("nStages", nstages)
    # the reference writes past the inflow face when BL_HIG > HIG - HIG/2
    bl = min(5, H - H // 2)
    orc.set("BL_HIG", bl)
    case = fsl.case_reference(L, H, D, nStages=nstages, BL_HIG=bl)
    cfg = fsl.config(sgs_model=fsl.FSL_SGS_DYNAMIC if dynamic else fsl.FSL_SGS_CONSTANT, **cfg_over)
    ctx = fsl.Context(case, cfg)
    if seed is None:
        orc.ic_reference()
        U = orc.state()
    else:
        U = random_state(L, H, D, seed)
        orc.set_state(U)
    ctx.upload_state(U)
    ctx.set_lcg(orc.get("X"), 0)
    return orc, ctx


def rel_maxnorm(a, b):
    """max |a - b| / max |a| per conserved field."""
    return [float(np.max(np.abs(a[f] - b[f])) / max(np.max(np.abs(a[f])), 1e-300))
            for f in range(a.shape[0])]


def assert_bitwise(a, b, what=""):
    """Equal bit patterns: floating-point arrays are compared as integers of
    their width, so a signed zero or a NaN payload that differs fails too."""
    a = np.asarray(a); b = np.asarray(b)
    if a.shape != b.shape:
        raise AssertionError("%s: shapes %s and %s" % (what, a.shape, b.shape))
    if a.dtype == b.dtype and a.dtype.kind == "f":
        ui = {2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
        if np.array_equal(a.view(ui), b.view(ui)):
            return
        if np.array_equal(a, b, equal_nan=True):
            n = int(np.count_nonzero(a.view(ui) != b.view(ui)))
            raise AssertionError("%s: equal values but %d entries differ in the bits of a signed zero or "
                                 "a NaN" % (what, n))
    elif np.array_equal(a, b):
        return
    d = np.abs(a - b)
    idx = np.unravel_index(np.nanargmax(d), d.shape)
    raise AssertionError("%s differs: %d entries, max |diff| %.3e at %s (%.17g vs %.17g)"
                         % (what, int(np.count_nonzero(a != b)), np.nanmax(d), idx, a[idx], b[idx]))
