"""Synthetic GRAMM fields for the terrain set-up tests: a coarse grid over a
Gaussian hill with terrain-following levels, and a smooth seeded wind."""
import numpy as np

F = np.float32


def gramm(nx=8, ny=7, nz=10, ddx=20.0, ddy=20.0, seed=0, hill=30.0, delta=(25, 15), uniform_wind=False):
    r = np.random.default_rng(seed)
    x = (np.arange(nx + 2) - nx / 2.0) * ddx
    y = (np.arange(ny + 2) - ny / 2.0) * ddy
    X, Y = np.meshgrid(x, y, indexing="ij")
    surf = (hill * np.exp(-(X ** 2 + Y ** 2) / (2 * (1.5 * ddx) ** 2)) + 100.0).astype(F)
    lev = np.cumsum(np.r_[0.0, 0.0, 4.0 * 1.3 ** np.arange(nz)])[:nz + 2].astype(F)     # level k >= 1 heights above ground
    ahe = (surf[:, :, None] + lev[None, None, :]).astype(F)                                # [nx+2][ny+2][nz+2]
    zsp = (0.5 * (ahe[:nx + 1, :ny + 1, :nz + 1] + ahe[:nx + 1, :ny + 1, 1:nz + 2])).astype(F)
    ah = ahe[:nx + 1, :ny + 1, 1].astype(F)
    prof = (2.0 + 0.15 * lev[:nz + 1]).astype(F)
    if uniform_wind:
        u = np.broadcast_to(prof, (nx + 1, ny + 1, nz + 1)).astype(F)
        v = (-0.5 * u).astype(F)
        w = (0.01 * u).astype(F)
    else:
        u = (prof[None, None, :] * (1 + 0.2 * r.standard_normal((nx + 1, ny + 1, 1)))).astype(F)
        v = (0.7 * prof[None, None, :] * (1 + 0.2 * r.standard_normal((nx + 1, ny + 1, 1)))).astype(F)
        w = (0.05 * r.standard_normal((nx + 1, ny + 1, nz + 1))).astype(F)
    return {"nx": nx, "ny": ny, "nz": nz, "ddx": ddx, "ddy": ddy, "delta_i": delta[0], "delta_j": delta[1],
            "ahe": ahe, "uwin": u, "vwin": v, "wwin": w, "zsp": zsp, "ah": ah, "zko1": 0.0, "zko2": 4.0}
