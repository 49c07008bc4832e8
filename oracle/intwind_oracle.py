"""CPU restatement of the particle model's wind interpolation on the
microscale flow field, Zeitschleife.IntWindCalculate and its BinarySearch
(Intwind.cs:31-170, :174-210, paths relative to the reference's src/).

TEST INFRASTRUCTURE ONLY: the checker of gral_pf_wind_at_particles.  The
product never imports it.

Two branches are restated, the two that read UK/VK/WK: terrain
(Program.Topo == TerrainAvailable, topo=1) and flat terrain with buildings
(topo=0).  The flat branch without buildings interpolates observation
profiles with MathF.Pow and reads no flow field; it is not part of this
path.  Float arithmetic is float32 and the reference's double expressions
float64, operation by operation: the cell origin IKOOAGRAL + (IndexI-1)*DXK
is an int plus a float, hence float; the particle coordinate is a double, so
the horizontal interpolation is evaluated in double and cast to float.
"""
from __future__ import annotations

import numpy as np

F = np.float32
D = np.float64


def binary_search(hok, nkk, height):
    """Zeitschleife.BinarySearch (Intwind.cs:174-210): the first level whose
    top HOKART[k] is at or above the height, clamped to NKK; the first probe
    is at NKK >> 3."""
    height = F(height)
    low, high = 1, nkk
    q = high >> 3
    if hok[q] >= height:
        high = q
    else:
        low = q
    while low <= high:
        mid = (low + high) >> 1
        if hok[mid] >= height:
            high = mid - 1
        else:
            low = mid + 1
    return min(low, nkk)


def wind_at_particle(uk, vk, wk, hok, dzk, nkk, dxk, dyk, ikoo, jkoo, ahmin, topo, i, j, ahint, x, y, z):
    """(UXint, UYint, UZint, IndexK) of one particle in cell (i, j)."""
    DXK, DYK = F(dxk), F(dyk)
    z = F(z)
    if topo:
        f1 = F(z - F(ahmin))
        k = binary_search(hok, nkk, f1)
    else:
        k = 1
        while k <= nkk and not (z <= hok[k]):
            k += 1
        k = min(k, nkk)
    u1, u2 = uk[i, j, k], uk[i + 1, j, k]
    ox = F(F(int(ikoo)) + F(F(i - 1) * DXK))
    ux = F(D(u1) + D(F(F(u2 - u1) / DXK)) * (D(x) - D(ox)))
    v1, v2 = vk[i, j, k], vk[i, j + 1, k]
    oy = F(F(int(jkoo)) + F(F(j - 1) * DYK))
    uy = F(D(v1) + D(F(F(v2 - v1) / DYK)) * (D(y) - D(oy)))
    w1, w2 = wk[i, j, k], wk[i, j, k + 1]
    dz = F(f1 - hok[k - 1]) if topo else F(F(z + F(ahint)) - hok[k - 1])
    uz = F(w1 + F(F(F(w2 - w1) / dzk[k]) * dz))
    return ux, uy, uz, k


def wind_at_particles(uk, vk, wk, hok, dzk, nkk, dxk, dyk, ikoo, jkoo, ahmin, topo, ci, cj, ahint, x, y, z):
    n = len(ci)
    ux, uy, uz = np.zeros(n, F), np.zeros(n, F), np.zeros(n, F)
    kk = np.zeros(n, np.int32)
    hok, dzk = np.asarray(hok, F), np.asarray(dzk, F)
    for p in range(n):
        ux[p], uy[p], uz[p], kk[p] = wind_at_particle(uk, vk, wk, hok, dzk, nkk, dxk, dyk, ikoo, jkoo, ahmin, topo,
                                                      int(ci[p]), int(cj[p]), ahint[p], x[p], y[p], z[p])
    return ux, uy, uz, kk
