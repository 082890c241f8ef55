"""Array-form FP64 restatement of the GRID passes, written from the pass
semantics of the reference shaders (not from the oracle's loops).

TEST INFRASTRUCTURE ONLY: used to cross-check the C++ oracle; neither it nor
the oracle is part of the product.

Every pass works on whole planes with shifted views instead of per-cell
loops, so a wrong index, sign, direction or operand order in either
implementation shows up as a disagreement.  State layout is the oracle's
(20, ny, nx) array in the HG_P_* plane order; index [y, x].
Reference files: glsl/hydro_flux.glsl:33-166, glsl/hydro_erosion.glsl:23-92,
glsl/sediment_transport.glsl:26-93, glsl/img_interpolation.glsl:3-22,
glsl/thermal_erosion.glsl:20-109, glsl/thermal_transport.glsl:15-65,
glsl/smoothing.glsl:22-103.
"""
from __future__ import annotations

import math

import numpy as np

ROCK, DIRT, WATER, FL, FR, FT, FB, S0, S1, VX, VY, VZ = range(12)
TC = [12, 13, 14, 15]   # thermal cross L, R, T, B
TD = [16, 17, 18, 19]   # thermal diagonal LT, RT, LB, RB
OOB = float(np.float32(999999999999.0))


def at(a: np.ndarray, dx: int, dy: int, fill: float) -> np.ndarray:
    """b[y, x] = a[y + dy, x + dx], `fill` where (x + dx, y + dy) is off the grid."""
    ny, nx = a.shape
    b = np.full_like(a, fill)
    y0, y1 = max(0, -dy), min(ny, ny - dy)
    x0, x1 = max(0, -dx), min(nx, nx - dx)
    if y0 < y1 and x0 < x1:
        b[y0:y1, x0:x1] = a[y0 + dy:y1 + dy, x0 + dx:x1 + dx]
    return b


def flux(S: np.ndarray, k: dict) -> np.ndarray:
    S = S.copy()
    r, g, b = S[ROCK], S[DIRT], S[WATER]
    _, ny, nx = S.shape
    w = r + g + b
    old = {c: S[p].copy() for c, p in zip("LRTB", (FL, FR, FT, FB))}
    step = {"L": (-1, 0), "R": (1, 0), "T": (0, 1), "B": (0, -1)}
    back = {"L": "R", "R": "L", "T": "B", "B": "T"}
    new = {}
    for c, (dx, dy) in step.items():
        new[c] = np.maximum(0.0, k["E"] * old[c] + k["dt"] * (k["G"] * (w - at(w, dx, dy, OOB))))
    xs = np.arange(nx)[None, :].repeat(ny, 0)
    ys = np.arange(ny)[:, None].repeat(nx, 1)
    new["L"][xs == 0] = 0.0
    new["R"][(xs == nx - 1) & (xs != 0)] = 0.0
    new["B"][ys == 0] = 0.0
    new["T"][(ys == ny - 1) & (ys != 0)] = 0.0
    out_sum = new["L"] + new["R"] + new["T"] + new["B"]
    with np.errstate(divide="ignore", invalid="ignore"):
        K = np.where(out_sum == 0, 1.0, np.minimum(1.0, b / (out_sum * k["dt"])))
    for c in new:
        new[c] = new[c] * K
    out_sum = out_sum * K
    # inflow: each neighbour's old component pointing back at this cell
    inflow = sum(at(old[back[c]], dx, dy, 0.0) for c, (dx, dy) in step.items())
    d2 = np.maximum(0.0, b + k["dt"] * (inflow - out_sum))
    vz = b + d2
    ux = at(old["R"], -1, 0, 0.0) - old["L"] + old["R"] - at(old["L"], 1, 0, 0.0)
    uy = at(old["T"], 0, -1, 0.0) - old["B"] + old["T"] - at(old["B"], 0, 1, 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        S[VX] = np.where(vz > 0, ux / vz, 0.0)
        S[VY] = np.where(vz > 0, uy / vz, 0.0)
    S[VZ] = vz
    S[WATER] = d2
    S[FL], S[FR], S[FT], S[FB] = new["L"], new["R"], new["T"], new["B"]
    return S


def sin_slope(S: np.ndarray) -> np.ndarray:
    h = S[ROCK] + S[DIRT]
    dx = at(h, 1, 0, 0.0) - at(h, -1, 0, 0.0)
    dz = at(h, 0, 1, 0.0) - at(h, 0, -1, 0.0)
    q = dx * dx + dz * dz
    return np.sqrt(q / (q + 4.0))


def erosion(S: np.ndarray, k: dict) -> np.ndarray:
    S = S.copy()
    sa = np.maximum(float(np.float32(0.02)), sin_slope(S))
    spd = np.hypot(S[VX], S[VY])
    vz = S[VZ]
    thin = vz < float(np.float32(1e-3))
    dd = np.maximum(float(np.float32(5e-4)), vz)
    t = np.clip((dd - float(np.float32(1e-3))) / (float(np.float32(5e-4)) - float(np.float32(1e-3))), 0, 1)
    spd = np.where(thin, spd * (1 - t * t * (3 - 2 * t)), spd)
    h = {0: S[ROCK].copy(), 1: S[DIRT].copy()}
    s = {0: S[S0].copy(), 1: S[S1].copy()}
    live = np.ones_like(vz, dtype=bool)
    carried = np.zeros_like(vz)
    for i in (1, 0):
        c = np.maximum(0.0, k["Kc"] * sa * spd - carried)
        take = live & (c > s[i])
        give = live & ~(c > s[i])
        d_take = k["dt"] * k["Ks"][i] * (c - s[i])
        d_give = k["dt"] * k["Kd"][i] * (s[i] - c)
        before = h[i].copy()
        h[i] = np.where(take, h[i] - d_take, np.where(give, h[i] + d_give, h[i]))
        s[i] = np.where(take, s[i] + d_take, np.where(give, s[i] - d_give, s[i]))
        dug = take & (h[i] < 0)
        s[i] = np.where(dug, s[i] + h[i], s[i])
        h[i] = np.where(dug, 0.0, h[i])
        carried = np.where(dug, carried + before, carried)
        live = live & ~(take & ~dug)
    conv = s[0] * k["Kconv"] * k["dt"]
    S[ROCK], S[DIRT] = h[0], h[1]
    S[S0], S[S1] = s[0] - conv, s[1] + conv
    return S


def transport(S: np.ndarray, k: dict) -> np.ndarray:
    S = S.copy()
    _, ny, nx = S.shape
    xs = np.arange(nx, dtype=np.float64)[None, :] - S[VX] * k["dt"]
    ys = np.arange(ny, dtype=np.float64)[:, None] - S[VY] * k["dt"]
    xs = np.clip(xs, 0, nx - 1)
    ys = np.clip(ys, 0, ny - 1)
    ix, iy = np.floor(xs).astype(int), np.floor(ys).astype(int)
    fx, fy = xs - ix, ys - iy

    def tap(p, x, y):
        ok = (x < nx) & (y < ny)
        return np.where(ok, p[np.minimum(y, ny - 1), np.minimum(x, nx - 1)], 0.0)

    for pl in (S0, S1):
        p = S[pl].copy()
        top = tap(p, ix, iy) * (1 - fx) + tap(p, ix + 1, iy) * fx
        bot = tap(p, ix, iy + 1) * (1 - fx) + tap(p, ix + 1, iy + 1) * fx
        S[pl] = top * (1 - fy) + bot * fy
    S[WATER] = S[WATER] * (1 - k["Ke"] * k["dt"])
    return S


# direction order of the thermal outflow planes: L, R, T, B, LT, RT, LB, RB
DIRS = [(-1, 0), (1, 0), (0, 1), (0, -1), (-1, 1), (1, 1), (-1, -1), (1, -1)]


def thermal_flux(S: np.ndarray, k: dict, layer: int) -> np.ndarray:
    S = S.copy()
    D = []
    for dx, dy in DIRS:
        acc = np.zeros_like(S[ROCK])
        for l in range(layer + 1):
            hl = S[ROCK + l]
            acc = acc + (hl - at(hl, dx, dy, OOB))
        D.append(acc)
    H = np.minimum(S[ROCK + layer], np.maximum(0.0, np.max(D, axis=0)))
    Ka = k["Kalpha"][layer]
    sharp = np.ones_like(H)
    bk = np.zeros_like(H)
    marks = []
    diag = float(np.float32(math.sqrt(2.0)))
    for j, Dk in enumerate(D):
        dist = 1.0 if j < 4 else diag
        with np.errstate(invalid="ignore"):
            ang = np.arctan(Dk / dist)
        m = (Dk > 0) & (ang > Ka)
        sharp = np.where(m, np.maximum(sharp, 1 + ang - Ka), sharp)
        bk = np.where(m, bk + Dk, bk)
        marks.append(m)
    Sv = k["dt"] * k["Kspeed"][layer] * sharp ** 3 * H / 2
    for j, (Dk, m) in enumerate(zip(D, marks)):
        with np.errstate(divide="ignore", invalid="ignore"):
            S[(TC + TD)[j]] = np.where(m, Sv * Dk / bk, 0.0)
    return S


def thermal_transport(S: np.ndarray, layer: int) -> np.ndarray:
    S = S.copy()
    planes = TC + TD
    opposite = [1, 0, 3, 2, 7, 6, 5, 4]      # L<->R, T<->B, LT<->RB, RT<->LB
    gain = np.zeros_like(S[ROCK])
    for j, (dx, dy) in enumerate(DIRS):
        gain = gain + at(S[planes[opposite[j]]], dx, dy, 0.0) - S[planes[j]]
    S[ROCK + layer] = S[ROCK + layer] + gain
    return S


def smooth(S: np.ndarray, k: dict) -> np.ndarray:
    S = S.copy()
    _, ny, nx = S.shape
    r, g = S[ROCK], S[DIRT]
    nbs = [(-1, 0), (1, 0), (0, 1), (0, -1)]      # l, r, t, b
    dr = [r - at(r, dx, dy, 0.0) for dx, dy in nbs]
    dg = [(g - at(g, dx, dy, 0.0)) + d for (dx, dy), d in zip(nbs, dr)]

    def extremum(d):
        rho = np.abs((d[0] + d[1] + d[2] + d[3]) / 4)
        horiz = ((-d[0] > rho) | (-d[1] > rho)) & (d[0] * d[1] > 0)
        vert = ((-d[2] > rho) | (-d[3] > rho)) & (d[2] * d[3] > 0)
        return horiz | vert

    mean_r = (r + sum(at(r, dx, dy, 0.0) for dx, dy in nbs)) / 5
    mean_g = (g + sum(at(g, dx, dy, 0.0) for dx, dy in nbs)) / 5
    rt = np.where(extremum(dr), mean_r, r)
    gt = np.where(extremum(dg), mean_g, g)
    m = min(max(k["Kspeed"][1] * k["dt"], 0.0), 1.0)
    inner = np.zeros((ny, nx), dtype=bool)
    inner[1:-1, 1:-1] = True
    S[ROCK] = np.where(inner, m * rt + (1 - m) * r, r)
    S[DIRT] = np.where(inner, m * gt + (1 - m) * g, g)
    return S


def grid_step(S: np.ndarray, k: dict, thermal: bool = True, smoothing: bool = True) -> np.ndarray:
    S = transport(erosion(flux(S, k), k), k)
    if thermal:
        for layer in (0, 1):
            S = thermal_transport(thermal_flux(S, k, layer), layer)
    if smoothing:
        S = smooth(S, k)
    return S


def settings_dict(s) -> dict:
    """The FP32 values of an oracle ErosionSettings structure, as floats."""
    f = lambda v: float(np.float32(v))  # noqa: E731
    return {"Kc": f(s.Kc), "Kalpha": [f(s.Kalpha[0]), f(s.Kalpha[1])], "Kconv": f(s.Kconv),
            "Ks": [f(s.Ks[0]), f(s.Ks[1])], "Kd": [f(s.Kd[0]), f(s.Kd[1])], "Ke": f(s.Ke),
            "E": f(s.energy_kept), "Kspeed": [f(s.Kspeed[0]), f(s.Kspeed[1])], "G": f(s.G),
            "dt": f(s.d_t)}
