"""The golden cases: law texts written for this project, grids, steps and seeded initial states.

make_golden.py runs each through the reference's CPU engine and stores inputs and outputs as tests/golden/NAME.npz.
The tests read only the .npz files (the law text, grid, dt and steps travel inside them).
"""
from __future__ import annotations

import numpy as np

HEAT_LAW = """# 3-D conduction with one conductivity: the 7-point flux-form step
field T init 0 range -1000000 1000000

param k = 0.1

term conduction
  d T = diffuse(k, T)
  conserves T
  max_dt = 0.15 * h * h / k
"""

# Zero-Dirichlet conduction: each boundary face holds a ghost of -T (the wall at 0), so a cell with nb such faces loses
# 2 k nb T / h^2 on top of the zero-flux diffuse().
HEAT_DIRICHLET_LAW = """field T init 0 range -1000000 1000000
field nb init 0 range 0 6 static

param k = 0.1

term conduction
  d T = diffuse(k, T) - 2 * k * nb * T / (h * h)
  max_dt = 0.15 * h * h / k
"""

BLACK_SCHOLES_LAW = """# Black-Scholes in x = ln S along x; every (y, z) line is its own option
field V init 0 range -1 1000000

param sig = 0.2
param r = 0.05

term black_scholes
  d V = 0.5 * sig * sig * (shift(V, 1, 0, 0) - 2 * V + shift(V, -1, 0, 0)) / (h * h) + (r - 0.5 * sig * sig) * gx(V) - r * V
  max_dt = 0.4 * h * h / (sig * sig)
"""

DIFF_REACT_RK4_LAW = """# two-field diffusion-reaction (FitzHugh-Nagumo), five-point Laplacian per plane, classical RK4 through lets
field u init 0 range -1000 1000
field v init 0 range -1000 1000

param du = 0.001
param dv = 0.005
param kr = 0.005

term reaction_diffusion
  let lu0 = (shift(u, 1, 0, 0) + shift(u, -1, 0, 0) + shift(u, 0, 1, 0) + shift(u, 0, -1, 0) - 4 * u) / (h * h)
  let lv0 = (shift(v, 1, 0, 0) + shift(v, -1, 0, 0) + shift(v, 0, 1, 0) + shift(v, 0, -1, 0) - 4 * v) / (h * h)
  let fu0 = du * lu0 + u - u * u * u - kr - v
  let fv0 = dv * lv0 + u - v
  let u1 = u + 0.5 * dt * fu0
  let v1 = v + 0.5 * dt * fv0
  let fu1 = du * (shift(u1, 1, 0, 0) + shift(u1, -1, 0, 0) + shift(u1, 0, 1, 0) + shift(u1, 0, -1, 0) - 4 * u1) / (h * h) + u1 - u1 * u1 * u1 - kr - v1
  let fv1 = dv * (shift(v1, 1, 0, 0) + shift(v1, -1, 0, 0) + shift(v1, 0, 1, 0) + shift(v1, 0, -1, 0) - 4 * v1) / (h * h) + u1 - v1
  let u2 = u + 0.5 * dt * fu1
  let v2 = v + 0.5 * dt * fv1
  let fu2 = du * (shift(u2, 1, 0, 0) + shift(u2, -1, 0, 0) + shift(u2, 0, 1, 0) + shift(u2, 0, -1, 0) - 4 * u2) / (h * h) + u2 - u2 * u2 * u2 - kr - v2
  let fv2 = dv * (shift(v2, 1, 0, 0) + shift(v2, -1, 0, 0) + shift(v2, 0, 1, 0) + shift(v2, 0, -1, 0) - 4 * v2) / (h * h) + u2 - v2
  let u3 = u + dt * fu2
  let v3 = v + dt * fv2
  let fu3 = du * (shift(u3, 1, 0, 0) + shift(u3, -1, 0, 0) + shift(u3, 0, 1, 0) + shift(u3, 0, -1, 0) - 4 * u3) / (h * h) + u3 - u3 * u3 * u3 - kr - v3
  let fv3 = dv * (shift(v3, 1, 0, 0) + shift(v3, -1, 0, 0) + shift(v3, 0, 1, 0) + shift(v3, 0, -1, 0) - 4 * v3) / (h * h) + u3 - v3
  d u = (fu0 + 2 * fu1 + 2 * fu2 + fu3) / 6
  d v = (fv0 + 2 * fv1 + 2 * fv2 + fv3) / 6
  max_dt = 0.25 * h * h / max(du, dv)
"""

CRANK_NICOLSON_LAW = """# Crank-Nicolson along x through tridiag_x; the rate is (T' - T) / dt
field T init 0 range -1000000 1000000

param k = 0.01

term crank_nicolson
  let a = 0.5 * k * dt / (h * h)
  d T = (tridiag_x(0 - a, 1 + 2 * a, 0 - a, T + a * (shift(T, 1, 0, 0) - 2 * T + shift(T, -1, 0, 0))) - T) / dt
"""

ADI_YZ_LAW = """# implicit steps along y then z (tridiag_y, tridiag_z) of a field-dependent operator
field T init 0 range -1000000 1000000
field w init 1 range 0 10 static

param k = 0.05

term implicit_yz
  let a = k * w * dt / (h * h)
  let y1 = tridiag_y(0 - a, 1 + 2 * a, 0 - a, T)
  d T = (tridiag_z(0 - 0.5 * a, 1 + a, 0 - 0.5 * a, y1) - T) / dt
"""

# Implicit lines on a lattice where x-lines span three 32-cell chunks (the last ragged) and a second, partial block of
# 64 lines: one number a coefficient (uniform) and field-dependent ones, along x and along z.
TRIDIAG_LINES_LAW = """# uniform and field-dependent implicit lines along x and z
field T init 0 range -1000000 1000000
field w init 1 range 0 10 static

param k = 0.04

term lines
  let a = k * dt / (h * h)
  let b = a * w
  let x1 = tridiag_x(0 - a, 1 + 2 * a, 0 - a, T)
  let x2 = tridiag_x(0 - b, 1 + 2 * b, 0 - b, x1 + 0.1 * w)
  let z1 = tridiag_z(0 - 0.5 * b, 1 + b, 0 - 0.5 * b, x2)
  d T = (tridiag_z(0 - 0.5 * a, 1 + a, 0 - 0.5 * a, z1) - T) / dt
"""

# Every op alone on a lattice uneven on every axis (each is its own law: A moved by the op, B static input).
OPS = {
    "lap": "d A = lap(A)",
    "grad": "d A = gx(A) + gy(A) + gz(A)",
    "diffuse_field_k": "d A = diffuse(0.5 + B, A)",
    "adv": "d A = adv(A, B - 0.5, 0.3, -0.2)",
    "shift": "d A = shift(A, 1, -1, 0) - A + 0.5 * shift(A, -2, 0, 1)",
    "elementwise": "d A = step(B - 0.5) * sqrt(abs(A)) - clamp(A, 0.2, 0.4) + min(A, B) * max(A, 0.1)",
    "transcendental": "d A = abs(B) ^ 1.5 - exp(-A) + log(1 + B) + pos(A - 0.6) / (B + 1)",
    "reductions": "d A = 0.01 * (maxof(B) - minof(A)) + 0.0001 * sumof(B)",
    "transposes": "d A = 0.1 * (shift_t(A, 1, 0, -1) + gx_t(B) - gy_t(A) + gz_t(A))",
    "rim_mul_zero": "d A = rim * (A - B) + 0 * lap(A) - 0.25 * A * B",
    "fall": "d B = fall(B, A, 2) + 0.5 * fall(A, B, 1)",
    "spread": "d B = spread(B, A * 0.5, 0.3) + spread(A, B, A)",
    "topple": "d B = topple(B, A, 1.5)\n  d A = topple(A, B, 0.5 + B)",
    "distance": "d A = 0.1 * distance(B) + 0.05 * distance(A)",
    "poisson": "d A = 0.01 * poisson(B - 0.4, 5) + 0.02 * poisson(A, 3) + 0.001 * poisson(0.3, 2)",
}

# A set that moves one cell along x a step (dt 1: B becomes shift(B)), so the distance is recomputed every step, and
# a Poisson source that stops changing after the first step (its later solves are skipped).
MOVING_LAW = """field A init 0 range -1000000000 1000000000
field B init 0 range -1000000000 1000000000
field S init 0 range -1000000000 1000000000 static
term t
  d B = shift(B, 1, 0, 0) - B
  d A = 0.01 * distance(B) + 0.001 * poisson(S, 4)
"""


# Heston by ADI, one option a z-plane (its model in static fields), x = ln S along x, var = sv (y + 1/2) h along y.
# The explicit part f0 = F(V) with the mixed, x and y operators (each of the last two carrying half of -rate V); two
# implicit sweeps of increments; then a second pair: Craig-Sneyd corrects only the mixed part (cs = 1/2).
_HESTON_HEAD = """field V init 0 range -1 1000000
field var init 0 static
field kap init 1 static
field th init 0.04 static
field sv init 0.3 static
field rho init -0.5 static
field rate init 0.02 static
param cs = {theta}

term {name}
  let edge = step(sv * h - var)
  let drift = kap * (th - var) / sv - 0.5 * sv
  let xl = 0.5 * var / (h * h) - (rate - 0.5 * var) / (2 * h)
  let xu = 0.5 * var / (h * h) + (rate - 0.5 * var) / (2 * h)
  let yl = 0.5 * (var - 0.5 * sv * h) / (h * h) - (1 - edge) * drift / (2 * h)
  let yu = 0.5 * (var + 0.5 * sv * h) / (h * h) + (1 + edge) * drift / (2 * h)
  let xd = 0 - xl - xu - 0.5 * rate
  let yd = 0 - yl - yu - 0.5 * rate
  let f0 = rho * var * gx(gy(V)) + xl * (shift(V, -1, 0, 0) - V) + xu * (shift(V, 1, 0, 0) - V) + yl * (shift(V, 0, -1, 0) - V) + yu * (shift(V, 0, 1, 0) - V) - rate * V
  let p1 = tridiag_x(0 - cs * dt * xl, 1 - cs * dt * xd, 0 - cs * dt * xu, dt * f0)
  let p2 = tridiag_y(0 - cs * dt * yl, 1 - cs * dt * yd, 0 - cs * dt * yu, p1)
"""
HESTON_CS_LAW = _HESTON_HEAD.format(theta=0.5, name="heston_cs") + """  let q1 = tridiag_x(0 - cs * dt * xl, 1 - cs * dt * xd, 0 - cs * dt * xu, 0.5 * dt * rho * var * gx(gy(p2)))
  let q2 = tridiag_y(0 - cs * dt * yl, 1 - cs * dt * yd, 0 - cs * dt * yu, q1)
  d V = (p2 + q2) / dt
"""
HESTON_HV_LAW = _HESTON_HEAD.format(theta=0.7886751345948129, name="heston_hv") + """  let fp = rho * var * gx(gy(p2)) + xl * (shift(p2, -1, 0, 0) - p2) + xu * (shift(p2, 1, 0, 0) - p2) + yl * (shift(p2, 0, -1, 0) - p2) + yu * (shift(p2, 0, 1, 0) - p2) - rate * p2
  let q1 = tridiag_x(0 - cs * dt * xl, 1 - cs * dt * xd, 0 - cs * dt * xu, dt * f0 - p2 + 0.5 * dt * fp)
  let q2 = tridiag_y(0 - cs * dt * yl, 1 - cs * dt * yd, 0 - cs * dt * yu, q1)
  d V = (p2 + q2) / dt
"""

# Three options (kappa, theta, sigma, rho, r, strike, maturity, v0) on 32 x 16 cells, ln 100 +- 3 along x.
HESTON_PLANES = [(1.5, 0.04, 0.3, -0.9, 0.025, 100.0, 0.5, 0.04), (3.0, 0.08, 0.5, -0.3, 0.01, 90.0, 0.5, 0.06),
                 (0.8, 0.02, 0.2, -0.6, 0.04, 115.0, 0.5, 0.03)]


def heston_state(nx, ny, planes):
    """The law's start: each plane's cell-averaged put payoff in V, its variance lattice and model in the statics."""
    h = 6.0 / nx
    x = np.log(100.0) - 3.0 + (np.arange(nx) + 0.5) * h
    out = {k: np.zeros((len(planes), ny, nx)) for k in ("V", "var", "kap", "th", "sv", "rho", "rate")}
    for z, (kap, th, sv, rho, r, strike, _, _) in enumerate(planes):
        lk = np.log(strike)
        lo, hi = np.minimum(x - 0.5 * h, lk), np.minimum(x + 0.5 * h, lk)
        out["V"][z] = ((strike * (hi - lo) - (np.exp(hi) - np.exp(lo))) / h)[None, :]
        out["var"][z] = (sv * (np.arange(ny) + 0.5) * h)[:, None]
        for key, val in (("kap", kap), ("th", th), ("sv", sv), ("rho", rho), ("rate", r)):
            out[key][z] = val
    return {k: v.astype(np.float32).reshape(-1) for k, v in out.items()}


def op_law(expr: str) -> str:
    return ("field A init 0 range -1000000000 1000000000\nfield B init 0 range -1000000000 1000000000\n"
            "term t\n  " + expr + "\n")


MULTI_LAW = """# several equations, several targets, a conserved group, a clamped field and derives
field a init 0 range -1000 1000
field b init 0 range 0 2 clamp
field c init 0 range -1000 1000
derive s = a + b
derive s += 0.5 * c
derive g = gx(a) * gy(b)

param ka = 0.3
param kb = 0.2

term exchange
  d a = diffuse(ka, a) - 0.1 * (a - c)
  d c = 0.1 * (a - c)
  d b = diffuse(kb, b) + 0.05 * a
  d b = 0.02 * (1 - b)
  conserves a + c
  max_dt = 0.2 * h * h / max(ka, kb)

term decay
  d c = -0.01 * c * abs(s)
"""


CLAMP_LAW = """# a clamped field a term pushes past both ends of its range every step, and a clamped derived field
field a init 0 range 0 1 clamp
field b init 0 range -1000 1000
field c init 0 range 0 2 clamp
derive c = 3 * a + b

term push
  d a = 4 * (b - 0.5) + diffuse(0.05, a)
"""


def _uniform(n, seed, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    return (lo + (hi - lo) * rng.random(int(n))).astype(np.float32)


def _start_value(n):
    # A deterministic spread of values 290..1290 (a splitmix-style hash of the cell index).
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = i * np.uint64(0x9E3779B97F4A7C15)
        h ^= h >> np.uint64(31)
        h *= np.uint64(0x94D049BB133111EB)
        h ^= h >> np.uint64(29)
    return (290.0 + 1000.0 * (h >> np.uint64(40)).astype(np.float32) / np.float32(16777216.0)).astype(np.float32)


def _boundary_faces(nx, ny, nz):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    nb = np.zeros((nz, ny, nx), dtype=np.float32)
    for c, n in ((x, nx), (y, ny), (z, nz)):
        nb += (c == 0).astype(np.float32) + (c == n - 1).astype(np.float32)
    return nb.ravel()


def _sine_mode(nx, ny, nz, h, noise_seed=None, amp=1e-3):
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) * h, (np.arange(ny) + 0.5) * h, (np.arange(nx) + 0.5) * h, indexing="ij")
    t = (np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)).ravel()
    if noise_seed is not None:
        t = t + amp * np.random.default_rng(noise_seed).random(t.size)
    return t.astype(np.float32)


def _bs_payoff(nx, planes, h, strike=100.0):
    # cell-averaged put payoff max(K - e^x, 0) on x in ln 100 +- 4
    x0 = np.log(100.0) - 4.0
    edges = x0 + h * np.arange(nx + 1)
    lo, hi = edges[:-1], edges[1:]
    k = np.log(strike)
    a = np.minimum(hi, k)
    avg = np.where(a > lo, (strike * (a - lo) - (np.exp(a) - np.exp(lo))) / h, 0.0)
    return np.tile(avg, planes).astype(np.float32)


def _ops_init(n, seed):
    return {"A": _uniform(n, seed, 0.1, 0.9), "B": (np.random.default_rng(seed + 1).random(n) < 0.4).astype(np.float32)}


G_OPS = (20, 14, 9, 0.5)


def net_text(name, sizes, act, seed, out_scale=1.0, shift=None, scale=None):
    """A .net file in the reference's text format (sections sizes / act / in_shift / in_scale / out_scale / weights;
    per layer W (out x in, row-major) then b), seeded uniform weights of Glorot scale with a nonzero last layer."""
    rng = np.random.default_rng(seed)
    weights = []
    for ni, no in zip(sizes[:-1], sizes[1:]):
        lim = np.sqrt(6.0 / (ni + no))
        weights += list(rng.uniform(-lim, lim, ni * no).astype(np.float32)) + list(rng.uniform(-0.2, 0.2, no).astype(np.float32))
    shift = [0.0] * sizes[0] if shift is None else shift
    scale = [1.0] * sizes[0] if scale is None else scale
    f32 = lambda v: repr(float(np.float32(v)))  # noqa: E731  shortest text that reads back as the same float
    lines = [f"net {name}", "sizes " + " ".join(str(n) for n in sizes), f"act {act}",
             "in_shift " + " ".join(f32(v) for v in shift), "in_scale " + " ".join(f32(v) for v in scale),
             f"out_scale {f32(out_scale)}", "weights"]
    for i in range(0, len(weights), 8):
        lines.append(" ".join(f32(v) for v in weights[i:i + 8]))
    return "\n".join(lines) + "\n"


def net_law(decls: str, body: str, a_range: str = "-1000000000 1000000000") -> str:
    return (decls + "field A init 0 range " + a_range + "\nfield B init 0 range -1000000000 1000000000\n"
            "param kk = 0.3\nterm t\n" + body)


# Many ops in one law: stages (distance, poisson, a net flux, a tridiagonal solve, reductions) beside inlined stencils,
# several terms, derives and a clamped field, as a world-style law mixes them.
KITCHEN_LAW = """net g in 4 hidden 6 out 1 act relu
field rock init 0 range 0 1 static
field water init 0 range 0 1
field heat init 0 range -1000 1000
field smoke init 0 range 0 2 clamp
derive wet = water * (1 - rock)
derive wet += 0.1 * smoke

param kh = 0.2
param g0 = 1.5

term flow
  let room = 1 - rock - water
  d water = fall(water, rock, g0) + spread(water, rock, 0.1) + topple(water, rock, 1) + 0.05 * flux(g, water, heat)
  d smoke = 0.02 * room * pos(heat - 0.5) - 0.01 * smoke + 0.1 * lap(smoke)
  conserves water
  max_dt = 0.5 * h

term warm
  d heat = diffuse(kh + 0.1 * rock, heat) + adv(heat, 0.2, -0.1, 0.05) - 0.01 * distance(rock) + 0.001 * poisson(water - 0.3, 6)
  d heat = 0.001 * (maxof(water) - sumof(smoke) / 1000) + 0.01 * tridiag_x(0 - 0.1, 1.2, 0 - 0.1, heat) - 0.01 * heat
"""

# Law nets (src/field/Law.cpp:1298-1443). relu and linear layers give the CPU's bits; tanh and silu go through the
# card's tanhf / expf, held to a tolerance (atol) instead.
NET_CASES = [
    dict(name="net_mlp_relu", law=net_law("net q in 2 hidden 8 8 out 2 act relu\n",
                                          "  d B = mlp(q, A, B) + 0.1 * mlp(q, kk, 2)\n  d A = 0.5 * mlp(q, B, A)\n"),
         nets={"q": net_text("q", [2, 8, 8, 2], "relu", 31, 0.7, [0.5, 0.2], [2.0, 1.5])}),
    dict(name="net_flux_relu", law=net_law("net g in 4 hidden 8 out 1 act relu\n", "  d A = flux(g, A, B)\n", "0 1"),
         nets={"g": net_text("g", [4, 8, 1], "relu", 32, 1.3, [0.5, 0.3, 0.5, 0.3])}, dt=0.05),
    dict(name="net_flux_limited", law=net_law("net g in 2 hidden 6 out 1 act relu\n", "  d A = flux(g, A)\n", "0 1"),
         nets={"g": net_text("g", [2, 6, 1], "relu", 33, 40.0)}, dt=0.05),
    dict(name="net_ldiffuse_relu", law=net_law("net w in 5 hidden 8 out 4 act relu\n",
                                               "  d B = ldiffuse(w, kk, B)\n  d A = ldiffuse(w, 0.5 + B, A)\n"),
         nets={"w": net_text("w", [5, 8, 4], "relu", 34, 0.3)}),
    dict(name="net_flux_tanh", law=net_law("net g in 4 hidden 8 out 1 act tanh\n", "  d A = flux(g, A, B)\n"),
         nets={"g": net_text("g", [4, 8, 1], "tanh", 35)}, atol=1e-6),
    dict(name="net_ldiffuse_silu", law=net_law("net w in 5 hidden 8 out 4 act silu\n", "  d A = ldiffuse(w, kk, A)\n"),
         nets={"w": net_text("w", [5, 8, 4], "silu", 36, 0.3)}, atol=1e-6),
    # `half` nets: the CPU evaluates them in fp32 as any net; the card on its matrix cores in fp16. Held as the
    # reference holds its tensor cores to its CPU (src/tests/nets_smoke.cpp:183): 0.03 of the largest change plus 1e-3.
    dict(name="net_mlp_half", law=net_law("net q in 2 hidden 32 32 out 2 act tanh half\n",
                                          "  d B = mlp(q, A, B)\n  d A = 0.5 * mlp(q, B, A)\n"),
         nets={"q": net_text("q", [2, 32, 32, 2], "tanh", 41, 0.7, [0.5, 0.2], [2.0, 1.5])}, atol_moved=0.03),
    dict(name="net_flux_half", law=net_law("net g in 4 hidden 16 out 1 act tanh half\n", "  d A = flux(g, A, B)\n", "0 1"),
         nets={"g": net_text("g", [4, 16, 1], "tanh", 42, 1.3, [0.5, 0.3, 0.5, 0.3])}, dt=0.05, atol_moved=0.03),
    dict(name="net_ldiffuse_half", law=net_law("net w in 5 hidden 16 out 4 act relu half\n",
                                               "  d B = ldiffuse(w, kk, B)\n  d A = ldiffuse(w, 0.5 + B, A)\n"),
         nets={"w": net_text("w", [5, 16, 4], "relu", 43, 0.3)}, atol_moved=0.03),
]

# Terms with nothing to do: the CPU engine does not run a term whose every rate is exactly zero by its interval rules
# (src/field/Solver.cpp:341-347), so fields it would touch keep their bits (a -0 stays -0, where adding a +0 rate would
# make it +0). `idle_cold`: no fire anywhere, so `spread` and `settle` are idle every step; `idle_spark`: one spark,
# so they wake as the fire spreads.
IDLE_LAW = """field fire init 0 range -100 100
field T init 0 range -1000 1000
field ash init 0 range -100 100

param burn = 3.0

term spread
  d fire = diffuse(0.2, fire)
  max_dt = 0.15 * h * h / 0.2

term heat
  d T = burn * fire + diffuse(0.1, T)

term settle
  d ash = 0.5 * fire * fire
"""


# A world of many terms with the op mix of the reference's shipped world law (threshold changes of state through step,
# min, max and pos; shifts; falling, spreading and toppling matter; diffusion with a field conductivity; gradients;
# upwind advection; a distance and a Poisson solve; reductions; clamped and static fields; derives), written for this
# project. The reference holds its CUDA backend to its CPU on 12 steps of its world law on a 48x40x24 scene
# (src/tests/cuda_parity_smoke.cpp:83-121); `valley` is that check on this law and scene.
VALLEY_LAW = """field rock  init 0   range 0 1.01 static
field sand  init 0   range 0 1.01
field water init 0   range 0 1.01
field ice   init 0   range 0 1.01
field steam init 0   range 0 1.01
field fuel  init 0   range 0 1
field flame init 0   range 0 1.01 clamp
field smoke init 0   range 0 100 clamp
field ash   init 0   range 0 1.01
field T     init 290 range 0 8000
field wet   init 0   range 0 3
field near  init 0   range -1000 1000
derive wet = water + ice
derive wet += 0.5 * steam
derive near = distance(flame)

param g = 2.0
param k_rock = 0.6
param k_air = 0.05
param burn_at = 600
param burn_rate = 0.8
param heat_of_burning = 900

term settle
  let room = pos(1 - rock - sand - water - ice)
  d sand = fall(sand, rock + ice, g) + 0.2 * topple(sand, rock, 1)
  conserves sand
  max_dt = 0.4 * h / g

term pour
  d water = fall(water, rock + sand + ice, g) + spread(water, rock + sand + ice, 0.3) + 0.1 * topple(water, rock + sand, 1)
  conserves water
  max_dt = 0.4 * h / g

term conduct
  d T = diffuse(k_air + (k_rock - k_air) * (rock + sand + 0.5 * water), T)
  max_dt = 0.15 * h * h / k_rock

term burn
  let hot = step(T - burn_at)
  let rate = burn_rate * min(fuel, hot * (0.2 + flame))
  d fuel = -rate
  d flame = 2 * rate - 0.5 * flame * step(0.05 - fuel)
  d T = heat_of_burning * rate
  d ash = 0.3 * rate

term catch
  let neighbours = max(max(shift(flame, 1, 0, 0), shift(flame, -1, 0, 0)), max(shift(flame, 0, 1, 0), shift(flame, 0, -1, 0)))
  d flame = 0.5 * fuel * step(neighbours - 0.3) * step(T - 450)

term rise
  d smoke = 0.6 * flame - 0.3 * gz(smoke) - 0.02 * smoke + 0.05 * lap(smoke)
  d T = -0.1 * adv(T, 0, 0, 0.4 * flame)

term freeze
  let cold = step(273 - T)
  d ice = 0.5 * water * cold - 0.5 * ice * step(T - 273)
  d water = -0.5 * water * cold + 0.5 * ice * step(T - 273)
  conserves water + ice

term boil
  let boiling = pos(T - 373) / 100
  d steam = 0.4 * min(water, boiling) - 0.05 * steam * step(350 - T)
  d water = -0.4 * min(water, boiling) + 0.05 * steam * step(350 - T)
  d T = -40 * min(water, boiling)

term vapour
  d steam = -0.5 * gz(steam) + 0.1 * lap(steam) + 0.01 * (shift(steam, 0, 0, -1) - steam)

term wind
  d smoke = 0.05 * (maxof(T) - T) / (maxof(T) + 1) * gx(smoke) + 0.02 * gy(smoke)

term radiate
  d T = -0.0002 * pos(T - 290) * (1 - rock) + 0.01 * sqrt(pos(T - 290)) * step(0.5 - near)

term press
  d water = 0.0005 * poisson(wet - 0.4, 8) * (1 - rock)

term crumble
  d sand = 0.02 * rock * step(T - 1200) * (1 - sand)

term cool_ash
  d ash = -0.01 * ash * step(350 - T) + 0.005 * max(shift(ash, 0, 0, 1) - ash, 0)
"""


def valley_init(nx=48, ny=40, nz=24):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rock = (z < 5).astype(np.float32)
    sand = ((z >= 5) & (z < 7) & (x < 20)).astype(np.float32)
    fuel = ((z >= 5) & (z < 8) & (x >= 12) & (x < 30) & (y >= 10) & (y < 30)).astype(np.float32)
    water = ((z >= 5) & (z < 10) & (x >= 34) & (x < 44) & (y >= 4) & (y < 14)).astype(np.float32)
    ice = ((z >= 5) & (z < 7) & (x >= 40) & (y >= 30)).astype(np.float32)
    t = np.full(x.shape, 290.0)
    t[(x - 20) ** 2 + (y - 20) ** 2 + (z - 7) ** 2 <= 9] = 1400.0
    t[ice > 0] = 250.0
    t = t + np.random.default_rng(71).random(x.shape) * 0.5
    flame = np.where((x - 20) ** 2 + (y - 20) ** 2 + (z - 7) ** 2 <= 4, 0.6, 0.0)
    out = {"rock": rock, "sand": sand, "water": water, "ice": ice, "fuel": fuel, "T": t, "flame": flame}
    return {k: np.ascontiguousarray(v, np.float32).ravel() for k, v in out.items()}


def _idle_init(n, spark):
    fire = np.zeros(n, np.float32)
    if spark:
        fire[n // 2 + 3] = 5.0
    t = _uniform(n, 61, -1.0, 1.0)
    t[::7] = -0.0
    return {"fire": fire, "T": t, "ash": np.full(n, -0.0, np.float32)}


CASES = [
    dict(name="heat_20x14x9", law=HEAT_LAW, grid=(20, 14, 9, 0.5), dt=0.9, steps=4,
         params={"k": 0.1}, init={"T": _uniform(20 * 14 * 9, 1)}),
    dict(name="heat_16_k1", law=HEAT_LAW, grid=(16, 16, 16, 1.0), dt=0.1, steps=5,
         params={"k": 1.0}, init={"T": _start_value(16 * 16 * 16)}),
    dict(name="heat_32_cfg", law=HEAT_LAW, grid=(32, 32, 32, 1.0 / 32), dt=0.4 * (1.0 / 32) ** 2 / 0.1, steps=6,
         params={"k": 0.1}, init={"T": _uniform(32 ** 3, 7)}),
    dict(name="heat_64x48x1", law=HEAT_LAW, grid=(64, 48, 1, 1.0), dt=0.1, steps=7,
         params={"k": 0.7}, init={"T": _uniform(64 * 48, 3, -2.0, 5.0)}),
    dict(name="dirichlet_16", law=HEAT_DIRICHLET_LAW, grid=(16, 16, 16, 1.0 / 16), dt=0.4 * (1.0 / 16) ** 2 / 0.1, steps=20,
         params={}, init={"T": _sine_mode(16, 16, 16, 1.0 / 16, noise_seed=5), "nb": _boundary_faces(16, 16, 16)}),
    dict(name="bs_256x4", law=BLACK_SCHOLES_LAW, grid=(256, 1, 4, 8.0 / 256), dt=0.4 * (8.0 / 256) ** 2 / 0.04, steps=30,
         params={}, init={"V": _bs_payoff(256, 4, 8.0 / 256)}),
    dict(name="diffreact_rk4", law=DIFF_REACT_RK4_LAW, grid=(32, 24, 3, 2.0 / 32), dt=0.05, steps=5,
         params={}, init={"u": np.random.default_rng(8).standard_normal(32 * 24 * 3).astype(np.float32),
                          "v": np.random.default_rng(9).standard_normal(32 * 24 * 3).astype(np.float32)}),
    dict(name="crank_nicolson", law=CRANK_NICOLSON_LAW, grid=(48, 5, 3, 1.0 / 48), dt=0.02, steps=6,
         params={}, init={"T": _uniform(48 * 5 * 3, 10)}),
    dict(name="tridiag_lines", law=TRIDIAG_LINES_LAW, grid=(70, 9, 8, 0.5), dt=0.4, steps=3,
         params={}, init={"T": _uniform(70 * 9 * 8, 31), "w": _uniform(70 * 9 * 8, 32, 0.5, 2.0)}),
    dict(name="crank_nicolson_long", law=CRANK_NICOLSON_LAW, grid=(4500, 2, 1, 1.0 / 4500), dt=2e-5, steps=3,
         params={}, init={"T": _uniform(4500 * 2, 33)}),
    dict(name="adi_yz", law=ADI_YZ_LAW, grid=(6, 17, 11, 0.25), dt=0.3, steps=4,
         params={}, init={"T": _uniform(6 * 17 * 11, 12), "w": _uniform(6 * 17 * 11, 13, 0.5, 2.0)}),
    dict(name="multi", law=MULTI_LAW, grid=(18, 11, 7, 0.4), dt=0.25, steps=5, params={},
         init={"a": _uniform(18 * 11 * 7, 14, -1.0, 1.0), "b": _uniform(18 * 11 * 7, 15, 0.0, 1.9),
               "c": _uniform(18 * 11 * 7, 16, 0.0, 2.0)}),
    dict(name="clamp_every_step", law=CLAMP_LAW, grid=(16, 9, 5, 0.5), dt=0.1, steps=6, params={},
         init={"a": _uniform(16 * 9 * 5, 17), "b": _uniform(16 * 9 * 5, 18)}),
    dict(name="heston_cs_3", law=HESTON_CS_LAW, grid=(32, 16, 3, 6.0 / 32), dt=0.5 / 8, steps=8, params={},
         init=heston_state(32, 16, HESTON_PLANES)),
    dict(name="heston_hv_3", law=HESTON_HV_LAW, grid=(32, 16, 3, 6.0 / 32), dt=0.5 / 8, steps=8, params={},
         init=heston_state(32, 16, HESTON_PLANES)),
    dict(name="valley", law=VALLEY_LAW, grid=(48, 40, 24, 0.5), dt=0.25, steps=12, params={}, init=valley_init()),
    dict(name="idle_cold", law=IDLE_LAW, grid=(12, 10, 6, 0.5), dt=0.2, steps=6, params={}, init=_idle_init(12 * 10 * 6, False)),
    dict(name="idle_spark", law=IDLE_LAW, grid=(12, 10, 6, 0.5), dt=0.2, steps=6, params={}, init=_idle_init(12 * 10 * 6, True)),
] + [
    dict(name="op_" + key, law=op_law(expr), grid=G_OPS, dt=0.02, steps=4, params={},
         init=_ops_init(int(np.prod(G_OPS[:3])), 100 + i))
    for i, (key, expr) in enumerate(OPS.items())
] + [
    dict(name="kitchen", law=KITCHEN_LAW, grid=(24, 18, 12, 0.5), dt=0.1, steps=6, params={},
         nets={"g": net_text("g", [4, 6, 1], "relu", 40, 0.8, [0.5, 0.2, 0.5, 0.2])},
         init={"rock": (np.random.default_rng(41).random(24 * 18 * 12) < 0.3).astype(np.float32),
               "water": _uniform(24 * 18 * 12, 42, 0.0, 0.6), "heat": _uniform(24 * 18 * 12, 43, -1.0, 2.0),
               "smoke": _uniform(24 * 18 * 12, 44, 0.0, 0.5)}),
    dict(name="op_distance_moving", law=MOVING_LAW, grid=(12, 7, 5, 0.25), dt=1.0, steps=5, params={},
         init=dict(_ops_init(12 * 7 * 5, 150), S=_uniform(12 * 7 * 5, 151, -1.0, 1.0))),
    # x-lines longer than 2048 cells (the distance stage's outward-search form) beside short y and z lines (its
    # line-envelope form), a sparse set
    dict(name="op_distance_long", law=op_law(OPS["distance"]), grid=(2100, 3, 2, 0.5), dt=0.02, steps=2, params={},
         init={"A": _uniform(2100 * 6, 160, 0.1, 0.9),
               "B": (np.random.default_rng(161).random(2100 * 6) < 0.01).astype(np.float32)}),
] + [
    dict(grid=G_OPS, steps=4, params={}, init=_ops_init(int(np.prod(G_OPS[:3])), 200 + i), **dict(dict(dt=0.02), **case))
    for i, case in enumerate(NET_CASES)
]
