"""Exact solution of the 1D Riemann problem for two stiffened gases (Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", ch. 4, with p replaced by p + p_inf on each side).
Independent of the oracle and the kernels: used to pin them against the mathematics."""
import math


class Side:
    def __init__(self, rho, u, p, gamma, pinf):
        self.rho, self.u, self.p, self.g, self.pinf = rho, u, p, gamma, pinf
        self.c = math.sqrt(gamma * (p + pinf) / rho)

    def f(self, ps):
        g, pi = self.g, self.pinf
        if ps > self.p:  # shock
            A = 2.0 / ((g + 1.0) * self.rho)
            B = (g - 1.0) / (g + 1.0) * (self.p + pi)
            q = math.sqrt(A / (ps + pi + B))
            return (ps - self.p) * q, q * (1.0 - 0.5 * (ps - self.p) / (ps + pi + B))
        r = (ps + pi) / (self.p + pi)
        e = (g - 1.0) / (2.0 * g)
        return (2.0 * self.c / (g - 1.0)) * (r ** e - 1.0), (1.0 / (self.rho * self.c)) * r ** (-(g + 1.0) / (2.0 * g))


def star(L, R):
    p = max(1e-12 - min(L.pinf, R.pinf), 0.5 * (L.p + R.p))
    lo = -min(L.pinf, R.pinf) + 1e-14
    for _ in range(200):
        fl, dl = L.f(p)
        fr, dr = R.f(p)
        F = fl + fr + R.u - L.u
        pn = p - F / (dl + dr)
        if pn <= lo:
            pn = 0.5 * (p + lo)
        if abs(pn - p) <= 1e-14 * max(1.0, abs(p)):
            p = pn
            break
        p = pn
    fl, _ = L.f(p)
    fr, _ = R.f(p)
    return p, 0.5 * (L.u + R.u) + 0.5 * (fr - fl)


def sample(L, R, ps, us, s):
    """(rho, u, p, side) at similarity coordinate s = x/t; side 0 = left material, 1 = right."""
    if s <= us:
        K, sgn, side = L, -1.0, 0
    else:
        K, sgn, side = R, 1.0, 1
    g, pi = K.g, K.pinf
    if ps > K.p:  # shock
        r = (ps + pi) / (K.p + pi)
        gm = (g - 1.0) / (g + 1.0)
        rho_s = K.rho * (r + gm) / (gm * r + 1.0)
        S = K.u + sgn * K.c * math.sqrt((g + 1.0) / (2.0 * g) * r + (g - 1.0) / (2.0 * g))
        if (sgn < 0 and s <= S) or (sgn > 0 and s >= S):
            return K.rho, K.u, K.p, side
        return rho_s, us, ps, side
    # rarefaction
    rho_s = K.rho * ((ps + pi) / (K.p + pi)) ** (1.0 / g)
    c_s = math.sqrt(g * (ps + pi) / rho_s)
    head = K.u + sgn * K.c
    tail = us + sgn * c_s
    if (sgn < 0 and s <= head) or (sgn > 0 and s >= head):
        return K.rho, K.u, K.p, side
    if (sgn < 0 and s >= tail) or (sgn > 0 and s <= tail):
        return rho_s, us, ps, side
    # inside the fan
    u = 2.0 / (g + 1.0) * (-sgn * K.c + (g - 1.0) / 2.0 * K.u + s)
    c = -sgn * (u - s)
    rho = K.rho * (c / K.c) ** (2.0 / (g - 1.0))
    p = (K.p + pi) * (c / K.c) ** (2.0 * g / (g - 1.0)) - pi
    return rho, u, p, side


def solve(L, R, x0, t, xs):
    ps, us = star(L, R)
    return [sample(L, R, ps, us, (x - x0) / t) for x in xs]
