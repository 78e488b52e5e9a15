"""Exact solution of the ideal-gas Riemann problem (E. F. Toro, Riemann
Solvers and Numerical Methods for Fluid Dynamics, 3rd ed., chapter 4: the
pressure function f_L + f_R + du = 0 solved by Newton's method from the
two-rarefaction guess, then sampled along x/t). Test infrastructure."""

from __future__ import annotations

import math


class RiemannProblem:
    def __init__(self, gamma: float, left, right):
        self.g = gamma
        self.rl, self.ul, self.pl = left
        self.rr, self.ur, self.pr = right
        self.cl = math.sqrt(gamma * self.pl / self.rl)
        self.cr = math.sqrt(gamma * self.pr / self.rr)
        self.p_star, self.u_star = self._star()

    def _f(self, p, r, pk, ck):
        """Toro (4.6)-(4.7): velocity jump across the wave of side k and its
        derivative with respect to p."""
        g = self.g
        if p > pk:  # shock
            A = 2.0 / ((g + 1.0) * r)
            B = (g - 1.0) / (g + 1.0) * pk
            q = math.sqrt(A / (B + p))
            return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
        e = (g - 1.0) / (2.0 * g)  # rarefaction
        return (2.0 * ck / (g - 1.0) * ((p / pk) ** e - 1.0),
                (p / pk) ** (-(g + 1.0) / (2.0 * g)) / (r * ck))

    def _star(self):
        g = self.g
        du = self.ur - self.ul
        e = (g - 1.0) / (2.0 * g)
        # two-rarefaction approximation (Toro 4.46) as the starting value
        p = ((self.cl + self.cr - 0.5 * (g - 1.0) * du)
             / (self.cl / self.pl ** e + self.cr / self.pr ** e)) ** (1.0 / e)
        for _ in range(100):
            fl, dl = self._f(p, self.rl, self.pl, self.cl)
            fr, dr = self._f(p, self.rr, self.pr, self.cr)
            p_new = max(p - (fl + fr + du) / (dl + dr), 1e-14)
            if abs(p_new - p) <= 1e-15 * (p_new + p):
                p = p_new
                break
            p = p_new
        fl, _ = self._f(p, self.rl, self.pl, self.cl)
        fr, _ = self._f(p, self.rr, self.pr, self.cr)
        return p, 0.5 * (self.ul + self.ur) + 0.5 * (fr - fl)

    def density(self, s: float) -> float:
        """Density at similarity coordinate s = x / t (Toro 4.4 sampling)."""
        g, ps, us = self.g, self.p_star, self.u_star
        if s <= us:
            r, u, p, c, sgn = self.rl, self.ul, self.pl, self.cl, 1.0
        else:
            r, u, p, c, sgn = self.rr, self.ur, self.pr, self.cr, -1.0
        if ps > p:  # shock on this side
            shock = u - sgn * c * math.sqrt((g + 1.0) / (2.0 * g) * ps / p + (g - 1.0) / (2.0 * g))
            if sgn * (s - shock) <= 0.0:
                return r
            k = (g - 1.0) / (g + 1.0)
            return r * (ps / p + k) / (k * ps / p + 1.0)
        c_star = c * (ps / p) ** ((g - 1.0) / (2.0 * g))
        head, tail = u - sgn * c, us - sgn * c_star
        if sgn * (s - head) <= 0.0:
            return r
        if sgn * (s - tail) >= 0.0:
            return r * (ps / p) ** (1.0 / g)
        w = 2.0 / (g + 1.0) + sgn * (g - 1.0) / ((g + 1.0) * c) * (u - s)
        return r * w ** (2.0 / (g - 1.0))

    def cell_average_density(self, a: float, b: float, x0: float, t: float, n: int = 400):
        """Mean density over [a, b] by the n-point composite midpoint rule."""
        h = (b - a) / n
        return sum(self.density((a + (i + 0.5) * h - x0) / t) for i in range(n)) / n
