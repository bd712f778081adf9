"""Reduced spherical-harmonic angular basis (fixed m), host-side setup.

Follows core/transform/reducedspherical/reducedsphericaltools.cpp:284-365:
lmax+1 Gauss-Legendre points in cos(theta), associated Legendre functions
normalised so that sum_q w_q P~_l P~_l' = delta (norm
sqrt((2l+1)/(4 pi) (l-|m|)!/(l+|m|)!) * sqrt(2 pi)), and the derivative table
sin(theta) d/dtheta P~_l = l cos(theta) P~_l - (l+m) c_l P~_{l-1}.
"""
import math

import numpy as np


class ReducedSphericalTools:
    def __init__(self, lmax, m=0):
        self.LMax = int(lmax)
        self.M = int(m)
        am = abs(self.M)
        nth = self.LMax + 1
        x, w = np.polynomial.legendre.leggauss(nth)
        self.ThetaGrid = np.arccos(x)
        self.Weights = w
        xx = np.cos(self.ThetaGrid)
        L = self.LMax + 1 - am
        somx2 = np.sqrt((1.0 - xx) * (1.0 + xx))
        # P_m^m, then upward recurrence in l (shtools.cpp:381+)
        pmm = np.ones(nth)
        fact = 1.0
        for _ in range(am):
            pmm = pmm * (-fact * somx2)
            fact += 2.0
        P = np.zeros((nth, L))
        prev2 = np.zeros(nth)
        prev = pmm.copy()
        for l in range(am, self.LMax + 1):
            if l == am:
                cur = pmm
            elif l == am + 1:
                cur = xx * (2.0 * am + 1.0) * pmm
            else:
                cur = (xx * (2.0 * l - 1.0) * prev - (l + am - 1.0) * prev2) / (l - am)
            if l > am:
                prev2, prev = prev, cur
            sign = (-1.0) ** self.M if self.M > 0 else 1.0
            norm = sign * math.sqrt((2.0 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            norm *= math.sqrt(2 * math.pi)
            P[:, l - am] = cur * norm
        D = np.zeros_like(P)
        for l in range(am + 1, self.LMax + 1):
            li = l - am
            c = math.sqrt((2 * l + 1.0) / (2 * l - 1.0) * (l - am) / float(l + am))
            D[:, li] = l * xx * P[:, li] - (l + self.M) * c * P[:, li - 1]
        self.AssocLegendrePoly = P
        self.AssocLegendrePolyDerivative = D

    def GetSize(self):
        return self.LMax - abs(self.M)

    def LValues(self):
        return np.arange(abs(self.M), self.LMax + 1)

    def project(self, source, pairs, differentiation):
        """<l_row| f(theta) (d) |l_col> for each pair; source shape (..., ntheta)
        (basis_reducedspherical.h:12-73)."""
        left = self.AssocLegendrePoly
        right = self.AssocLegendrePoly if differentiation == 0 else self.AssocLegendrePolyDerivative
        if differentiation not in (0, 1):
            raise RuntimeError("Reduced spherical harmonics does not support higher derivatives than 1")
        rows = pairs[:, 0]
        cols = pairs[:, 1]
        W = left[:, rows] * right[:, cols] * self.Weights[:, None]    # (ntheta, P)
        return np.tensordot(source, W, axes=([source.ndim - 1], [0]))
