"""Coupled spherical harmonics |(l1 l2) L M> of two electrons, host side.

Channel enumeration follows DefaultCoupledIndexIterator
(pyprop/CoupledSphericalHarmonics.py:15-67): M outermost, then L, l1, l2,
keeping |l1 - l2| <= L <= l1 + l2.  The Clebsch-Gordan convention is the
reference's (core/representation/coupledspherical/coupledrange.h:17-32):
<l1 m1 l2 m2 | L M> = (-1)^(M + l1 - l2) sqrt(2L + 1) (l1 l2 L; m1 m2 -M).

Wigner 3j and 6j symbols are evaluated by Racah's formulas in exact integer
arithmetic and rounded once to double.  The electron-electron multipole
coefficient uses the closed form
  <(l1' l2') L | C^k(1) . C^k(2) | (l1 l2) L>
      = (-1)^(l1 + l2' + L) <l1'||C^k||l1> <l2'||C^k||l2> {l1' l2' L; l2 l1 k},
  <l'||C^k||l> = (-1)^l' sqrt((2l'+1)(2l+1)) (l' k l; 0 0 0),
which is the value the reference obtains by its double sum over magnetic
quantum numbers (examples/tensor/helium_stabilization/coupledspherical.cpp:
53-90).  The velocity-gauge dipole couplings restate
examples/tensor/helium_stabilization/coupledvelocity.cpp:96-146 (1/r part)
and :258-276 (d/dr parts) with the one-electron matrix elements of
laserhelper.h:9-27.
"""
import math
from fractions import Fraction
from functools import lru_cache


class CoupledIndex:
    __slots__ = ("l1", "l2", "L", "M")

    def __init__(self, l1, l2, L, M):
        self.l1, self.l2, self.L, self.M = int(l1), int(l2), int(L), int(M)

    def __iter__(self):
        yield self.l1
        yield self.l2
        yield self.L
        yield self.M

    def __eq__(self, other):
        return tuple(self) == tuple(other)

    def __hash__(self):
        return hash(tuple(self))

    def __repr__(self):
        return "CoupledIndex(%i, %i, %i, %i)" % tuple(self)

    __str__ = lambda self: "l1=%i, l2=%i, L=%i, M=%i" % tuple(self)


class DefaultCoupledIndexIterator:
    """All (l1, l2, L, M) with l1, l2 <= lmax and |l1-l2| <= L <= l1+l2
    (CoupledSphericalHarmonics.py:15-67); L None means 0..lmax."""

    def __init__(self, lmax, L=None, M=[0], fullTensor=False):
        if L is None:
            L = range(lmax + 1)
        if not hasattr(L, "__iter__"):
            L = [L]
        if not hasattr(M, "__iter__"):
            M = [M]
        self.L, self.M, self.lmax, self.FullTensor = list(L), list(M), int(lmax), bool(fullTensor)

    def __iter__(self):
        for curM in self.M:
            for curL in self.L:
                for l1 in range(self.lmax + 1):
                    for l2 in range(self.lmax + 1):
                        if abs(l1 - l2) <= curL <= l1 + l2 or self.FullTensor:
                            yield CoupledIndex(l1, l2, curL, curM)

    def __repr__(self):
        # evaluable wherever sys is imported, as the reference's
        # (CoupledSphericalHarmonics.py:65-67), so Config.SetValue accepts it
        return "sys.modules['pyprop_amd'].DefaultCoupledIndexIterator(%s, L=%s, M=%s, fullTensor=%s)" % (
            self.lmax, self.L, self.M, self.FullTensor)


# ---------------------------------------------------------------- 3j / 6j
_f = math.factorial


def _tri(a, b, c):
    return a + b >= c and a + c >= b and b + c >= a


@lru_cache(maxsize=None)
def wigner_3j(j1, j2, j3, m1, m2, m3):
    """(j1 j2 j3; m1 m2 m3) for integer arguments, Racah's formula."""
    if m1 + m2 + m3 != 0 or not _tri(j1, j2, j3):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    pre = Fraction(_f(j1 + j2 - j3) * _f(j1 - j2 + j3) * _f(-j1 + j2 + j3), _f(j1 + j2 + j3 + 1))
    pre *= _f(j1 + m1) * _f(j1 - m1) * _f(j2 + m2) * _f(j2 - m2) * _f(j3 + m3) * _f(j3 - m3)
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    s = Fraction(0)
    for k in range(kmin, kmax + 1):
        den = (_f(k) * _f(j3 - j2 + k + m1) * _f(j3 - j1 + k - m2) * _f(j1 + j2 - j3 - k)
               * _f(j1 - k - m1) * _f(j2 - k + m2))
        s += Fraction((-1) ** k, den)
    sign = -1.0 if (j1 - j2 - m3) % 2 else 1.0
    # sqrt(pre) * s, rounded once: sqrt(pre * s^2) with the sign of s
    v = math.sqrt(pre * s * s)
    return sign * (v if s >= 0 else -v)


def _delta(a, b, c):
    return Fraction(_f(a + b - c) * _f(a - b + c) * _f(-a + b + c), _f(a + b + c + 1))


@lru_cache(maxsize=None)
def wigner_6j(j1, j2, j3, j4, j5, j6):
    """{j1 j2 j3; j4 j5 j6} for integer arguments, Racah's formula."""
    for t in ((j1, j2, j3), (j1, j5, j6), (j4, j2, j6), (j4, j5, j3)):
        if not _tri(*t):
            return 0.0
    pre = _delta(j1, j2, j3) * _delta(j1, j5, j6) * _delta(j4, j2, j6) * _delta(j4, j5, j3)
    a = (j1 + j2 + j3, j1 + j5 + j6, j4 + j2 + j6, j4 + j5 + j3)
    b = (j1 + j2 + j4 + j5, j2 + j3 + j5 + j6, j3 + j1 + j6 + j4)
    s = Fraction(0)
    for t in range(max(a), min(b) + 1):
        den = 1
        for x in a:
            den *= _f(t - x)
        for y in b:
            den *= _f(y - t)
        s += Fraction((-1) ** t * _f(t + 1), den)
    v = math.sqrt(pre * s * s)
    return v if s >= 0 else -v


def clebsch_gordan(l1, l2, m1, m2, L, M):
    """<l1 m1 l2 m2 | L M> in the reference's convention (coupledrange.h:21-31)."""
    sign = -1.0 if (M + l1 - l2) % 2 else 1.0
    return sign * math.sqrt(2 * L + 1.0) * wigner_3j(l1, l2, L, m1, m2, -M)


def reduced_ck(lp, k, l):
    """<l'||C^k||l> of the renormalised spherical harmonic C^k."""
    sign = -1.0 if lp % 2 else 1.0
    return sign * math.sqrt((2 * lp + 1.0) * (2 * l + 1.0)) * wigner_3j(lp, k, l, 0, 0, 0)


# ---------------------------------------------------------------- selection rules
def _pairs(channels, rule):
    """All (left, right) channel index pairs with rule(left, right), left
    outer (coupledsphericalselectionrule.h:17-51)."""
    out = []
    for i, a in enumerate(channels):
        for j, b in enumerate(channels):
            if rule(a, b):
                out.append((i, j))
    return out


def rule_diagonal(a, b):
    """CoupledSphericalSelectionRuleDiagonal (coupledsphericalselectionrule.h:176-187)."""
    return a.L == b.L and a.M == b.M and a.l1 == b.l1 and a.l2 == b.l2


def rule_r12_old(a, b):
    """CoupledSphericalSelectionRuleR12Old (coupledsphericalselectionrule.h:58-82)."""
    L = b.L
    return (a.L == b.L and a.M == b.M and L <= a.l1 + a.l2 and abs(a.l1 - a.l2) <= L
            and L <= b.l1 + b.l2 and abs(b.l1 - b.l2) <= L)


def make_rule_r12(cutoff):
    """CoupledSphericalSelectionRuleR12 with a multipole cutoff
    (coupledsphericalselectionrule.h:85-146): R12Old and a nonzero sum over
    l3 <= cutoff of the multipole coefficients."""
    def rule(a, b):
        if not rule_r12_old(a, b):
            return False
        lo, hi = multipole_range(a, b)
        return any(abs(ee_coefficient(a, b, k)) > 1e-14 for k in range(lo, min(hi, cutoff) + 1))
    return rule


def rule_linear_polarized(a, b):
    """CoupledSphericalSelectionRuleLinearPolarizedField
    (coupledsphericalselectionrule.h:149-172)."""
    if abs(a.L - b.L) == 1 and a.M == b.M:
        if abs(a.l1 - b.l1) == 1 and a.l2 == b.l2:
            return True
        if abs(a.l2 - b.l2) == 1 and a.l1 == b.l1:
            return True
    return False


def basis_pairs(channels, rule_name):
    """Angular basis pairs for a geometry name of the coupled basis
    (pyprop/tensorpotential/basis/CoupledSphericalHarmonic.py:132-154)."""
    g = rule_name.lower().strip()
    if g == "diagonal":
        return _pairs(channels, rule_diagonal)
    if g.startswith("selectionrule_"):
        name = g[len("selectionrule_"):]
        if name == "r12":
            return _pairs(channels, rule_r12_old)
        if name[:3] == "r12":
            return _pairs(channels, make_rule_r12(int(name[4:])))
        if name == "linearpolarizedfield":
            return _pairs(channels, rule_linear_polarized)
        raise ValueError("Unkonwn selection rule %s" % name)
    raise ValueError("Geometry '%s' not supported by the coupled spherical basis" % rule_name)


# ---------------------------------------------------------------- coefficients
def multipole_range(a, b):
    """l3 range of the e-e expansion between channels a (left) and b (right)
    (coupledspherical.cpp:64-65)."""
    return max(abs(b.l1 - a.l1), abs(b.l2 - a.l2)), min(b.l1 + a.l1, b.l2 + a.l2)


@lru_cache(maxsize=None)
def _ee(l1p, l2p, l1, l2, L, k):
    sign = -1.0 if (l1 + l2p + L) % 2 else 1.0
    return sign * reduced_ck(l1p, k, l1) * reduced_ck(l2p, k, l2) * wigner_6j(l1p, l2p, L, l2, l1, k)


def ee_coefficient(a, b, k):
    """Angular factor of r<^k / r>^(k+1) in <a| 1/r12 |b> (a left, b right)."""
    if a.L != b.L or a.M != b.M:
        return 0.0
    return _ee(a.l1, a.l2, b.l1, b.l2, a.L, int(k))


def _lh_C(l, m):
    return l * math.sqrt(((l + 1 + m) * (l + 1 - m)) / ((2 * l + 1.0) * (2 * l + 3)))


def _lh_D(l, m):
    return -(l + 1) * math.sqrt(((l + m) * (l - m)) / ((2 * l + 1.0) * (2 * l - 1)))


def _lh_E(l, m):
    return math.sqrt(((l + 1 + m) * (l + 1 - m)) / ((2 * l + 1.0) * (2 * l + 3)))


def _lh_F(l, m):
    return math.sqrt(((l + m) * (l - m)) / ((2 * l + 1.0) * (2 * l - 1)))


def velocity_couplings(a, b):
    """(c1, c2, e1, e2) for left channel a and right channel b: the
    velocity-gauge dipole is -i (c1/r1 + c2/r2) + -i e1 d/dr1 + -i e2 d/dr2
    (coupledvelocity.cpp:96-146, :258-276 and the r2 counterpart).  c1, c2
    combine <l'|-(sin(theta) d/dtheta + cos(theta))|l> and e1, e2 <l'|cos(theta)|l>
    of each electron, summed over the magnetic quantum numbers of the
    coupling."""
    if a.M != b.M:
        return 0.0, 0.0, 0.0, 0.0
    l1, l2, L, M = a.l1, a.l2, a.L, a.M
    l1p, l2p, Lp, Mp = b.l1, b.l2, b.L, b.M
    ls = max(l1, l1p, l2, l2p)
    c1 = c2 = e1 = e2 = 0.0
    for m in range(-ls, ls + 1):
        cg = clebsch_gordan(l1, l2, m, M - m, L, M) * clebsch_gordan(l1p, l2p, m, Mp - m, Lp, Mp)
        if cg == 0.0:
            continue
        if l2 == l2p and abs(m) <= l1 and abs(m) <= l1p:
            up, dn = l1 == l1p + 1, l1 == l1p - 1
            C = _lh_C(l1p, m) if up else 0.0
            D = _lh_D(l1p, m) if dn else 0.0
            E = _lh_E(l1p, m) if up else 0.0
            F = _lh_F(l1p, m) if dn else 0.0
            c1 += (-(C + D) - (E + F)) * cg
            e1 += (E + F) * cg
        if l1 == l1p and abs(m) <= l2 and abs(m) <= l2p:
            mm = M - m
            up, dn = l2 == l2p + 1, l2 == l2p - 1
            C = _lh_C(l2p, mm) if up else 0.0
            D = _lh_D(l2p, mm) if dn else 0.0
            E = _lh_E(l2p, mm) if up else 0.0
            F = _lh_F(l2p, mm) if dn else 0.0
            c2 += (-(C + D) - (E + F)) * cg
            e2 += (E + F) * cg
    return c1, c2, e1, e2
