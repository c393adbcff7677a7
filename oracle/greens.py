"""The layered-media Green's tables and via kernels restated in NumPy. TEST INFRASTRUCTURE ONLY.

Follows src/mom25d/multilayer.py:58-318 (LayerStack network, GreenTableML) and src/mom25d/via.py:36-235
(sommerfeld_contour, j0_sum / hankel, Table1D, TMNetwork, ViaKernels). The real part of every Hankel sum goes
through the oracle's C hankel_real (glibc j0), the lifted part through its bessel_jn_complex, as the reference's
compiled core does (via.py:90-101). Internal PEC planes (stack.pec) are out of scope: none of the configurations
has one.
"""
import numpy as np
from scipy.interpolate import CubicSpline
from scipy.special import kv

from . import core as ocore

MU0 = 4e-7 * np.pi
C0 = 299792458.0
EPS0 = 1.0 / (MU0 * C0 ** 2)


class Layer:
    """h (m), er, tand (multilayer.py:74-88; constant permittivity)."""
    def __init__(self, h, er, tand=0.0):
        self.h, self.er, self.tand = h, er, tand

    @property
    def er_c(self):
        return self.er * (1 - 1j * self.tand)


class LayerStack:
    """Layers bottom -> top, interface k the top of layer k (multilayer.py:91-250)."""
    def __init__(self, layers, top='open', top_er=1.0, bottom='pec', bottom_er=1.0):
        self.layers, self.top, self.top_er, self.bottom, self.bottom_er = list(layers), top, top_er, bottom, bottom_er
        self.N = len(self.layers)
        self.z = np.concatenate([[0.0], np.cumsum([l.h for l in self.layers])])

    def at(self, f):
        return self

    def is_pec(self, k):
        return (k == 0 and self.bottom == 'pec') or (k == self.N and self.top == 'pec')

    @property
    def metal_interfaces(self):
        return [k for k in range(0 if self.bottom == 'open' else 1, self.N + 1 if self.top == 'open' else self.N)]

    @property
    def hmin(self):
        return min(l.h for l in self.layers)

    def eps(self, j):
        if 1 <= j <= self.N:
            return self.layers[j - 1].er_c
        return complex(self.bottom_er if j <= 0 else self.top_er)

    def eps_sum(self, k):
        return self.eps(k) + self.eps(k + 1)

    def nearest_discontinuity(self, k):
        """multilayer.py:163-169."""
        down = next((self.z[k] - self.z[j] for j in range(k - 1, -1, -1)
                     if self.is_pec(j) or self.eps(j) != self.eps(j + 1)), np.inf)
        up = next((self.z[j] - self.z[k] for j in range(k + 1, self.N + 1)
                   if self.is_pec(j) or self.eps(j) != self.eps(j + 1)), np.inf)
        return min(down, up)

    def max_er(self):
        return max([abs(l.er_c) for l in self.layers] + ([abs(self.top_er)] if self.top == 'open' else []) +
                   ([abs(self.bottom_er)] if self.bottom == 'open' else []))

    def _network(self, kr, w, mode):
        """Line admittances, tan / sec of each layer, and the admittances looking down / up (multilayer.py:183-211)."""
        k0, N = w / C0, self.N

        def y0(e):
            kz = -1j * np.sqrt(kr ** 2 - k0 ** 2 * e)
            return (kz / (w * MU0) if mode == 'h' else w * EPS0 * e / kz), kz
        Y0, tn, sc = [None] * (N + 1), [None] * (N + 1), [None] * (N + 1)
        for j in range(1, N + 1):
            Y0[j], kz = y0(self.eps(j))
            th = kz * self.layers[j - 1].h
            E = np.exp(-2j * th)
            tn[j], sc[j] = -1j * (1 - E) / (1 + E), 2 * np.exp(-1j * th) / (1 + E)
        Yd = [None] * (N + 1)
        if self.bottom == 'pec':
            Yd[1], kfirst = -1j * Y0[1] / tn[1], 2
        else:
            Yd[0], kfirst = y0(complex(self.bottom_er))[0], 1
        for k in range(kfirst, N + 1):
            Yd[k] = Y0[k] * (Yd[k - 1] + 1j * Y0[k] * tn[k]) / (Y0[k] + 1j * Yd[k - 1] * tn[k])
        Yu = [None] * (N + 1)
        if self.top == 'open':
            Yu[N], kstart = y0(complex(self.top_er))[0], N
        else:
            Yu[N - 1], kstart = -1j * Y0[N] / tn[N], N - 1
        for k in range(kstart, 0 if self.bottom == 'open' else 1, -1):
            Yu[k - 1] = Y0[k] * (Yu[k] + 1j * Y0[k] * tn[k]) / (Y0[k] + 1j * Yu[k] * tn[k])
        return Y0, tn, sc, Yd, Yu

    def voltage(self, kr, w, o, s, mode):
        """TM ('e') / TE ('h') voltage at interface o for a unit current source at s (multilayer.py:234-245)."""
        if self.is_pec(o) or self.is_pec(s):
            return np.zeros_like(np.asarray(kr, complex))
        Y0, tn, sc, Yd, Yu = self._network(kr, w, mode)
        v = 1.0 / (Yu[s] + Yd[s])
        for k in range(s + 1, o + 1):
            v = v * sc[k] / (1 + 1j * (Yu[k] / Y0[k]) * tn[k])
        for k in range(s, o, -1):
            v = v * sc[k] / (1 + 1j * (Yd[k - 1] / Y0[k]) * tn[k])
        return v

    def spectral(self, kr, w, o, s):
        """G~_A = V_h / (jw), G~_q = jw (V_e - V_h) / kr^2 (multilayer.py:247-250)."""
        Vh, Ve = self.voltage(kr, w, o, s, 'h'), self.voltage(kr, w, o, s, 'e')
        return Vh / (1j * w), 1j * w * (Ve - Vh) / kr ** 2


def sommerfeld_contour(f, stack, rho_max, K=None, nq=8):
    """Gauss-Legendre nodes on the lifted contour (via.py:36-53): (kr, dkr, K)."""
    k0 = 2 * np.pi * f / C0
    kmax = k0 * np.sqrt(stack.max_er())
    K = K or (max(1e5, 30.0 / stack.hmin) + 40 * kmax)
    t1, alpha = 1.5 * kmax, min(0.5 * k0, 4.0 / rho_max)
    wid1, wid2 = min(0.3 * k0, 0.8 * alpha), 2 * np.pi / rho_max
    b1 = np.linspace(0, t1, int(np.ceil(t1 / wid1)) + 1)
    b2 = np.append(np.arange(t1, K, wid2), K)
    br = np.concatenate([b1, b2[1:]])
    xg, wg = np.polynomial.legendre.leggauss(nq)
    a, b = br[:-1, None], br[1:, None]
    t = (0.5 * (b - a) * xg[None, :] + 0.5 * (a + b)).ravel()
    wt = (0.5 * (b - a) * wg[None, :]).ravel()
    s = np.where(t < t1, alpha * np.sin(np.pi * t / t1), 0.0)
    ds = np.where(t < t1, alpha * np.pi / t1 * np.cos(np.pi * t / t1), 0.0)
    return t + 1j * s, wt * (1 + 1j * ds), K


def j0_sum(rho, kr, W):
    """sum_k J0(rho_i kr_k) W[s, k] (via.py:90-101): real kr by hankel_real, the lifted part by complex J0."""
    rho, W = np.atleast_1d(np.asarray(rho, float)), np.atleast_2d(np.asarray(W, complex))
    real = np.abs(kr.imag) < 1e-300
    out = ocore.hankel_real(rho, np.ascontiguousarray(kr.real[real]), np.ascontiguousarray(W[:, real]))
    if (~real).any():
        out = out + (ocore.bessel_jn_complex(0, rho[:, None] * kr[None, ~real]) @ W[:, ~real].T).T
    return out


def hankel(spec, kr, dk, rho):
    """(1/2 pi) int spec J0(kr rho) kr dkr (via.py:104-108)."""
    spec = np.asarray(spec)
    out = j0_sum(rho, kr, np.atleast_2d(spec) * (kr * dk)[None, :]) / (2 * np.pi)
    return out[0] if spec.ndim == 1 else out


class Table1D:
    """Not-a-knot spline from the coarse grid to nfine uniform points, linear lookup (via.py:111-123)."""
    def __init__(self, rho_coarse, vals, rho_max, nfine=8192):
        self.rho = np.linspace(0, rho_max, nfine)
        self.v = CubicSpline(rho_coarse, vals)(self.rho)
        self._dr = self.rho[1]

    def __call__(self, r):
        u = np.minimum(np.asarray(r) / self._dr, len(self.rho) - 1.000001)
        i = u.astype(np.int64)
        t = (u - i).reshape(u.shape + (1,) * (self.v.ndim - 1))
        return self.v[i] * (1 - t) + self.v[i + 1] * t


class GreenTableML:
    """Smooth remainders of G_A, G_q between interfaces o and s and the extracted parts (multilayer.py:253-318)."""
    def __init__(self, f, stack, o, s, rho_max, nrho=260, nfine=8192, K=None, nq=8, cross_sing=False,
                 image_sing=None):
        w = 2 * np.pi * f
        self.o, self.s, self.f = o, s, f
        same, self.d = o == s, abs(stack.z[o] - stack.z[s])
        self.cA, self.cq = (MU0 / (4 * np.pi), 1.0 / (2 * np.pi * EPS0 * stack.eps_sum(o))) if same else (0.0, 0.0)
        if cross_sing and not same:
            K1 = min(25.0 / stack.hmin, 300.0 / self.d)
            ga, gq = stack.spectral(np.array([K1 + 0j]), w, o, s)
            fac = K1 * np.exp(K1 * self.d) / (2 * np.pi)
            self.cA, self.cq = complex(ga[0] * fac), complex(gq[0] * fac)
        kr, dk, K = sommerfeld_contour(f, stack, rho_max, K, nq)
        GA, Gq = stack.spectral(kr, w, o, s)
        self.dimg, self.cAi, self.cqi = 0.0, 0.0, 0.0
        if same and image_sing:
            self.dimg = 2 * stack.nearest_discontinuity(o)
            K1 = 18.0 / self.dimg
            ga, gq = stack.spectral(np.array([K1 + 0j]), 2 * np.pi, o, s)
            fac = K1 * np.exp(K1 * self.dimg) / (2 * np.pi)
            self.cAi = complex((ga[0] - MU0 / (2 * K1)) * fac).real
            self.cqi = complex((gq[0] - 1.0 / (EPS0 * stack.eps_sum(o) * K1)) * fac)
        for d, cA, cq in self.extracted():
            e = 2 * np.pi * np.exp(-kr * d) / kr
            GA, Gq = GA - cA * e, Gq - cq * e
        rho = rho_max * np.linspace(0, 1, nrho) ** 2
        self._tab = Table1D(rho, hankel(np.stack([GA, Gq]), kr, dk, rho).T, rho_max, nfine)
        self.rho, self.A_rest, self.Q_rest = self._tab.rho, self._tab.v[:, 0], self._tab.v[:, 1]
        self.npts, self.K = len(kr), K

    def extracted(self):
        if self.o != self.s:
            return [(self.d, self.cA, self.cq)] if self.cA != 0 else []
        return [(0.0, self.cA, self.cq)] + ([(self.dimg, self.cAi, self.cqi)] if self.dimg else [])

    def total(self, rho):
        GA, GQ = self.AQ(rho)
        for d, cA, cq in self.extracted():
            R = np.hypot(rho, d)
            inv = np.divide(1.0, R, out=np.zeros(np.shape(R)), where=R > 0)
            GA, GQ = GA + cA * inv, GQ + cq * inv
        return GA, GQ

    def AQ(self, r):
        v = self._tab(r)
        return v[..., 0], v[..., 1]


def core_tables(tables, layers):
    """The kernels' table dict from {(a, b): GreenTableML} (fast.py:556-568)."""
    keys = sorted(tables)
    T = [tables[k] for k in keys]
    lidx = {k: i for i, k in enumerate(layers)}
    tab = np.zeros((len(layers),) * 2, np.int32)
    for i, (a, b) in enumerate(keys):
        tab[lidx[a], lidx[b]] = tab[lidx[b], lidx[a]] = i
    get = lambda attr, dt=float: np.array([dt(getattr(t, attr)) for t in T])
    return dict(A=np.stack([t.A_rest for t in T]).astype(complex), Q=np.stack([t.Q_rest for t in T]).astype(complex),
                dr=np.array([t.rho[1] for t in T]), d=get('d'), dimg=get('dimg'), cA=get('cA', complex),
                cq=get('cq', complex), cAi=get('cAi', complex), cqi=get('cqi', complex), tab=tab,
                kind=np.array([0 if a == b else (1 if t.cA != 0 else 2) for (a, b), t in zip(keys, T)], np.int32))


def green_tables(f, stack, layers, rho_max, dmax):
    """Every pair a <= b of metal interfaces (sim.py:736-745)."""
    out = {}
    for i, a in enumerate(layers):
        for b in layers[i:]:
            xs = a != b and abs(stack.z[a] - stack.z[b]) < 3 * dmax
            im = a == b and stack.nearest_discontinuity(a) < 1.5 * dmax
            out[(a, b)] = GreenTableML(f, stack, a, b, rho_max, cross_sing=xs, image_sing=im)
    return out


# ---- via kernels (via.py:127-235) -----------------------------------------------------------------------------
class TMNetwork:
    """V[s][i]: TM voltage at interface i for a unit source at s (via.py:127-166)."""
    def __init__(self, stack, kr, w):
        self.st, self.kr, self.w, N = stack, kr, w, stack.N
        Y0, tn, sc, Yd, Yu = stack._network(kr, w, 'e')
        self.k2 = [None] + [(w / C0) ** 2 * stack.eps(j) for j in range(1, N + 1)]
        self.kz2 = [None] + [k2 - kr ** 2 for k2 in self.k2[1:]]
        self.V = {}
        for s in stack.metal_interfaces:
            v = [None] * (N + 1)
            v[s] = 1.0 / (Yu[s] + Yd[s])
            for k in range(s + 1, N + 1):
                v[k] = np.zeros_like(kr) if stack.is_pec(k) else v[k - 1] * sc[k] / (1 + 1j * (Yu[k] / Y0[k]) * tn[k])
            for k in range(s, 0, -1):
                v[k - 1] = np.zeros_like(kr) if stack.is_pec(k - 1) else \
                    v[k] * sc[k] / (1 + 1j * (Yd[k - 1] / Y0[k]) * tn[k])
            self.V[s] = v

    def is_pec(self, k):
        return self.st.is_pec(k)

    def dV(self, i, s):
        return 0.0 if self.is_pec(s) else self.V[s][i] - self.V[s][i - 1]

    def S(self, k, layers):
        return sum((self.dV(j, k) / self.kz2[j] for j in layers), 0.0)

    def P(self, li, lj, with_homog=True):
        out = 0.0
        kr2 = self.kr ** 2
        for i in li:
            for j in lj:
                out = out + kr2 / (self.kz2[i] * self.kz2[j]) * (self.dV(i, j) - self.dV(i, j - 1))
                if with_homog and i == j:
                    out = out + 1j * self.w * MU0 * self.st.layers[j - 1].h / (kr2 - self.k2[j])
        return out


def chain_ends(k_bot, k_top):
    return [(k_bot, +1.0), (k_top, -1.0)]


def _layers(span):
    return range(span[0] + 1, span[1] + 1)


class ViaKernels:
    """D tables per (interface, span), the via-via remainder R and the K0 part (via.py:179-235)."""
    def __init__(self, f, stack, spans, interfaces, rho_max, nrho=200):
        w = 2 * np.pi * f
        kr, dk, _ = sommerfeld_contour(f, stack, rho_max)
        net = TMNetwork(stack, kr, w)
        self.f, self.st, self.w, self.kr, self.dk, self.net = f, stack, w, kr, dk, net
        self._gq = {}

        def gq(a, b):
            if net.is_pec(a) or net.is_pec(b):
                return 0.0
            key = (min(a, b), max(a, b))
            if key not in self._gq:
                self._gq[key] = stack.spectral(kr, w, a, b)[1]
            return self._gq[key]
        self.gq, self.rho_max, self.D = gq, rho_max, {}
        rho = rho_max * np.linspace(0, 1, nrho) ** 2
        keys, specs = [], []
        for span in sorted(set(spans)):
            for k in interfaces:
                spec = net.S(k, _layers(span))
                for ke, div in chain_ends(*span):
                    spec = spec - div * gq(k, ke) / (1j * w)
                keys.append((k, span)); specs.append(spec * np.ones_like(kr))
        if keys:
            vals = hankel(np.array(specs), kr, dk, rho)
            for key, v in zip(keys, vals):
                self.D[key] = Table1D(rho, v, rho_max)

    def R(self, sm, sn, rho):
        spec = self.net.P(_layers(sm), _layers(sn), with_homog=False)
        for ka, da in chain_ends(*sm):
            for kb, db in chain_ends(*sn):
                spec = spec - da * db * self.gq(ka, kb) / (1j * self.w)
        return hankel(spec, self.kr, self.dk, rho)

    def R_table(self, sm, sn, nrho=200):
        rho = self.rho_max * np.linspace(0, 1, nrho) ** 2
        return Table1D(rho, self.R(sm, sn, rho), self.rho_max)

    def homog(self, sm, sn, rho):
        rho, out = np.atleast_1d(rho), np.zeros(np.size(rho), complex)
        for j in set(_layers(sm)) & set(_layers(sn)):
            kj = np.sqrt(self.net.k2[j] + 0j)
            if kj.imag > 0:
                kj = -kj
            out += 1j * self.w * MU0 * self.st.layers[j - 1].h / (2 * np.pi) * kv(0, 1j * kj * rho)
        return out
