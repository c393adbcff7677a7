"""Layered stack, its spectral transmission-line network and the per-frequency Green's tables, built on the GPU.

  Layer, LayerStack        src/mom25d/multilayer.py:58-250 (constant or Djordjevic-Sarkar permittivity)
  sommerfeld_contour       src/mom25d/via.py:36-53
  GreenTableML             src/mom25d/multilayer.py:253-318
  green_tables             src/mom25d/sim.py:736-745
  RunLog                   src/mom25d/multilayer.py:30-53

The network is evaluated on the device in complex128 (PyTorch tensors over the contour's nodes); the Hankel
transforms J0 sums run in the HIP kernel m25x_hankel; the not-a-knot cubic spline from the 260-point quadratic
grid to the 8192-point uniform one (via.py:111-123) is a fixed linear map of the grid alone, precomputed once per
grid as a matrix and applied on the device. Internal PEC planes of the stack (LayerStack(pec=...)) are not
supported: none of this build's boards has one.
"""
import copy
import time
from contextlib import contextmanager

import numpy as np
import torch
from scipy.interpolate import CubicSpline

from .base import C0, EPS0, MU0
from . import device as dv

DS_M1, DS_M2 = 3.0, 12.0


class RunLog:
    """Stage timings and notes (records: [(stage, seconds, detail)]); printed live when verbose."""

    def __init__(self, verbose=True):
        self.verbose, self.records, self.notes = verbose, [], []

    @contextmanager
    def timer(self, stage, detail='', sync=False):
        if sync and torch.cuda.is_available():
            torch.cuda.synchronize()
        t = time.perf_counter()
        yield
        if sync and torch.cuda.is_available():
            torch.cuda.synchronize()
        dt = time.perf_counter() - t
        self.records.append((stage, dt, detail))
        if self.verbose:
            print(f"  [{dt:7.3f} s] {stage}" + (f"  ({detail})" if detail else ''))

    def note(self, msg):
        self.notes.append(msg)
        if self.verbose:
            print("  " + msg)

    def totals(self):
        d = {}
        for s, t, _ in self.records:
            d[s] = d.get(s, 0.0) + t
        return d


def _ds_k(f):
    w = 2 * np.pi * np.asarray(f, float)
    return np.log10((2 * np.pi * 10 ** DS_M2 + 1j * w) / (2 * np.pi * 10 ** DS_M1 + 1j * w)) / (DS_M2 - DS_M1)


def djordjevic_sarkar(er, tand, fref, f):
    """The wideband Debye permittivity at f, er (1 - j tand) at fref (multilayer.py:65-71)."""
    kr = _ds_k(fref)
    d_eps = -er * tand / kr.imag
    return er - d_eps * kr.real + d_eps * _ds_k(f)


class Layer:
    """A dielectric layer: h (m), er, tand; fref (Hz) for a wideband layer (multilayer.py:74-88)."""

    def __init__(self, h, er, tand=0.0, name=None, fref=None):
        self.h, self.er, self.tand, self.name, self.fref = h, er, tand, name, fref

    @property
    def er_c(self):
        return self.er * (1 - 1j * self.tand)

    def er_at(self, f):
        return self.er_c if self.fref is None else complex(djordjevic_sarkar(self.er, self.tand, self.fref, f))


class LayerStack:
    """Dielectric layers bottom -> top, interface k the top of layer k; a PEC ground at z = 0 (bottom='pec') or a
    half-space below; an open (half-space) or PEC top (multilayer.py:91-250)."""

    def __init__(self, layers, top='open', top_er=1.0, bottom='pec', bottom_er=1.0, pec=()):
        if top not in ('open', 'pec') or bottom not in ('open', 'pec'):
            raise ValueError("top / bottom: 'open' or 'pec'")
        if tuple(pec):
            raise ValueError("PEC planes inside the stack are not supported by this build")
        self.layers, self.top, self.top_er, self.bottom, self.bottom_er = list(layers), top, top_er, bottom, bottom_er
        self.N = len(self.layers)
        self.pec = ()
        self.z = np.concatenate([[0.0], np.cumsum([l.h for l in self.layers])])
        for i, l in enumerate(self.layers):
            if l.name is None:
                l.name = f"D{i + 1}"

    @property
    def dispersive(self):
        return any(l.fref is not None for l in self.layers)

    def at(self, f):
        if not self.dispersive:
            return self
        out = copy.copy(self)
        out.layers = []
        for l in self.layers:
            e = l.er_at(f)
            out.layers.append(Layer(l.h, e.real, -e.imag / e.real, l.name))
        out.f = f
        return out

    @property
    def metal_interfaces(self):
        return [k for k in range(0 if self.bottom == 'open' else 1, self.N + 1 if self.top == 'open' else self.N)]

    def is_pec(self, k):
        return (k == 0 and self.bottom == 'pec') or (k == self.N and self.top == 'pec')

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
        down = next((self.z[k] - self.z[j] for j in range(k - 1, -1, -1)
                     if self.is_pec(j) or self.eps(j) != self.eps(j + 1)), np.inf)
        up = next((self.z[j] - self.z[k] for j in range(k + 1, self.N + 1)
                   if self.is_pec(j) or self.eps(j) != self.eps(j + 1)), np.inf)
        return min(down, up)

    def max_er(self):
        return max([abs(l.er_c) for l in self.layers] + ([abs(self.top_er)] if self.top == 'open' else []) +
                   ([abs(self.bottom_er)] if self.bottom == 'open' else []))

    def key(self):
        """The parameters the spectral functions depend on (at this stack's frequency)."""
        return (self.top, self.bottom, complex(self.top_er), complex(self.bottom_er),
                tuple((complex(l.er_c), float(l.h)) for l in self.layers))

    # ---- the spectral network (kr: a complex128 device tensor, or a NumPy array for a few host points) ------
    def _network(self, kr, w, mode):
        """The layers' admittances and transfer factors at kr. On the device the result is kept on the kr tensor
        for the stack's parameters, w and the mode: every table (and the via kernels) of a frequency shares its
        contour, so each network is evaluated once per frequency rather than once per interface pair."""
        xp = torch if isinstance(kr, torch.Tensor) else np
        key = (float(w), mode) + self.key()
        cache = None
        if xp is torch:
            cache = getattr(kr, '_m25x_networks', None)
            if cache is None:
                cache = {}
                kr._m25x_networks = cache
            if key in cache:
                return cache[key]
        out = self._network_eval(kr, w, mode, xp)
        if cache is not None:
            cache[key] = out
        return out

    def _layer_factors(self, kr, w, xp):
        """Per layer j: kz_j and the transfer factors tn_j, sc_j, which the two modes share (kept on a device kr
        tensor with the networks)."""
        cache = getattr(kr, '_m25x_networks', None) if xp is torch else None
        key = ('layers', float(w)) + self.key()
        if cache is not None and key in cache:
            return cache[key]
        k0, N = w / C0, self.N
        kz, tn, sc = [None] * (N + 1), [None] * (N + 1), [None] * (N + 1)
        for j in range(1, N + 1):
            kz[j] = -1j * xp.sqrt(kr * kr - k0 * k0 * self.eps(j))
            th = kz[j] * self.layers[j - 1].h
            E = xp.exp(-2j * th)
            tn[j], sc[j] = -1j * (1 - E) / (1 + E), 2 * xp.exp(-1j * th) / (1 + E)
        if xp is torch:
            if cache is None:
                cache = kr._m25x_networks = {}
            cache[key] = (kz, tn, sc)
        return kz, tn, sc

    def _network_eval(self, kr, w, mode, xp):
        k0, N = w / C0, self.N

        def y0(e):
            kz = -1j * xp.sqrt(kr * kr - k0 * k0 * e)
            return (kz / (w * MU0) if mode == 'h' else w * EPS0 * e / kz), kz
        kzs, tn, sc = self._layer_factors(kr, w, xp)
        Y0 = [None] * (N + 1)
        for j in range(1, N + 1):
            e, kz = self.eps(j), kzs[j]
            Y0[j] = kz / (w * MU0) if mode == 'h' else w * EPS0 * e / kz
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
        if self.is_pec(o) or self.is_pec(s):
            return torch.zeros_like(kr) if isinstance(kr, torch.Tensor) else np.zeros_like(kr)
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
        return Vh / (1j * w), 1j * w * (Ve - Vh) / (kr * kr)

    def spectral1(self, k, w, o, s):
        """The spectral kernels at one real kr (a Python complex pair), evaluated on the host."""
        ga, gq = self.spectral(np.array([complex(k)]), w, o, s)
        return complex(ga[0]), complex(gq[0])


def sommerfeld_contour(f, stack, rho_max, K=None, nq=8):
    """(kr, dkr, K): Gauss-Legendre nodes of the contour lifted by alpha sin(pi t / t1) to t1 = 1.5 k_max, then
    real to K (via.py:36-53). Host NumPy: geometry of the integral, a few thousand nodes."""
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


_CONTOURS = {}


def device_contour(f, stack, rho_max):
    """(kr, dkr, K) of sommerfeld_contour on the device, kept for the last few (f, stack, rho_max): the tables and
    the via kernels of a frequency share one contour tensor, and with it the spectral networks kept on it."""
    st = stack.at(f)
    key = (float(f), float(rho_max), st.key(), torch.cuda.current_device())
    c = _CONTOURS.get(key)
    if c is None:
        kr_h, dk_h, K = sommerfeld_contour(f, st, rho_max)
        c = (dv.on_dev(kr_h, dv.c128), dv.on_dev(dk_h, dv.c128), K)
        if len(_CONTOURS) >= 4:
            _CONTOURS.clear()
        _CONTOURS[key] = c
    return c


_SPLINES = {}


def spline_operator(rho_max, nrho, nfine):
    """The (nfine x nrho) matrix S with S @ y = CubicSpline(rho_max u^2, y)(uniform nfine grid) (not-a-knot):
    the spline is linear in its data, so its action is fixed by the grid. Built once per grid on the host,
    kept on the device."""
    key = (float(rho_max), int(nrho), int(nfine), torch.cuda.current_device())
    S = _SPLINES.get(key)
    if S is None:
        rho = rho_max * np.linspace(0, 1, nrho) ** 2
        fine = np.linspace(0, rho_max, nfine)
        S = torch.from_numpy(np.ascontiguousarray(CubicSpline(rho, np.eye(nrho))(fine))).to(dv.device())
        _SPLINES[key] = S
    return S


def resample(S, vals):
    """S (nfine x n real) applied to complex vals (n, m) on the device."""
    return torch.complex(S @ vals.real.contiguous(), S @ vals.imag.contiguous())


def hankel(spec, kr, dk, rho):
    """(1/2 pi) int spec(kr) J0(kr rho) kr dkr for spec (ns, npts) on the device (via.py:104-108)."""
    W = (spec * (kr * dk)[None, :]).contiguous()
    return dv.hankel_sum(rho, kr, W) / (2 * np.pi)


class GreenTableML:
    """The smooth remainders A_rest(rho), Q_rest(rho) on the uniform grid of nfine points between interfaces o and
    s, and the extracted parts c / sqrt(rho^2 + d^2) (multilayer.py:253-318), computed on the device."""

    def __init__(self, f, stack, o, s, rho_max, nrho=260, nfine=8192, K=None, nq=8, cross_sing=False,
                 image_sing=None, contour=None):
        w = 2 * np.pi * f
        stack = stack.at(f)
        self.o, self.s, self.f = o, s, f
        same, self.d = o == s, abs(stack.z[o] - stack.z[s])
        self.cA, self.cq = (MU0 / (4 * np.pi), 1.0 / (2 * np.pi * EPS0 * stack.eps_sum(o))) if same else (0.0, 0.0)
        if cross_sing and not same:
            K1 = min(25.0 / stack.hmin, 300.0 / self.d)
            ga, gq = stack.spectral1(K1 + 0j, w, o, s)
            fac = K1 * np.exp(K1 * self.d) / (2 * np.pi)
            self.cA, self.cq = complex(ga * fac), complex(gq * fac)
        if contour is None:
            kr_h, dk_h, K = sommerfeld_contour(f, stack, rho_max, K, nq)
            kr, dk = dv.on_dev(kr_h, dv.c128), dv.on_dev(dk_h, dv.c128)
        else:
            kr, dk, K = contour
        GA, Gq = stack.spectral(kr, w, o, s)
        self.dimg, self.cAi, self.cqi = 0.0, 0.0, 0.0
        if same and image_sing:
            self.dimg = 2 * stack.nearest_discontinuity(o)
            K1 = 18.0 / self.dimg
            ga, gq = stack.spectral1(K1 + 0j, 2 * np.pi, o, s)
            fac = K1 * np.exp(K1 * self.dimg) / (2 * np.pi)
            self.cAi = complex((ga - MU0 / (2 * K1)) * fac).real
            self.cqi = complex((gq - 1.0 / (EPS0 * stack.eps_sum(o) * K1)) * fac)
        for d, cA, cq in self.extracted():
            e = 2 * np.pi * torch.exp(-kr * d) / kr
            GA, Gq = GA - cA * e, Gq - cq * e
        rho = dv.on_dev(rho_max * np.linspace(0, 1, nrho) ** 2, dv.f64)
        coarse = hankel(torch.stack([GA, Gq]), kr, dk, rho).T          # (nrho, 2)
        fine = resample(spline_operator(rho_max, nrho, nfine), coarse)  # (nfine, 2)
        self.rho = np.linspace(0, rho_max, nfine)
        self.dr = self.rho[1]
        self.A_rest, self.Q_rest = fine[:, 0], fine[:, 1]
        self.npts, self.K = int(kr.shape[0]), K

    def extracted(self):
        if self.o != self.s:
            return [(self.d, self.cA, self.cq)] if self.cA != 0 else []
        return [(0.0, self.cA, self.cq)] + ([(self.dimg, self.cAi, self.cqi)] if self.dimg else [])

    def AQ(self, r):
        """Linear lookup of the rests at distances r (host NumPy), u = min(r / dr, nfine - 1.000001)."""
        A, Q = self.A_rest.cpu().numpy(), self.Q_rest.cpu().numpy()
        u = np.minimum(np.asarray(r) / self.dr, len(self.rho) - 1.000001)
        i = u.astype(np.int64)
        t = u - i
        return A[i] * (1 - t) + A[i + 1] * t, Q[i] * (1 - t) + Q[i + 1] * t

    def total(self, rho):
        GA, GQ = self.AQ(rho)
        for d, cA, cq in self.extracted():
            R = np.hypot(rho, d)
            inv = np.divide(1.0, R, out=np.zeros(np.shape(R)), where=R > 0)
            GA, GQ = GA + cA * inv, GQ + cq * inv
        return GA, GQ


def green_tables(f, stack, layers, rho_max, dmax):
    """{(a, b): GreenTableML} for every pair a <= b of metal interfaces (sim.py:736-745); the tables of one
    frequency share its contour (uploaded once)."""
    st = stack.at(f)
    contour = device_contour(f, stack, rho_max)
    out = {}
    for i, a in enumerate(layers):
        for b in layers[i:]:
            xs = a != b and abs(stack.z[a] - stack.z[b]) < 3 * dmax
            im = a == b and st.nearest_discontinuity(a) < 1.5 * dmax
            out[(a, b)] = GreenTableML(f, stack, a, b, rho_max, cross_sing=xs, image_sing=im, contour=contour)
    return out


def device_tables(tables, layers):
    """The kernels' table view (DevTables) of {(a, b): GreenTableML} (fast.py:556-568)."""
    keys = sorted(tables)
    T = [tables[k] for k in keys]
    lidx = {k: i for i, k in enumerate(layers)}
    tab = np.zeros((len(layers),) * 2, np.int32)
    for i, (a, b) in enumerate(keys):
        tab[lidx[a], lidx[b]] = tab[lidx[b], lidx[a]] = i
    get = lambda attr, dt: torch.tensor([dt(getattr(t, attr)) for t in T],
                                        dtype=torch.complex128 if dt is complex else torch.float64)
    kind = torch.tensor([0 if a == b else (1 if t.cA != 0 else 2) for (a, b), t in zip(keys, T)], dtype=torch.int32)
    return dv.DevTables(torch.stack([t.A_rest for t in T]), torch.stack([t.Q_rest for t in T]),
                        get('dr', float), get('d', float), get('dimg', float),
                        get('cA', complex), get('cq', complex), get('cAi', complex), get('cqi', complex),
                        torch.from_numpy(tab), kind)
