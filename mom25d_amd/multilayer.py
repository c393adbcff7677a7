"""The layered stack and the per-frequency Green's tables, built on the GPU.

  Layer, LayerStack        the stack (src/mom25d/multilayer.py:58-250): layers bottom -> top, a PEC ground or a
                           half-space below, an open or PEC top; its parameters at one frequency as the C ABI's
                           m25x_stack
  sommerfeld_contour       the integration path of the Sommerfeld integrals (src/mom25d/via.py:36-53)
  green_tables             every interface pair's GreenTableML of a frequency (src/mom25d/multilayer.py:253-318,
                           src/mom25d/sim.py:736-745) in one batch: one launch of m25x_spectra for all spectra, one
                           Hankel transform (m25x_hankel) and one spline product for all tables
  RunLog                   stage timings and notes (src/mom25d/multilayer.py:30-53)

The spectral network itself (layer admittances, the voltages of a shunt source, the spectra) is the HIP kernel
m25x_spectra; the host only evaluates the same functions at single points through m25x_spectral_point (the
extraction coefficients). Internal PEC planes (LayerStack(pec=...)) are not supported: none of this build's boards
has one.
"""
import ctypes as C
import time
from contextlib import contextmanager

import numpy as np
import torch
from scipy.interpolate import CubicSpline

from .base import C0, EPS0, MU0
from . import _lib as L
from . import device as dv

DS_M1, DS_M2 = 3.0, 12.0          # decades of the wideband (Djordjevic-Sarkar) model's relaxation band
NRHO, NFINE = 260, 8192           # the tables' quadratic grid and the uniform grid they are resampled to


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


def djordjevic_sarkar(er, tand, fref, f):
    """The wideband Debye-sum permittivity at f that equals er (1 - j tand) at fref (multilayer.py:65-71): the
    relaxation spread evenly over 10^DS_M1 .. 10^DS_M2 rad/s, so er' falls and tand stays nearly flat with f."""
    def shape(x):
        jw = 2j * np.pi * np.asarray(x, float)
        return np.log10((2 * np.pi * 10 ** DS_M2 + jw) / (2 * np.pi * 10 ** DS_M1 + jw)) / (DS_M2 - DS_M1)
    at_ref = shape(fref)
    strength = -er * tand / at_ref.imag
    return er - strength * at_ref.real + strength * shape(f)


class Layer:
    """A dielectric layer: h (m), er, tand; fref (Hz) for a wideband (Djordjevic-Sarkar) layer."""

    def __init__(self, h, er, tand=0.0, name=None, fref=None):
        self.h, self.er, self.tand, self.name, self.fref = h, er, tand, name, fref

    @property
    def er_c(self):
        return self.er * (1 - 1j * self.tand)

    def er_at(self, f):
        return self.er_c if self.fref is None else complex(djordjevic_sarkar(self.er, self.tand, self.fref, f))


class LayerStack:
    """Dielectric layers bottom -> top; interface k is the top of layer k (0 the bottom of layer 1). bottom 'pec'
    (a ground at z = 0) or 'open' (a half-space of bottom_er); top 'open' (a half-space of top_er) or 'pec'."""

    def __init__(self, layers, top='open', top_er=1.0, bottom='pec', bottom_er=1.0, pec=()):
        if top not in ('open', 'pec') or bottom not in ('open', 'pec'):
            raise ValueError("top / bottom: 'open' or 'pec'")
        if tuple(pec):
            raise ValueError("PEC planes inside the stack are not supported by this build")
        if not layers:
            raise ValueError("a stack needs one layer or more")
        if len(layers) > L.MAX_LAYERS:
            raise ValueError(f"at most {L.MAX_LAYERS} layers")
        self.layers, self.top, self.top_er, self.bottom, self.bottom_er = list(layers), top, top_er, bottom, bottom_er
        self.pec = ()
        for i, l in enumerate(self.layers):
            l.name = l.name if l.name is not None else f"D{i + 1}"
        self.z = np.concatenate([[0.0], np.cumsum([l.h for l in self.layers])])

    @property
    def N(self):
        return len(self.layers)

    @property
    def dispersive(self):
        return any(l.fref is not None for l in self.layers)

    def at(self, f):
        """The stack with every wideband layer's permittivity frozen at f (itself when none is wideband)."""
        if not self.dispersive:
            return self
        frozen = []
        for l in self.layers:
            e = l.er_at(f)
            frozen.append(Layer(l.h, e.real, -e.imag / e.real, l.name))
        out = LayerStack(frozen, self.top, self.top_er, self.bottom, self.bottom_er)
        out.f = f
        return out

    # ---- media -----------------------------------------------------------------------------------------------
    def _media(self):
        """Permittivity of every medium from below the bottom (index 0) to above the top (index N + 1)."""
        return [complex(self.bottom_er)] + [l.er_c for l in self.layers] + [complex(self.top_er)]

    def eps(self, j):
        """Layer j's permittivity (1..N); j <= 0 the half-space below, j > N the one above."""
        return self._media()[min(max(j, 0), self.N + 1)]

    def eps_sum(self, k):
        """The permittivities on both sides of interface k (its 1/rho coefficient's denominator)."""
        return self.eps(k) + self.eps(k + 1)

    def is_pec(self, k):
        return (k == 0 and self.bottom == 'pec') or (k == self.N and self.top == 'pec')

    @property
    def metal_interfaces(self):
        first = 1 if self.bottom == 'pec' else 0
        last = self.N - 1 if self.top == 'pec' else self.N
        return list(range(first, last + 1))

    @property
    def hmin(self):
        return min(l.h for l in self.layers)

    def max_er(self):
        """The largest |er| of the media the fields live in (the layers and the open half-spaces)."""
        m = self._media()
        inner = [abs(e) for e in m[1:-1]]
        if self.top == 'open':
            inner.append(abs(m[-1]))
        if self.bottom == 'open':
            inner.append(abs(m[0]))
        return max(inner)

    def nearest_discontinuity(self, k):
        """The distance from interface k to the nearest interface that is a PEC plane or a permittivity step."""
        m = self._media()
        steps = [j for j in range(self.N + 1) if j != k and (self.is_pec(j) or m[j] != m[j + 1])]
        return min((abs(self.z[k] - self.z[j]) for j in steps), default=np.inf)

    def key(self):
        """The parameters the spectral functions depend on (at this stack's frequency)."""
        return (self.top, self.bottom, complex(self.top_er), complex(self.bottom_er),
                tuple((complex(l.er_c), float(l.h)) for l in self.layers))

    # ---- the C ABI's view --------------------------------------------------------------------------------------
    def struct(self, w):
        """m25x_stack at angular frequency w (this stack's permittivities: call on stack.at(f))."""
        s = L.Stack()
        s.N, s.bottom_pec, s.top_pec = self.N, int(self.bottom == 'pec'), int(self.top == 'pec')
        for j, l in enumerate(self.layers):
            e = l.er_c
            s.eps[j][0], s.eps[j][1] = e.real, e.imag
            s.h[j] = l.h
        eb, et = complex(self.bottom_er), complex(self.top_er)
        s.eps_bot[0], s.eps_bot[1], s.eps_top[0], s.eps_top[1] = eb.real, eb.imag, et.real, et.imag
        s.k0, s.wmu, s.weps = w / C0, w * MU0, w * EPS0
        return s

    def spectral_point(self, kr, w, o, s):
        """(G~_A, G~_q) between interfaces o and s at one kr, on the host."""
        out = (C.c_double * 4)()
        kr = complex(kr)
        L.call('m25x_spectral_point', C.byref(self.struct(w)), float(w), kr.real, kr.imag, int(o), int(s), out)
        return complex(out[0], out[1]), complex(out[2], out[3])


def sommerfeld_contour(f, stack, rho_max, K=None, nq=8):
    """(kr, dkr, K): nq-point Gauss-Legendre panels along real t from 0 to K, lifted into Im kr > 0 by
    lift sin(pi t / t1) up to t1 = 1.5 k_max (clear of the branch points and surface-wave poles), real beyond; the
    panels are at most min(0.3 k0, 0.8 lift) wide before t1 and 2 pi / rho_max after it (via.py:36-53)."""
    k0 = 2 * np.pi * f / C0
    kmax = k0 * np.sqrt(stack.max_er())
    if not K:
        K = max(1e5, 30.0 / stack.hmin) + 40 * kmax
    t1, lift = 1.5 * kmax, min(0.5 * k0, 4.0 / rho_max)
    near = np.linspace(0, t1, int(np.ceil(t1 / min(0.3 * k0, 0.8 * lift))) + 1)
    far = np.arange(t1, K, 2 * np.pi / rho_max)[1:]
    edges = np.concatenate([near, far, [K]])
    x, wq = np.polynomial.legendre.leggauss(nq)
    lo, hi = edges[:-1, None], edges[1:, None]
    half, mid = 0.5 * (hi - lo), 0.5 * (lo + hi)
    t, wt = (half * x[None, :] + mid).ravel(), (half * wq[None, :]).ravel()
    arc = np.pi * t / t1
    on = t < t1
    lifted = np.where(on, lift * np.sin(arc), 0.0)
    slope = np.where(on, lift * np.pi / t1 * np.cos(arc), 0.0)
    return t + 1j * lifted, wt * (1 + 1j * slope), K


_CONTOURS = {}


def device_contour(f, stack, rho_max):
    """(kr, dkr, K) of sommerfeld_contour on the device, kept for the last few (f, stack, rho_max): the tables and
    the via kernels of a frequency share one contour."""
    st = stack.at(f)
    key = (float(f), float(rho_max), st.key(), torch.cuda.current_device())
    c = _CONTOURS.get(key)
    if c is None:
        kr_h, dk_h, K = sommerfeld_contour(f, st, rho_max)
        both = dv.on_dev(np.stack([kr_h, dk_h]), dv.c128)
        c = (both[0], both[1], K)
        if len(_CONTOURS) >= 4:
            _CONTOURS.clear()
        _CONTOURS[key] = c
    return c


_SPLINES = {}


def spline_operator(rho_max, nrho, nfine):
    """The (nfine x nrho) matrix S with S @ y = CubicSpline(rho_max u^2, y)(uniform nfine grid) (not-a-knot, the
    reference's Table1D, via.py:111-123): the spline is linear in its data, so its action is fixed by the grid, and
    a cubic spline is unchanged by a scaling of its abscissae, so one S (built on rho_max = 1) serves every rho_max.
    Built once per (nrho, nfine) on the host, kept on the device."""
    key = (int(nrho), int(nfine), torch.cuda.current_device())
    S = _SPLINES.get(key)
    if S is None:
        rho = np.linspace(0, 1, nrho) ** 2
        fine = np.linspace(0, 1, nfine)
        S = torch.from_numpy(np.ascontiguousarray(CubicSpline(rho, np.eye(nrho))(fine))).to(dv.device())
        _SPLINES[key] = S
    return S


def resample(S, vals):
    """S (nfine x n real) applied to complex vals (n, m) on the device."""
    return torch.complex(S @ vals.real.contiguous(), S @ vals.imag.contiguous())


def quad_grid(rho_max, nrho):
    return rho_max * np.linspace(0, 1, nrho) ** 2


class Requests:
    """A batch of m25x_spectra requests (M25X_SPEC_NI ints and M25X_SPEC_ND doubles each) and their output rows."""

    def __init__(self):
        self.ints, self.dbls, self.rows = [], [], 0

    def add(self, kind, args, dbl=()):
        row, n_out = self.rows, 2 if kind == L.SPEC_TABLE else 1
        self.ints.append([kind, row] + list(args) + [0] * (L.SPEC_NI - 2 - len(args)))
        self.dbls.append(list(dbl) + [0.0] * (L.SPEC_ND - len(dbl)))
        self.rows += n_out
        return row

    def run(self, stack, w, contour):
        """W (rows x npts complex128, device): every request's spectra times kr dkr on the contour."""
        kr, dk, _ = contour
        npts = int(kr.shape[0])
        W = torch.empty((self.rows, npts), dtype=dv.c128, device=kr.device)
        if not self.ints:
            return W
        ri_h = np.ascontiguousarray(self.ints, dtype=np.int32)
        ri = torch.from_numpy(ri_h).to(kr.device)
        rd = torch.from_numpy(np.ascontiguousarray(self.dbls, dtype=np.float64)).to(kr.device)
        L.call('m25x_spectra', C.byref(stack.struct(w)), float(w), dv.ptr(kr), dv.ptr(dk), npts, dv.ptr(ri),
               dv.ptr(rd), len(self.ints), ri_h.ctypes.data_as(C.c_void_p), dv.ptr(W), self.rows, dv.stream())
        return W


def hankel(W, contour, rho):
    """(1/2 pi) sum_k W[:, k] J0(kr_k rho) for Hankel weights W (rows, npts) on the device (via.py:104-108)."""
    kr = contour[0]
    return dv.hankel_sum(rho, kr, W) / (2 * np.pi)


class GreenTableML:
    """The smooth remainders A_rest(rho), Q_rest(rho) of G_A and G_q between interfaces o and s (device, on the
    uniform grid rho of nfine points) and the parts extracted from them, c / sqrt(rho^2 + d^2) (multilayer.py:
    253-318): the 1/rho singularity of a same-interface pair (cA = mu0 / 4 pi, cq = 1 / (2 pi eps0 eps_sum)) and,
    where image_sing, its nearest image at dimg = 2 x the nearest discontinuity; for a cross pair with
    cross_sing, the near-singular part at its separation d."""

    def __init__(self, f, stack, o, s, rho_max, nrho=NRHO, nfine=NFINE, K=None, nq=8, cross_sing=False,
                 image_sing=None, contour=None):
        self.f, self.o, self.s = f, o, s
        st = stack.at(f)
        self._extraction(st, 2 * np.pi * f, cross_sing, image_sing)
        if contour is None:
            kr_h, dk_h, K = sommerfeld_contour(f, st, rho_max, K, nq)
            both = dv.on_dev(np.stack([kr_h, dk_h]), dv.c128)
            contour = (both[0], both[1], K)
        _fill_tables([self], st, 2 * np.pi * f, contour, rho_max, nrho, nfine)

    @classmethod
    def _blank(cls, f, o, s):
        t = cls.__new__(cls)
        t.f, t.o, t.s = f, o, s
        return t

    def _extraction(self, st, w, cross_sing, image_sing):
        o, s = self.o, self.s
        self.d = abs(st.z[o] - st.z[s])
        self.dimg, self.cAi, self.cqi = 0.0, 0.0, 0.0
        if o == s:
            self.cA, self.cq = MU0 / (4 * np.pi), 1.0 / (2 * np.pi * EPS0 * st.eps_sum(o))
            if image_sing:
                # the nearest image's coefficients from the static network (w = 2 pi: 1 Hz), at a kr where the
                # image term is e^-18 and the farther ones are gone
                self.dimg = 2 * st.nearest_discontinuity(o)
                k1 = 18.0 / self.dimg
                ga, gq = st.spectral_point(k1, 2 * np.pi, o, s)
                scale = k1 * np.exp(k1 * self.dimg) / (2 * np.pi)
                self.cAi = ((ga - MU0 / (2 * k1)) * scale).real
                self.cqi = (gq - 1.0 / (EPS0 * st.eps_sum(o) * k1)) * scale
        elif cross_sing:
            # close interfaces: the large-kr limit of the network at this frequency gives the 1/R coefficients
            k1 = min(25.0 / st.hmin, 300.0 / self.d)
            ga, gq = st.spectral_point(k1, w, o, s)
            scale = k1 * np.exp(k1 * self.d) / (2 * np.pi)
            self.cA, self.cq = ga * scale, gq * scale
        else:
            self.cA, self.cq = 0.0, 0.0

    def extracted(self):
        """[(d, cA, cq)]: the parts taken out of the tables (added back by the fill kernels)."""
        if self.o != self.s:
            return [(self.d, self.cA, self.cq)] if self.cA != 0 else []
        parts = [(0.0, self.cA, self.cq)]
        if self.dimg:
            parts.append((self.dimg, self.cAi, self.cqi))
        return parts

    def AQ(self, r):
        """Linear lookup of the rests at distances r (host NumPy), u = min(r / dr, nfine - 1.000001)."""
        A, Q = self.A_rest.cpu().numpy(), self.Q_rest.cpu().numpy()
        u = np.minimum(np.asarray(r) / self.dr, len(self.rho) - 1.000001)
        i = u.astype(np.int64)
        t = u - i
        return A[i] * (1 - t) + A[i + 1] * t, Q[i] * (1 - t) + Q[i + 1] * t

    def total(self, rho):
        """G_A, G_q at distances rho with the extracted parts added back (0 for a part at rho = d = 0)."""
        GA, GQ = self.AQ(rho)
        for d, cA, cq in self.extracted():
            R = np.hypot(rho, d)
            inv = np.divide(1.0, R, out=np.zeros(np.shape(R)), where=R > 0)
            GA, GQ = GA + cA * inv, GQ + cq * inv
        return GA, GQ


def _fill_tables(tabs, st, w, contour, rho_max, nrho, nfine):
    """A_rest and Q_rest of every table in tabs (their extraction set): one spectra launch, one Hankel transform on
    the quadratic grid, one spline product to the uniform grid."""
    req = Requests()
    for t in tabs:
        parts = t.extracted()
        dbl = []
        for d, cA, cq in parts:
            cA, cq = complex(cA), complex(cq)
            dbl += [d, cA.real, cA.imag, cq.real, cq.imag]
        t._row = req.add(L.SPEC_TABLE, (t.o, t.s, len(parts)), dbl)
    W = req.run(st, w, contour)
    rho = dv.on_dev(quad_grid(rho_max, nrho), dv.f64)
    coarse = hankel(W, contour, rho)                                       # (2 ntab, nrho)
    fine = resample(spline_operator(rho_max, nrho, nfine), coarse.T)       # (nfine, 2 ntab)
    grid = np.linspace(0, rho_max, nfine)
    for t in tabs:
        t.A_rest, t.Q_rest = fine[:, t._row], fine[:, t._row + 1]
        t.rho, t.dr = grid, grid[1]
        t.npts, t.K = int(contour[0].shape[0]), contour[2]
        t._fine = (fine, t._row)


def green_tables(f, stack, layers, rho_max, dmax):
    """{(a, b): GreenTableML} for every pair a <= b of metal interfaces (sim.py:736-745): a cross pair closer than
    3 dmax has its near-singular part extracted, a same-interface pair within 1.5 dmax of a discontinuity its image.
    The tables of a frequency share one contour and are computed in one batch."""
    st, w = stack.at(f), 2 * np.pi * f
    contour = device_contour(f, stack, rho_max)
    out = {}
    for i, a in enumerate(layers):
        for b in layers[i:]:
            t = GreenTableML._blank(f, a, b)
            t._extraction(st, w, cross_sing=a != b and abs(st.z[a] - st.z[b]) < 3 * dmax,
                          image_sing=a == b and st.nearest_discontinuity(a) < 1.5 * dmax)
            out[(a, b)] = t
    if out:
        _fill_tables(list(out.values()), st, w, contour, rho_max, NRHO, NFINE)
    return out


def device_tables(tables, layers):
    """The kernels' table view (DevTables) of {(a, b): GreenTableML} (fast.py:556-568), packed from the batch's
    spline output in one pass; the scalars come from the host."""
    keys = sorted(tables)
    T = [tables[k] for k in keys]
    lidx = {k: i for i, k in enumerate(layers)}
    tab = np.zeros((len(layers),) * 2, np.int32)
    for i, (a, b) in enumerate(keys):
        tab[lidx[a], lidx[b]] = tab[lidx[b], lidx[a]] = i
    kind = np.array([0 if a == b else (1 if t.cA != 0 else 2) for (a, b), t in zip(keys, T)], np.int32)
    fine, _ = T[0]._fine
    if all(t._fine[0] is fine for t in T):
        cols = torch.tensor(sum(([t._row, t._row + 1] for t in T), []), device=fine.device)
        AQ = torch.view_as_real(fine[:, cols]).reshape(fine.shape[0], len(T), 4).transpose(0, 1)
    else:
        AQ = torch.stack([torch.stack([t.A_rest.real, t.A_rest.imag, t.Q_rest.real, t.Q_rest.imag], -1) for t in T])
    scal = dict(dr=np.array([t.dr for t in T]), d=np.array([t.d for t in T], float),
                dimg=np.array([t.dimg for t in T], float),
                cA=np.array([complex(t.cA) for t in T]), cq=np.array([complex(t.cq) for t in T]),
                cAi=np.array([complex(t.cAi) for t in T]), cqi=np.array([complex(t.cqi) for t in T]))
    return dv.DevTables.packed(AQ, scal, tab, kind)
