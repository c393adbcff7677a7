"""Via (vertical wire) kernels on the device (src/mom25d/via.py:127-235).

A via's current is constant along each chain (the layers between two interfaces where it touches metal); its end
charges are spread over the fan triangles around the via vertex and enter through G_q like the strips' charges.
Per frequency: the TM network voltages for every source interface (TMNetwork), the smooth via -> interface kernels
D(rho) tabulated per (interface, span) through one batched Hankel transform, the via-via remainder R (directly at
the pair distances, or tabulated above 64 pairs) and the homogeneous K0 part.
"""
import numpy as np
import torch
from scipy.special import kv

from .base import C0, MU0
from . import device as dv
from .multilayer import device_contour, hankel, resample, spline_operator


class TMNetwork:
    """V[s][i](kr): TM voltage at interface i for a unit shunt current at interface s (via.py:127-166)."""

    def __init__(self, stack, kr, w):
        self.st, self.kr, self.w, N = stack, kr, w, stack.N
        Y0, tn, sc, Yd, Yu = stack._network(kr, w, 'e')
        self.k2 = [None] + [(w / C0) ** 2 * stack.eps(j) for j in range(1, N + 1)]
        self.kz2 = [None] + [k2 - kr * kr for k2 in self.k2[1:]]
        self.V = {}
        for s in stack.metal_interfaces:
            v = [None] * (N + 1)
            v[s] = 1.0 / (Yu[s] + Yd[s])
            for k in range(s + 1, N + 1):
                v[k] = torch.zeros_like(kr) if stack.is_pec(k) else v[k - 1] * sc[k] / (1 + 1j * (Yu[k] / Y0[k]) * tn[k])
            for k in range(s, 0, -1):
                v[k - 1] = torch.zeros_like(kr) if stack.is_pec(k - 1) else \
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
        kr2 = self.kr * self.kr
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


class DevTable1D:
    """A radial table on the uniform grid of nfine points (values on the device), linear lookup (via.py:111-123)."""

    def __init__(self, v, rho_max, nfine=8192):
        self.v, self.rho_max, self.nfine = v, rho_max, nfine
        self._dr = np.linspace(0, rho_max, nfine)[1]

    def __call__(self, r):
        v = self.v.cpu().numpy()
        u = np.minimum(np.asarray(r) / self._dr, self.nfine - 1.000001)
        i = u.astype(np.int64)
        t = u - i
        return v[i] * (1 - t) + v[i + 1] * t


class ViaKernels:
    """Per-frequency via kernels for chain spans [(k_bot, k_top), ...] and the metal interfaces (via.py:179-235)."""

    def __init__(self, f, stack, spans, interfaces, rho_max, nrho=200):
        w = 2 * np.pi * f
        kr, dk, _ = device_contour(f, stack, rho_max)            # the contour of the frequency's tables
        stack = stack.at(f)
        net = TMNetwork(stack, kr, w)
        self.f, self.st, self.w, self.kr, self.dk, self.net = f, stack, w, kr, dk, net
        self._gq = {}
        self.rho_max, self.nrho, self.D = rho_max, nrho, {}
        self.rho = dv.on_dev(rho_max * np.linspace(0, 1, nrho) ** 2, dv.f64)
        keys, specs = [], []
        for span in sorted(set(spans)):
            for k in interfaces:
                spec = net.S(k, _layers(span))
                for ke, div in chain_ends(*span):
                    spec = spec - div * self.gq(k, ke) / (1j * w)
                keys.append((k, span))
                specs.append(spec * torch.ones_like(kr))
        if keys:
            vals = hankel(torch.stack(specs), kr, dk, self.rho)              # one transform for every table
            fine = resample(spline_operator(rho_max, nrho, 8192), vals.T)     # (8192, ntables)
            for j, key in enumerate(keys):
                self.D[key] = DevTable1D(fine[:, j].contiguous(), rho_max)

    def gq(self, a, b):
        """G~_q between interfaces a and b (0 at a PEC one), cached per pair."""
        if self.net.is_pec(a) or self.net.is_pec(b):
            return 0.0
        key = (min(a, b), max(a, b))
        if key not in self._gq:
            self._gq[key] = self.st.spectral(self.kr, self.w, a, b)[1]
        return self._gq[key]

    def _R_spec(self, sm, sn):
        spec = self.net.P(_layers(sm), _layers(sn), with_homog=False)
        for ka, da in chain_ends(*sm):
            for kb, db in chain_ends(*sn):
                spec = spec - da * db * self.gq(ka, kb) / (1j * self.w)
        return spec * torch.ones_like(self.kr)

    def R(self, sm, sn, rho):
        """The via-via reaction remainder at distances rho (host array in, host array out)."""
        rho = np.atleast_1d(np.asarray(rho, float))
        return hankel(self._R_spec(sm, sn)[None], self.kr, self.dk, dv.on_dev(rho, dv.f64))[0].cpu().numpy()

    def R_table(self, sm, sn):
        vals = hankel(self._R_spec(sm, sn)[None], self.kr, self.dk, self.rho)      # (1, nrho)
        return DevTable1D(resample(spline_operator(self.rho_max, self.nrho, 8192), vals.T)[:, 0].contiguous(),
                          self.rho_max)

    def homog(self, sm, sn, rho):
        """The analytic K0 part of the reaction between spans (host; a handful of values per board)."""
        rho, out = np.atleast_1d(rho), np.zeros(np.size(rho), complex)
        for j in set(_layers(sm)) & set(_layers(sn)):
            kj = np.sqrt(self.net.k2[j] + 0j)
            if kj.imag > 0:
                kj = -kj
            out += 1j * self.w * MU0 * self.st.layers[j - 1].h / (2 * np.pi) * kv(0, 1j * kj * rho)
        return out
