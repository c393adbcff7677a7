"""Via (vertical wire) kernels on the device (src/mom25d/via.py:127-235).

A via's current is constant along each chain (the layers between two interfaces where it touches metal); its end
charges spread over the fan triangles around the via vertex and enter through G_q like the strips' charges. Per
frequency, one m25x_spectra launch gives every spectrum the vias need on the frequency's contour:
  D(k, span)   the smooth via -> interface kernel of a chain span seen on interface k, tabulated on the uniform
               grid (one Hankel transform and one spline product for all of them)
  R(sm, sn)    the via -> via reaction remainder between two spans, transformed at the pairs' distances (or
               tabulated, for many pairs)
and homog() the analytic K0 part of the reaction, on the host.
"""
import numpy as np
import torch
from scipy.special import kv

from .base import C0, MU0
from . import _lib as L
from . import device as dv
from .multilayer import NFINE, Requests, device_contour, hankel, quad_grid, resample, spline_operator


class DevTable1D:
    """A radial table on the uniform grid of nfine points (values on the device), linear lookup (via.py:111-123)."""

    def __init__(self, v, rho_max, nfine=NFINE):
        self.v, self.rho_max, self.nfine = v, rho_max, nfine
        self._dr = np.linspace(0, rho_max, nfine)[1]

    def __call__(self, r):
        v = self.v.cpu().numpy()
        u = np.minimum(np.asarray(r) / self._dr, self.nfine - 1.000001)
        i = u.astype(np.int64)
        t = u - i
        return v[i] * (1 - t) + v[i + 1] * t


def span_layers(span):
    """The dielectric layers a chain from interface span[0] up to span[1] runs through."""
    return range(span[0] + 1, span[1] + 1)


class ViaKernels:
    """Per-frequency via kernels for chain spans [(k_bot, k_top), ...] and the metal interfaces (via.py:179-235)."""

    def __init__(self, f, stack, spans, interfaces, rho_max, nrho=200):
        self.f, self.w = f, 2 * np.pi * f
        self.st = stack.at(f)
        self.contour = device_contour(f, stack, rho_max)
        self.rho_max, self.nrho, self.D = rho_max, nrho, {}
        self.spans = sorted(set(spans))
        req, d_rows, self._r_rows = Requests(), [], {}
        for span in self.spans:
            for k in interfaces:
                d_rows.append(((k, span), req.add(L.SPEC_VIA_D, (k, span[0], span[1]))))
        for sm in self.spans:
            for sn in self.spans:
                self._r_rows[(sm, sn)] = req.add(L.SPEC_VIA_R, (sm[0], sm[1], sn[0], sn[1]))
        self._W = req.run(self.st, self.w, self.contour)
        if d_rows:
            rows = torch.tensor([r for _, r in d_rows], device=self._W.device)
            grid = dv.on_dev(quad_grid(rho_max, nrho), dv.f64)
            fine = resample(spline_operator(rho_max, nrho, NFINE), hankel(self._W[rows], self.contour, grid).T)
            for j, (key, _) in enumerate(d_rows):
                self.D[key] = DevTable1D(fine[:, j].contiguous(), rho_max)

    def _r_row(self, sm, sn):
        row = self._r_rows.get((tuple(sm), tuple(sn)))
        if row is None:
            raise KeyError(f"no via reaction between spans {sm} and {sn} at this frequency")
        return self._W[row:row + 1]

    def R(self, sm, sn, rho):
        """The via-via reaction remainder at distances rho (host array in, host array out)."""
        rho = np.atleast_1d(np.asarray(rho, float))
        return hankel(self._r_row(sm, sn), self.contour, dv.on_dev(rho, dv.f64))[0].cpu().numpy()

    def R_table(self, sm, sn):
        """R tabulated on the uniform grid (one transform on the quadratic grid): for many pairs of one span pair."""
        grid = dv.on_dev(quad_grid(self.rho_max, self.nrho), dv.f64)
        vals = hankel(self._r_row(sm, sn), self.contour, grid)
        return DevTable1D(resample(spline_operator(self.rho_max, self.nrho, NFINE), vals.T)[:, 0].contiguous(),
                          self.rho_max)

    def homog(self, sm, sn, rho):
        """The analytic part of the reaction of two spans through their shared layers j: a line current in the
        homogeneous medium of layer j, (j w mu0 h_j / 2 pi) K0(j k_j rho), k_j with Im k_j <= 0 (host)."""
        rho = np.atleast_1d(np.asarray(rho, float))
        out = np.zeros(rho.shape, complex)
        k0sq = (self.w / C0) ** 2
        for j in sorted(set(span_layers(sm)) & set(span_layers(sn))):
            kj = np.sqrt(k0sq * self.st.eps(j) + 0j)
            kj = -kj if kj.imag > 0 else kj
            out += 1j * self.w * MU0 * self.st.layers[j - 1].h / (2 * np.pi) * kv(0, 1j * kj * rho)
        return out
