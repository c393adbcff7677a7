"""The oracle's Green's tables and via kernels against the reference's own values (tests/golden/tables_c1_c2.npz,
written by tests/golden/make_tables.py from the reference's multilayer.py / via.py with its compiled core): the
microstrip stack (C1) and the seeded 4-layer stack (C2) at 1, 3.7 and 8 GHz, every interface pair, to 1e-10 of each
table's largest value (A_rest, Q_rest, D) and of each value (R + homog), the extraction scalars to 1e-12."""
import os

import numpy as np
import pytest

from oracle import greens as og

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'tables_c1_c2.npz')


def gold():
    return np.load(GOLD, allow_pickle=False)


def cases(G):
    for name in ('c1', 'c2'):
        lay = G[f'{name}/layers']
        yield (name, [(float(h), float(er), float(td)) for h, er, td in lay], [int(k) for k in G[f'{name}/interfaces']],
               float(G[f'{name}/dmax']), float(G[f'{name}/rho_max']), [tuple(int(x) for x in s) for s in G[f'{name}/spans']])


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize('name', ['c1', 'c2'])
def test_oracle_tables_match_reference(name):
    G = gold()
    idx = G['idx']
    for nm, lay, ifs, dmax, rho_max, spans in cases(G):
        if nm != name:
            continue
        st = og.LayerStack([og.Layer(h, er, td) for h, er, td in lay])
        for fi, f in enumerate(G['freqs']):
            tabs = og.green_tables(f, st, ifs, rho_max, dmax)
            for (a, b), t in tabs.items():
                key = f'{name}/f{fi}/t{a}{b}'
                assert rel(t._tab.v[idx, 0], G[key + '/A']) < 1e-10, key
                assert rel(t._tab.v[idx, 1], G[key + '/Q']) < 1e-10, key
                sc = G[key + '/scal']
                mine = np.array([t.d, t.dimg, t.cA, t.cq, t.cAi, t.cqi], complex)
                assert np.all(np.abs(mine - sc[:6]) <= 1e-12 * np.maximum(np.abs(sc[:6]), 1e-300)), key
                assert t.npts == int(sc[6].real) and t.K == sc[7].real, key
            vk = og.ViaKernels(f, st, spans, ifs, rho_max)
            for (k, span), tab in vk.D.items():
                key = f'{name}/f{fi}/D{k}_{span[0]}{span[1]}'
                assert rel(tab.v[idx], G[key]) < 1e-10, key
            for sm in sorted(set(spans)):
                for sn in sorted(set(spans)):
                    key = f'{name}/f{fi}/R{sm[0]}{sm[1]}_{sn[0]}{sn[1]}'
                    ref = G[key]
                    mine = vk.R(sm, sn, G['rho_r']) + vk.homog(sm, sn, G['rho_r'])
                    assert np.all(np.abs(mine - ref) <= 1e-10 * np.abs(ref)), key
