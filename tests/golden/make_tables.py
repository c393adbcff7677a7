"""Writes tests/golden/tables_c1_c2.npz: the reference's own Green's tables and via kernels for the stacks of the
microstrip (C1) and the seeded 4-layer board (C2) at 1, 3.7 and 8 GHz. Run in the build container only, where the
reference is mounted at /root/reference (read-only) and its compiled kernel module is in oracle/_ref/:

    sh oracle/build_ref.sh && python tests/golden/make_tables.py

It imports the reference's mom25d.multilayer / mom25d.via / mom25d.fast (NumPy and SciPy only) and installs the
reference's compiled core through the reference's own fast.set_core, as its solver runs. What it records, per
stack and frequency:
  GreenTableML (src/mom25d/multilayer.py:253-318) of every interface pair a <= b with the flags of
  Simulation.green_tables (src/mom25d/sim.py:736-745): A_rest and Q_rest at the indices IDX of the 8192-point
  grid (the first 64 points, then every 32nd), and d, dimg, cA, cq, cAi, cqi, the contour's node count and K;
  ViaKernels (src/mom25d/via.py:179-235) for the board's chain spans: D[(k, span)] at IDX, and R + homog at the
  distances RHO_R for every span pair.
The inputs are written beside the outputs so the tests rebuild the same stacks.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = '/root/reference/src'
OUT = os.path.join(ROOT, 'tests', 'golden', 'tables_c1_c2.npz')
MM = 1e-3
FREQS = (1e9, 3.7e9, 8e9)
IDX = np.r_[0:64, 64:8192:32]
RHO_R = np.array([0.15e-3, 0.5e-3, 2e-3, 8e-3, 20e-3])


def stacks():
    """(name, layers [(h, er, tand)], metal interfaces, dmax, rho_max, chain spans)."""
    rng = np.random.default_rng(2)                 # the seeded board's draws, in its order
    c2 = [(rng.uniform(0.1, 0.4) * MM, rng.uniform(3.5, 4.6), 0.02) for _ in range(4)]
    return [('c1', [(0.8 * MM, 4.4, 0.02)], [1], 0.3 * MM, 33e-3, [(0, 1)]),
            ('c2', c2, [2, 4], 0.25 * MM, 27e-3, [(0, 2), (2, 4)])]


def main():
    sys.path.insert(0, REF)
    sys.path.insert(0, ROOT)
    from mom25d import fast as rfast, multilayer as rml, via as rvia
    from oracle import ref
    core = ref.core()
    if core is None:
        raise SystemExit("oracle/_ref is not built: sh oracle/build_ref.sh")
    rfast.set_core(core)
    out = dict(idx=IDX, freqs=np.array(FREQS), rho_r=RHO_R)
    for name, lay, ifs, dmax, rho_max, spans in stacks():
        out[f'{name}/layers'] = np.array(lay)
        out[f'{name}/interfaces'] = np.array(ifs)
        out[f'{name}/dmax'] = dmax
        out[f'{name}/rho_max'] = rho_max
        out[f'{name}/spans'] = np.array(spans)
        st = rml.LayerStack([rml.Layer(h, er, td) for h, er, td in lay])
        for fi, f in enumerate(FREQS):
            for i, a in enumerate(ifs):
                for b in ifs[i:]:
                    xs = a != b and abs(st.z[a] - st.z[b]) < 3 * dmax
                    im = a == b and st.at(f).nearest_discontinuity(a) < 1.5 * dmax
                    g = rml.GreenTableML(f, st, a, b, rho_max, cross_sing=xs, image_sing=im)
                    key = f'{name}/f{fi}/t{a}{b}'
                    out[key + '/A'] = g.A_rest[IDX]
                    out[key + '/Q'] = g.Q_rest[IDX]
                    out[key + '/scal'] = np.array([g.d, g.dimg, g.cA, g.cq, g.cAi, g.cqi, g.npts, g.K], complex)
            vk = rvia.ViaKernels(f, st, spans, ifs, rho_max)
            for (k, span), tab in vk.D.items():
                out[f'{name}/f{fi}/D{k}_{span[0]}{span[1]}'] = tab.v[IDX]
            for sm in sorted(set(spans)):
                for sn in sorted(set(spans)):
                    out[f'{name}/f{fi}/R{sm[0]}{sm[1]}_{sn[0]}{sn[1]}'] = vk.R(sm, sn, RHO_R) + vk.homog(sm, sn, RHO_R)
            print(f"{name} {f / 1e9:.2f} GHz done", flush=True)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == '__main__':
    main()
