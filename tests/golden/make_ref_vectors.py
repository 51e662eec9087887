"""Write per-step golden vectors from the reference's own compiled Fortran step.

Runs in the development container only: it needs oracle/_ref/*.so, which
``make -f oracle/ref.mk`` builds from the reference sources under
/root/reference.  Each case stores seeded inputs (qbc with ghosts filled by the
reference BC rules) and the reference routine's outputs:

  classic/step2.f90      unsplit, transverse 0/1/2, limiters minmod/vanleer/MC
  classic/step2ds.f90    split sweeps 1 and 2
  classic/step3ds.f90    split sweeps 1, 2, 3
  sharpclaw/flux2.f90    char_decomp 0 (weno.f90 weno5) and 2 (weno5_char)

and, with aux fields, step2ds sweeps 1 and 2 driving the advection_reaction_2d
example's own Riemann solver (rpn2_vc_advection_vector.f90, built by ref.mk).

Output: tests/golden/ref_step_vectors.npz, tests/golden/ref_step_vectors_aux.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402
from tests._cases import seeded_state  # noqa: E402


def main():
    if not O.ref_available():
        sys.exit("reference build missing: run make -f oracle/ref.mk")
    rng = np.random.default_rng(20240611)
    out = {}
    n = 0
    for kind in (O.ACOUSTICS, O.EULER, O.SHALLOW, O.ADVECTION):
        rp = O.make_rp(kind)
        mw = O.MWAVES[kind]
        bcl, bcu = ([O.EXTRAP, O.WALL], [O.WALL, O.PERIODIC]) if kind != O.ADVECTION \
            else ([O.PERIODIC, O.EXTRAP], [O.PERIODIC, O.EXTRAP])
        for ids, trans, lim in [(0, 0, 4), (0, 1, 3), (0, 2, 4), (0, 2, 1), (1, 0, 4), (2, 0, 4)]:
            ng, mx, my = 2, 14, 11
            qbc = O.with_ghosts(seeded_state(rng, kind, (mx, my)), ng)
            O.bc2(qbc, ng, bcl, bcu)
            meth = O.method_array(transverse=trans, dimensional_split=ids > 0)
            dx, dy, dt = 0.1, 0.13, 0.021
            qnew = qbc.copy(order="F")
            cfl = O.ref_step2(rp, qbc, qnew, ng, dx, dy, dt, meth, [lim] * mw, ids=ids)
            key = f"c{n:02d}"
            out[key + "_meta"] = np.array([kind, ids, trans, lim, ng, mx, my, 0], dtype=np.int64)
            out[key + "_real"] = np.array([dx, dy, dt, cfl, 0.0])
            out[key + "_bc"] = np.array(bcl + bcu, dtype=np.int64)
            out[key + "_qin"] = qbc
            out[key + "_qout"] = qnew
            n += 1
    rp = O.make_rp(O.EULER3)
    for idir in (1, 2, 3):
        ng, mx, my, mz = 2, 7, 6, 5
        qbc = O.with_ghosts(seeded_state(rng, O.EULER3, (mx, my, mz)), ng)
        O.bc3(qbc, ng, [O.EXTRAP, O.EXTRAP, O.WALL], [O.EXTRAP, O.WALL, O.EXTRAP])
        meth = O.method_array(dimensional_split=True)
        qnew = qbc.copy(order="F")
        cfl = O.ref_step3ds(rp, qbc, qnew, ng, 0.1, 0.12, 0.13, 0.02, meth, [4] * 3, idir)
        key = f"c{n:02d}"
        out[key + "_meta"] = np.array([O.EULER3, 10 + idir, -1, 4, ng, mx, my, mz], dtype=np.int64)
        out[key + "_real"] = np.array([0.1, 0.12, 0.02, cfl, 0.13])
        out[key + "_bc"] = np.array([O.EXTRAP, O.EXTRAP, O.WALL, O.EXTRAP, O.WALL, O.EXTRAP],
                                    dtype=np.int64)
        out[key + "_qin"] = qbc
        out[key + "_qout"] = qnew
        n += 1
    for kind, cd in [(O.SHALLOW, 0), (O.SHALLOW, 2), (O.EULER, 0), (O.EULER, 2), (O.ACOUSTICS, 0)]:
        rp = O.make_rp(kind)
        ng, mx, my = 3, 12, 12
        qbc = O.with_ghosts(seeded_state(rng, kind, (mx, my)), ng)
        O.bc2(qbc, ng, [O.PERIODIC, O.PERIODIC], [O.PERIODIC, O.PERIODIC])
        dq, cfl = O.ref_sharp_flux2(rp, qbc, ng, 0.1, 0.1, 0.012, cd)
        key = f"c{n:02d}"
        out[key + "_meta"] = np.array([kind, 20 + cd, -1, 0, ng, mx, my, 0], dtype=np.int64)
        out[key + "_real"] = np.array([0.1, 0.1, 0.012, cfl, 0.0])
        out[key + "_bc"] = np.array([2, 2, 2, 2], dtype=np.int64)
        out[key + "_qin"] = qbc
        out[key + "_qout"] = dq
        n += 1
    dst = os.path.join(HERE, "ref_step_vectors.npz")
    np.savez_compressed(dst, **out)
    print(dst, n, "cases")
    aux_vectors()


def aux_vectors():
    """step2ds with edge-velocity aux fields: the example's rpn2, meqn 1..3,
    velocities of both signs, aux BCs (never negated) as the q BCs; then the
    unsplit step2 with aux and with a capacity field."""
    rng = np.random.default_rng(20240612)
    rpn2 = O.ref_vcadv_rpn2()
    out = {}
    n = 0
    for meqn, ids, lim, bcl, bcu in [(2, 1, 4, [O.EXTRAP, O.WALL], [O.PERIODIC, O.EXTRAP]),
                                     (2, 2, 4, [O.EXTRAP, O.WALL], [O.PERIODIC, O.EXTRAP]),
                                     (1, 1, 1, [O.WALL, O.PERIODIC], [O.WALL, O.PERIODIC]),
                                     (3, 2, 3, [O.PERIODIC, O.EXTRAP], [O.PERIODIC, O.WALL])]:
        ng, mx, my = 2, 13, 10
        q = rng.uniform(-1.0, 2.0, (meqn, mx, my))
        aux = rng.uniform(-1.5, 1.5, (2, mx, my))
        qbc = O.with_ghosts(np.asfortranarray(q), ng)
        abc = O.with_ghosts(np.asfortranarray(aux), ng)
        O.bc2(qbc, ng, bcl, bcu)
        O.bc2(abc, ng, bcl, bcu, reflect=(-1, -1))
        meth = O.method_array(dimensional_split=True, num_aux=2)
        dx, dy, dt = 0.1, 0.13, 0.021
        qnew = qbc.copy(order="F")
        cfl = O.ref_step2ds_aux(rpn2, qbc, qnew, abc, ng, dx, dy, dt, meth, [lim], ids)
        key = f"a{n:02d}"
        out[key + "_meta"] = np.array([O.VC_ADVECTION, ids, -1, lim, ng, mx, my, meqn],
                                      dtype=np.int64)
        out[key + "_real"] = np.array([dx, dy, dt, cfl, 0.0])
        out[key + "_bc"] = np.array(bcl + bcu, dtype=np.int64)
        out[key + "_qin"] = qbc
        out[key + "_aux"] = abc
        out[key + "_qout"] = qnew
        n += 1
    # the unsplit step2 with aux (vc_advection_2D's rpt2 from the oracle) and a
    # capacity field: order, transverse, limiter, index_capa, num_aux
    for kind, meqn, order, trans, lim, capa, maux, bcl, bcu in [
            (7, 1, 2, 2, 3, 2, 3, [O.EXTRAP, O.PERIODIC], [O.EXTRAP, O.PERIODIC]),
            (7, 1, 2, 1, 4, -1, 2, [O.WALL, O.EXTRAP], [O.PERIODIC, O.WALL]),
            (7, 1, 1, 2, 1, 2, 3, [O.PERIODIC, O.WALL], [O.PERIODIC, O.EXTRAP]),
            (6, 2, 2, 0, 4, 2, 3, [O.EXTRAP, O.EXTRAP], [O.WALL, O.PERIODIC])]:
        ng, mx, my = 2, 11, 14
        q = rng.uniform(-1.0, 2.0, (meqn, mx, my))
        aux = rng.uniform(-1.5, 1.5, (maux, mx, my))
        if capa >= 0:
            aux[capa] = rng.uniform(0.5, 2.0, (mx, my))
        qbc = O.with_ghosts(np.asfortranarray(q), ng)
        abc = O.with_ghosts(np.asfortranarray(aux), ng)
        O.bc2(qbc, ng, bcl, bcu)
        O.bc2(abc, ng, bcl, bcu, reflect=(-1, -1))
        meth = O.method_array(order=order, transverse=trans, num_aux=maux, index_capa=capa)
        dx, dy, dt = 0.1, 0.13, 0.021
        qnew = qbc.copy(order="F")
        cfl = O.ref_step2_aux(O.make_rp(kind, meqn=meqn), rpn2, qbc, qnew, abc, ng, dx, dy, dt,
                              meth, [lim])
        key = f"a{n:02d}"
        out[key + "_meta"] = np.array([kind, 0, trans, lim, ng, mx, my, meqn], dtype=np.int64)
        out[key + "_ext"] = np.array([order, capa, maux], dtype=np.int64)
        out[key + "_real"] = np.array([dx, dy, dt, cfl, 0.0])
        out[key + "_bc"] = np.array(bcl + bcu, dtype=np.int64)
        out[key + "_qin"] = qbc
        out[key + "_aux"] = abc
        out[key + "_qout"] = qnew
        n += 1
    dst = os.path.join(HERE, "ref_step_vectors_aux.npz")
    np.savez_compressed(dst, **out)
    print(dst, n, "cases")


if __name__ == "__main__":
    main()
