"""Writes the golden fixtures under tests/golden/ from the closed-form NumPy restatement
(tests/npref.py).  Grid's own tests hold no golden vectors for this path (they assert
identities and print norms), so these fixtures are produced by a second, independent
formulation: explicit gamma matrices and np.roll for the hop, dense Ls x Ls inverses for
Mooee^{-1}, a dense LU solve for the Wilson propagator.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import npref as R  # noqa: E402


def wilson_case():
    dims = (4, 4, 4, 4)
    rng = np.random.default_rng(1001)
    U = R.random_gauge(rng, dims, 0.5)
    psi = R.gaussian_fermion(rng, dims, 1)
    mass = 0.1
    out = dict(dims=np.array(dims), Ls=np.array(1), mass=np.array(mass), eps=np.array(0.5))
    out['U'] = U
    out['psi'] = psi
    out['dhop'] = R.dhop(U, psi)
    out['dhop_dag'] = R.dhop(U, psi, dag=True)
    out['M'] = R.wilson_M(U, psi, mass)
    out['Mdag'] = R.wilson_M(U, psi, mass, dag=True)
    # antiperiodic time boundary
    Uap = R.apply_phases(U, [1, 1, 1, -1])
    out['dhop_antiperiodic_t'] = R.dhop(Uap, psi)
    # even-odd pieces: DhopEO (odd -> even) and DhopOE (even -> odd)
    psi_o = R.pick(psi, dims, 1)
    psi_e = R.pick(psi, dims, 0)
    out['dhop_eo'] = R.pick(R.dhop(U, R.place(psi_o, dims, 1)), dims, 0)
    out['dhop_oe'] = R.pick(R.dhop(U, R.place(psi_e, dims, 0)), dims, 1)
    out['dhop_eo_dag'] = R.pick(R.dhop(U, R.place(psi_o, dims, 1), dag=True), dims, 0)
    # dense solve M x = psi
    n = psi.size
    cols = []
    for k in range(n):
        e = np.zeros(n, dtype=complex)
        e[k] = 1.0
        cols.append(R.wilson_M(U, e.reshape(psi.shape), mass).reshape(-1))
    Mdense = np.stack(cols, axis=1)
    out['x'] = np.linalg.solve(Mdense, psi.reshape(-1)).reshape(psi.shape)
    return out


def mobius_case():
    dims = (4, 2, 2, 4)
    ls = 6
    mf, M5, b, c = 0.1, 1.8, 1.5, 0.5
    rng = np.random.default_rng(2002)
    U = R.random_gauge(rng, dims, 0.3)
    psi = R.gaussian_fermion(rng, dims, ls)
    out = dict(dims=np.array(dims), Ls=np.array(ls), mass=np.array(mf), M5=np.array(M5), b=np.array(b), c=np.array(c))
    out['U'] = U
    out['psi'] = psi
    out['M'] = R.mobius_M(U, psi, mf, M5, b, c)
    out['Mdag'] = R.mobius_M(U, psi, mf, M5, b, c, dag=True)
    psi_o = R.pick(psi, dims, 1)
    psi_e = R.pick(psi, dims, 0)
    full_o = R.place(psi_o, dims, 1)
    out['meooe_o'] = R.pick(R.mobius_meooe(U, full_o, mf, b, c), dims, 0)
    out['meooedag_o'] = R.pick(R.mobius_meooe(U, full_o, mf, b, c, dag=True), dims, 0)
    out['mooee_e'] = R.mobius_mooee(psi_e, mf, M5, b, c)
    out['mooeedag_e'] = R.mobius_mooee(psi_e, mf, M5, b, c, dag=True)
    out['mooeeinv_e'] = R.mobius_mooee_inv(psi_e, mf, M5, b, c)
    out['mooeeinvdag_e'] = R.mobius_mooee_inv(psi_e, mf, M5, b, c, dag=True)

    def meooe(h, cb_in, dag):
        return R.pick(R.mobius_meooe(U, R.place(h, dims, cb_in), mf, b, c, dag), dims, 1 - cb_in)

    # Schur Mpc on the odd checkerboard
    t = meooe(psi_o, 1, False)
    t = R.mobius_mooee_inv(t, mf, M5, b, c)
    t = meooe(t, 0, False)
    out['mpc_o'] = R.mobius_mooee(psi_o, mf, M5, b, c) - t
    t = meooe(psi_o, 1, True)
    t = R.mobius_mooee_inv(t, mf, M5, b, c, dag=True)
    t = meooe(t, 0, True)
    out['mpcdag_o'] = R.mobius_mooee(psi_o, mf, M5, b, c, dag=True) - t
    return out


def main():
    np.savez_compressed(os.path.join(HERE, 'wilson_4x4x4x4.npz'), **wilson_case())
    np.savez_compressed(os.path.join(HERE, 'mobius_4x2x2x4_ls6.npz'), **mobius_case())
    print('wrote fixtures')


if __name__ == '__main__':
    main()
