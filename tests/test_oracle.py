"""Pins the CPU oracle (oracle/cpu_ref.cpp) before anything is compared with it.

Checks, each chosen so that a plausible slip (a sign, a projector, a link direction, an
EO index, an LDU coefficient) fails at least one of them:
  * golden fixtures from the closed-form NumPy restatement (tests/npref.py);
  * Grid's own identities: Deo + Doe = D (Test_wilson_even_odd.cc:100-118), adjointness
    (:120-157), MooeeInv Mooee = 1 (:159-195), MpcDagMpc Hermitian (:197-224),
    Meo + Moe + Moo + Mee = M (Test_mobius_even_odd.cc:121-145);
  * gamma5-hermiticity, gauge covariance, the free-field plane-wave closed form
    (WilsonFermionImplementation.h:177-230 momentum-space form);
  * b=1, c=0 Mobius reduces to Shamir DWF (DomainWallFermion.h:125-130);
  * the CG solution against a dense LU solve (Wilson 4^4).
"""
import numpy as np
import pytest

import npref as R
from conftest import rel_err, max_site_rel_err
from oracle.pyoracle import Oracle, EVEN, ODD

TOL64 = 1e-13


def test_gamma_basis():
    g = R.GAMMA
    for mu in range(4):
        for nu in range(4):
            anti = g[mu] @ g[nu] + g[nu] @ g[mu]
            assert np.allclose(anti, 2 * np.eye(4) * (mu == nu))
        assert np.allclose(g[mu], g[mu].conj().T)
    assert np.allclose(g[0] @ g[1] @ g[2] @ g[3], R.GAMMA5)


def test_wilson_dhop_golden(wilson_golden):
    G = wilson_golden
    O = Oracle(G['U'], G['dims'], 1, 'wilson', mass=float(G['mass']))
    assert max_site_rel_err(O.dhop(G['psi']), G['dhop']) < TOL64
    assert max_site_rel_err(O.dhop(G['psi'], dag=True), G['dhop_dag']) < TOL64
    assert max_site_rel_err(O.apply('M', G['psi']), G['M']) < TOL64
    assert max_site_rel_err(O.apply('Mdag', G['psi']), G['Mdag']) < TOL64
    dims = tuple(G['dims'])
    psi_o = R.pick(G['psi'], dims, ODD)
    psi_e = R.pick(G['psi'], dims, EVEN)
    assert max_site_rel_err(O.dhop(psi_o, cb_out=EVEN), G['dhop_eo']) < TOL64
    assert max_site_rel_err(O.dhop(psi_e, cb_out=ODD), G['dhop_oe']) < TOL64
    assert max_site_rel_err(O.dhop(psi_o, cb_out=EVEN, dag=True), G['dhop_eo_dag']) < TOL64
    assert max_site_rel_err(O.apply('Meooe', psi_o, cb=ODD), G['dhop_eo']) < TOL64


def test_wilson_antiperiodic_golden(wilson_golden):
    G = wilson_golden
    O = Oracle(G['U'], G['dims'], 1, 'wilson', phases=(1, 1, 1, -1))
    assert max_site_rel_err(O.dhop(G['psi']), G['dhop_antiperiodic_t']) < TOL64


def test_wilson_solve_golden(wilson_golden):
    G = wilson_golden
    O = Oracle(G['U'], G['dims'], 1, 'wilson', mass=float(G['mass']))
    sol, st = O.schur_solve(G['psi'], tol=1e-12)
    assert st['converged']
    assert st['unprec_residual'] < 1e-10
    assert rel_err(sol, G['x']) < 1e-9


def test_mobius_golden(mobius_golden):
    G = mobius_golden
    dims = tuple(G['dims'])
    O = Oracle(G['U'], dims, int(G['Ls']), 'mobius', mass=float(G['mass']), M5=float(G['M5']),
               b=float(G['b']), c=float(G['c']))
    assert max_site_rel_err(O.apply('M', G['psi']), G['M']) < TOL64
    assert max_site_rel_err(O.apply('Mdag', G['psi']), G['Mdag']) < TOL64
    psi_o = R.pick(G['psi'], dims, ODD)
    psi_e = R.pick(G['psi'], dims, EVEN)
    assert max_site_rel_err(O.apply('Meooe', psi_o, ODD), G['meooe_o']) < TOL64
    assert max_site_rel_err(O.apply('MeooeDag', psi_o, ODD), G['meooedag_o']) < TOL64
    assert max_site_rel_err(O.apply('Mooee', psi_e, EVEN), G['mooee_e']) < TOL64
    assert max_site_rel_err(O.apply('MooeeDag', psi_e, EVEN), G['mooeedag_e']) < TOL64
    assert max_site_rel_err(O.apply('MooeeInv', psi_e, EVEN), G['mooeeinv_e']) < TOL64
    assert max_site_rel_err(O.apply('MooeeInvDag', psi_e, EVEN), G['mooeeinvdag_e']) < TOL64
    assert max_site_rel_err(O.schur('Mpc', psi_o), G['mpc_o']) < TOL64
    assert max_site_rel_err(O.schur('MpcDag', psi_o), G['mpcdag_o']) < TOL64


@pytest.fixture(scope='module')
def lat8():
    rng = np.random.default_rng(7)
    dims = (8, 4, 4, 8)
    U = R.random_gauge(rng, dims, 0.5)
    return dims, U, rng


def test_eo_decomposition_and_adjointness(lat8):
    dims, U, rng = lat8
    O = Oracle(U, dims, 1, 'wilson', mass=0.1)
    psi = R.gaussian_fermion(rng, dims, 1)
    phi = R.gaussian_fermion(rng, dims, 1)
    full = O.dhop(psi)
    deo = O.dhop(R.pick(psi, dims, ODD), cb_out=EVEN)
    doe = O.dhop(R.pick(psi, dims, EVEN), cb_out=ODD)
    recon = R.place(deo, dims, EVEN) + R.place(doe, dims, ODD)
    assert rel_err(recon, full) < TOL64
    # <phi|D chi>^* = <chi|D^dag phi>
    a = np.vdot(phi, O.dhop(psi))
    b = np.vdot(psi, O.dhop(phi, dag=True))
    assert abs(np.conj(a) - b) / abs(a) < 1e-13
    # gamma5 hermiticity: D^dag = g5 D g5
    g5 = lambda f: R.spin_mul(R.GAMMA5, f)
    assert rel_err(O.dhop(psi, dag=True), g5(O.dhop(g5(psi)))) < TOL64


def test_gauge_covariance(lat8):
    dims, U, rng = lat8
    lx, ly, lz, lt = dims
    g = R.random_su3(rng, (lt, lz, ly, lx), 1.0)
    Ug = np.empty_like(U)
    for mu in range(4):
        gshift = np.roll(g, -1, axis=R.axis_of(mu))
        Ug[..., mu, :, :] = g @ U[..., mu, :, :] @ np.conj(np.swapaxes(gshift, -1, -2))
    psi = R.gaussian_fermion(rng, dims, 1)
    gpsi = np.einsum('...ab,...sjb->...sja', g, psi)
    lhs = Oracle(Ug, dims).dhop(gpsi)
    rhs = np.einsum('...ab,...sjb->...sja', g, Oracle(U, dims).dhop(psi))
    assert rel_err(lhs, rhs) < TOL64


def test_free_field_plane_wave():
    dims = (8, 4, 6, 4)
    lx, ly, lz, lt = dims
    U = np.zeros((lt, lz, ly, lx, 4, 3, 3), dtype=complex)
    U[...] = np.eye(3)
    n = (1, 2, 1, 3)
    p = [2 * np.pi * n[mu] / dims[mu] for mu in range(4)]
    t, z, y, x = np.meshgrid(np.arange(lt), np.arange(lz), np.arange(ly), np.arange(lx), indexing='ij')
    phase = np.exp(1j * (p[0] * x + p[1] * y + p[2] * z + p[3] * t))
    rng = np.random.default_rng(3)
    u = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    psi = (phase[..., None, None, None] * u[None, None, None, None, None]).astype(complex)
    # D e^{ipx} u = -sum_mu [cos p_mu - i gamma_mu sin p_mu] e^{ipx} u
    K = -sum(np.cos(p[mu]) * np.eye(4) - 1j * R.GAMMA[mu] * np.sin(p[mu]) for mu in range(4))
    expect = phase[..., None, None, None] * (K @ u)[None, None, None, None, None]
    assert rel_err(Oracle(U, dims).dhop(psi), expect) < TOL64


def test_mobius_identities(lat8):
    dims, U, rng = lat8
    Ls = 8
    O = Oracle(U, dims, Ls, 'mobius', mass=0.01, M5=1.8, b=1.5, c=0.5)
    psi = R.gaussian_fermion(rng, dims, Ls)
    pe, po = R.pick(psi, dims, EVEN), R.pick(psi, dims, ODD)
    # Meo + Moe + Mee + Moo = M
    Mfull = O.apply('M', psi)
    r_e = O.apply('Mooee', pe, EVEN) + O.apply('Meooe', po, ODD)
    r_o = O.apply('Mooee', po, ODD) + O.apply('Meooe', pe, EVEN)
    assert rel_err(R.place(r_e, dims, EVEN) + R.place(r_o, dims, ODD), Mfull) < TOL64
    # MooeeInv Mooee = 1 and the dagger forms
    assert rel_err(O.apply('MooeeInv', O.apply('Mooee', pe, EVEN), EVEN), pe) < 1e-14
    assert rel_err(O.apply('MooeeInvDag', O.apply('MooeeDag', pe, EVEN), EVEN), pe) < 1e-14
    # dagger consistency <phi|M psi>^* = <psi|Mdag phi>
    phi = R.gaussian_fermion(rng, dims, Ls)
    a = np.vdot(phi, Mfull)
    b = np.vdot(psi, O.apply('Mdag', phi))
    assert abs(np.conj(a) - b) / abs(a) < 1e-13
    phio = R.pick(phi, dims, ODD)
    a = np.vdot(phio, O.schur('Mpc', po))
    b = np.vdot(po, O.schur('MpcDag', phio))
    assert abs(np.conj(a) - b) / abs(a) < 1e-13
    # MpcDagMpc Hermitian, positive
    h1 = np.vdot(phio, O.schur('MpcDagMpc', po))
    h2 = np.vdot(O.schur('MpcDagMpc', phio), po)
    assert abs(h1 - h2) / abs(h1) < 1e-13
    assert np.vdot(po, O.schur('MpcDagMpc', po)).real > 0


def test_fifth_dimension_ops_closed_form(lat8):
    """The standalone fifth-dimension operators against their closed forms, independently of the
    compositions (CayleyFermion5DImplementation.h:155-271 with Mobius bs = b, cs = c): M5D(psi, chi)
    adds the off-diagonal part -S psi to chi (upper -1 with mass_minus at s = Ls-1, lower -1 with
    mass_plus at s = 0); the oracle's entry point starts chi from psi, giving (1 - S) psi.
    Meooe5D = b + c S, and the dagger forms with S^dag."""
    dims, U, rng = lat8
    Ls, mf, b, c = 6, 0.07, 1.5, 0.5
    O = Oracle(U, dims, Ls, 'mobius', mass=mf, M5=1.8, b=b, c=c)
    po = R.pick(R.gaussian_fermion(rng, dims, Ls), dims, ODD)
    for dag, m5, meo in ((False, 'M5D', 'Meooe5D'), (True, 'M5Ddag', 'MeooeDag5D')):
        S = R.s_shift(po, mf, dag=dag)
        assert rel_err(O.apply(m5, po, ODD), po - S) < TOL64, m5
        assert rel_err(O.apply(meo, po, ODD), b * po + c * S) < TOL64, meo
    # a mistaken corner (mass on the wrong end) or a swapped chirality fails this: the two differ
    assert rel_err(R.s_shift(po, mf, dag=True), R.s_shift(po, mf)) > 0.1


def test_mobius_reduces_to_shamir(lat8):
    dims, U, rng = lat8
    Ls, mf, M5 = 6, 0.05, 1.8
    O = Oracle(U, dims, Ls, 'mobius', mass=mf, M5=M5, b=1.0, c=0.0)
    psi = R.gaussian_fermion(rng, dims, Ls)
    # Shamir: D = (DW + 1) - P- psi_{s+1} - P+ psi_{s-1}, with +mf on the wrap terms
    dw = R.dhop(U, psi) + (4.0 - M5) * psi
    expect = dw + psi - R.s_shift(psi, mf)
    assert rel_err(O.apply('M', psi), expect) < TOL64


def test_physical_source_maps(lat8):
    """Import (P+ at s=0, P- at s=Ls-1, then 1 - c DW) and export (P- psi(0) + P+ psi(Ls-1))."""
    dims, U, rng = lat8
    Ls, M5, c = 6, 1.8, 0.5
    O = Oracle(U, dims, Ls, 'mobius', mass=0.05, M5=M5, b=1.5, c=c)
    b4 = R.gaussian_fermion(rng, dims, 1)
    assert rel_err(O.import_physical(b4), R.import_physical(U, b4, Ls, M5, c)) < TOL64
    sol = R.gaussian_fermion(rng, dims, Ls)
    assert rel_err(O.export_physical(sol), R.export_physical(sol)) < 1e-15
    # with c = 0 the import is the bare chiral embedding and export undoes it
    O0 = Oracle(U, dims, Ls, 'mobius', mass=0.05, M5=M5, b=1.0, c=0.0)
    emb = O0.import_physical(b4)
    flipped = np.zeros_like(emb)
    flipped[..., 0, :, :] = emb[..., Ls - 1, :, :]
    flipped[..., Ls - 1, :, :] = emb[..., 0, :, :]
    assert rel_err(O0.export_physical(flipped), b4) < 1e-15


def test_cayley_coefficients_mobius():
    U = np.zeros((2, 2, 2, 2, 4, 3, 3), dtype=complex)
    U[...] = np.eye(3)
    O = Oracle(U, (2, 2, 2, 2), 16, 'mobius', mass=0.01, M5=1.8, b=1.5, c=0.5)
    c = O.coeffs()
    assert np.allclose(c['bs'], 1.5) and np.allclose(c['cs'], 0.5)
    assert np.allclose(c['bee'], 4.3) and np.allclose(c['cee'], -0.1)


def test_cg_matches_dense_schur(lat8):
    dims, U, rng = lat8
    O = Oracle(U, dims, 1, 'wilson', mass=0.1)
    src = R.pick(R.gaussian_fermion(rng, dims, 1), dims, ODD)
    sol, st = O.cg(src, tol=1e-10)
    assert st['converged'] and st['true_residual'] < 1e-9
    r = O.schur('MpcDagMpc', sol) - src
    assert np.linalg.norm(r) / np.linalg.norm(src) < 1e-9


def test_mobius_schur_solve(lat8):
    dims, U, rng = lat8
    Ls = 6
    O = Oracle(U, dims, Ls, 'mobius', mass=0.05, M5=1.8, b=1.5, c=0.5)
    src = R.gaussian_fermion(rng, dims, Ls)
    sol, st = O.schur_solve(src, tol=1e-10)
    assert st['converged']
    # residual with the independently formulated closed-form M
    r = R.mobius_M(U, sol, 0.05, 1.8, 1.5, 0.5) - src
    assert np.linalg.norm(r) / np.linalg.norm(src) < 1e-8
