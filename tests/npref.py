"""Independent NumPy restatement of the Wilson / Mobius operators in closed form.

Test infrastructure: it is formulated differently from the projector-form oracle
(oracle/cpu_ref.cpp) so that each pins the other.

* Wilson hop, closed form with explicit 4x4 gamma matrices
  D psi(x) = -1/2 sum_mu [(1 - g_mu) U_mu(x) psi(x+mu) + (1 + g_mu) U_mu(x-mu)^dag psi(x-mu)]
  (Grid: benchmarks/Benchmark_dwf_fp32.cc:205-239; gamma basis Grid/qcd/spin/TwoSpinor.h:67-139).
* Mobius as  M = (b DW + 1) + (c DW - 1) S  with  DW = Dhop + (4 - M5)  and the
  chiral s-shift S = P- S_up + P+ S_down carrying -mf on the wrap-around entries
  (Grid: CayleyFermion5DImplementation.h:155-329, MobiusFermion.h:45-71).
* Mooee^{-1} by dense Ls x Ls inverses per chirality (numpy.linalg.inv), independent
  of the LDU sweeps (CayleyFermion5Dcache.h:117-230).

Array layout: fields are (Lt, Lz, Ly, Lx, Ls, 4, 3) complex128, gauge (Lt, Lz, Ly, Lx, 4, 3, 3),
so a C-order flatten gives Grid's lexicographic order (x fastest, s innermost).
"""
import numpy as np

I2 = 1j
GAMMA = [
    np.array([[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]]),
    np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], dtype=complex),
    np.array([[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]]),
    np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex),
]
GAMMA5 = np.diag([1, 1, -1, -1]).astype(complex)
PPLUS = 0.5 * (np.eye(4) + GAMMA5)
PMINUS = 0.5 * (np.eye(4) - GAMMA5)


def axis_of(mu):
    # array axes are (t, z, y, x): direction mu=0 (x) is axis 3
    return 3 - mu


def random_su3(rng, shape, eps):
    """exp(i eps H), H Hermitian traceless with Gaussian(0,1) entries, by eigendecomposition."""
    g = rng.standard_normal(shape + (3, 3)) + 1j * rng.standard_normal(shape + (3, 3))
    h = 0.5 * (g + np.conj(np.swapaxes(g, -1, -2)))
    tr = np.trace(h, axis1=-2, axis2=-1).real / 3.0
    h = h - tr[..., None, None] * np.eye(3)
    w, v = np.linalg.eigh(h)
    return np.einsum('...ab,...b,...cb->...ac', v, np.exp(1j * eps * w), np.conj(v))


def random_gauge(rng, dims, eps):
    lx, ly, lz, lt = dims
    return random_su3(rng, (lt, lz, ly, lx, 4), eps)


def gaussian_fermion(rng, dims, ls):
    lx, ly, lz, lt = dims
    shp = (lt, lz, ly, lx, ls, 4, 3)
    return rng.standard_normal(shp) + 1j * rng.standard_normal(shp)


def apply_phases(U, phases):
    """Multiply U_mu at x_mu = L_mu - 1 by phases[mu] (Grid WilsonImpl.h:143-170)."""
    U = U.copy()
    for mu in range(4):
        if phases[mu] == 1:
            continue
        ax = axis_of(mu)
        idx = [slice(None)] * 4
        idx[ax] = U.shape[ax] - 1
        U[tuple(idx) + (mu,)] *= phases[mu]
    return U


def spin_mul(m, psi):
    return np.einsum('ij,...jc->...ic', m, psi)


def dhop(U, psi, dag=False):
    out = np.zeros_like(psi)
    sgn = 1.0 if dag else -1.0
    for mu in range(4):
        ax = axis_of(mu)
        Umu = U[..., mu, :, :][..., None, :, :]          # broadcast over s
        fwd = np.roll(psi, -1, axis=ax)                   # psi(x+mu)
        t1 = np.einsum('...ab,...jb->...ja', Umu, fwd)
        Ub = np.roll(U[..., mu, :, :], 1, axis=ax)[..., None, :, :]  # U_mu(x-mu)
        bwd = np.roll(psi, 1, axis=ax)
        t2 = np.einsum('...ba,...jb->...ja', np.conj(Ub), bwd)
        P1 = np.eye(4) + sgn * GAMMA[mu]
        P2 = np.eye(4) - sgn * GAMMA[mu]
        out += -0.5 * (spin_mul(P1, t1) + spin_mul(P2, t2))
    return out


def parity_mask(dims):
    lx, ly, lz, lt = dims
    t, z, y, x = np.meshgrid(np.arange(lt), np.arange(lz), np.arange(ly), np.arange(lx), indexing='ij')
    return ((x + y + z + t) % 2).astype(np.int64)


def pick(field, dims, cb):
    """Full field -> half field in Grid's red-black order (x halved)."""
    lx, ly, lz, lt = dims
    par = parity_mask(dims)
    out = np.zeros((lt, lz, ly, lx // 2) + field.shape[4:], dtype=field.dtype)
    for t in range(lt):
        for z in range(lz):
            for y in range(ly):
                xs = [x for x in range(lx) if par[t, z, y, x] == cb]
                out[t, z, y, :] = field[t, z, y, xs]
    return out


def place(half, dims, cb):
    """Half field -> full field with zeros on the other checkerboard."""
    lx, ly, lz, lt = dims
    par = parity_mask(dims)
    out = np.zeros((lt, lz, ly, lx) + half.shape[4:], dtype=half.dtype)
    for t in range(lt):
        for z in range(lz):
            for y in range(ly):
                xs = [x for x in range(lx) if par[t, z, y, x] == cb]
                out[t, z, y, xs] = half[t, z, y, :]
    return out


# ---------------- 5th dimension (Mobius) ----------------
def s_shift(psi, mf, dag=False):
    """S = P- S_up + P+ S_down (S^dag = P- S_down + P+ S_up), wrap entries carry -mf."""
    ls = psi.shape[4]
    up = np.roll(psi, -1, axis=4)     # psi_{s+1}
    up[..., ls - 1, :, :] *= -mf
    dn = np.roll(psi, 1, axis=4)      # psi_{s-1}
    dn[..., 0, :, :] *= -mf
    if not dag:
        return spin_mul(PMINUS, up) + spin_mul(PPLUS, dn)
    return spin_mul(PMINUS, dn) + spin_mul(PPLUS, up)


def mobius_M(U, psi, mf, M5, b, c, dag=False):
    dw = lambda f: dhop(U, f, dag) + (4.0 - M5) * f
    if not dag:
        return dw(b * psi + c * s_shift(psi, mf)) + psi - s_shift(psi, mf)
    x = dw(psi)
    return b * x + psi + s_shift(c * x - psi, mf, dag=True)


def mobius_mooee(psi, mf, M5, b, c, dag=False):
    bee = b * (4.0 - M5) + 1.0
    cee = 1.0 - c * (4.0 - M5)
    return bee * psi - cee * s_shift(psi, mf, dag)


def mobius_meooe(U, psi_full, mf, b, c, dag=False):
    """Hop part of M (whole lattice, caller masks checkerboards)."""
    if not dag:
        return dhop(U, b * psi_full + c * s_shift(psi_full, mf), False)
    x = dhop(U, psi_full, True)
    return b * x + c * s_shift(x, mf, dag=True)


def mooee_dense(ls, mf, M5, b, c):
    """Ls x Ls matrices of Mooee on the P+ (upper spins) and P- (lower spins) components."""
    bee = b * (4.0 - M5) + 1.0
    cee = 1.0 - c * (4.0 - M5)
    Sup = np.zeros((ls, ls))
    Sdn = np.zeros((ls, ls))
    for s in range(ls):
        Sup[s, (s + 1) % ls] = -mf if s == ls - 1 else 1.0
        Sdn[s, (s - 1) % ls] = -mf if s == 0 else 1.0
    Aplus = bee * np.eye(ls) - cee * Sdn     # P+ part sees S_down
    Aminus = bee * np.eye(ls) - cee * Sup    # P- part sees S_up
    return Aplus, Aminus


def mobius_mooee_inv(psi, mf, M5, b, c, dag=False):
    ls = psi.shape[4]
    Ap, Am = mooee_dense(ls, mf, M5, b, c)
    if dag:
        Ap, Am = Ap.T, Am.T
    Bp, Bm = np.linalg.inv(Ap), np.linalg.inv(Am)
    out = np.zeros_like(psi)
    out[..., 0:2, :] = np.einsum('st,...tjc->...sjc', Bp, psi[..., 0:2, :])
    out[..., 2:4, :] = np.einsum('st,...tjc->...sjc', Bm, psi[..., 2:4, :])
    return out


def import_physical(U, in4d, ls, M5, c):
    """5D source from a 4D one: P+ b at s=0, P- b at s=Ls-1, then (1 - c DW)."""
    shp = in4d.shape[:4] + (ls, 4, 3)
    tmp = np.zeros(shp, dtype=complex)
    tmp[..., 0, :, :] = spin_mul(PPLUS, in4d[..., 0, :, :])
    tmp[..., ls - 1, :, :] = spin_mul(PMINUS, in4d[..., 0, :, :])
    return tmp - c * (dhop(U, tmp) + (4.0 - M5) * tmp)


def export_physical(sol5d):
    """4D propagator from a 5D solution: P- psi(0) + P+ psi(Ls-1)."""
    ls = sol5d.shape[4]
    out = spin_mul(PMINUS, sol5d[..., 0, :, :]) + spin_mul(PPLUS, sol5d[..., ls - 1, :, :])
    return out[..., None, :, :]


def wilson_M(U, psi, mass, dag=False):
    return dhop(U, psi, dag) + (4.0 + mass) * psi


def to_flat(field):
    return np.ascontiguousarray(field).reshape(-1)


def gauge_flat(U):
    return np.ascontiguousarray(U).reshape(-1)
