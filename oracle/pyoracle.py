"""ctypes binding of the CPU oracle (oracle/cpu_ref.cpp).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as the checker,
and by bench.py for its CPU-baseline leg.  Field arrays use Grid's lexicographic order
with s innermost: shape (Lt, Lz, Ly, Lx or Lx/2, Ls, 4, 3) complex128.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

OP = dict(M=0, Mdag=1, Meooe=2, MeooeDag=3, Mooee=4, MooeeDag=5, MooeeInv=6, MooeeInvDag=7,
          M5D=8, M5Ddag=9, Meooe5D=10, MeooeDag5D=11, Dhop=12, DhopDag=13)
EVEN, ODD = 0, 1


class Params(ctypes.Structure):
    _fields_ = [('dims', ctypes.c_int * 4), ('Ls', ctypes.c_int), ('kind', ctypes.c_int),
                ('mass', ctypes.c_double), ('M5', ctypes.c_double), ('b', ctypes.c_double), ('c', ctypes.c_double),
                ('phase_re', ctypes.c_double * 4), ('phase_im', ctypes.c_double * 4)]


def build():
    subprocess.check_call(['make', '-s', '-C', HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(Params)
        _lib.oracle_double_store.argtypes = [pp, dp, dp]
        _lib.oracle_dhop.argtypes = [pp, dp, dp, dp, ctypes.c_int, ctypes.c_int]
        _lib.oracle_apply.argtypes = [pp, dp, dp, dp, ctypes.c_int, ctypes.c_int]
        _lib.oracle_apply.restype = ctypes.c_int
        _lib.oracle_schur.argtypes = [pp, dp, dp, dp, ctypes.c_int, ctypes.c_int]
        _lib.oracle_cayley_coeffs.argtypes = [pp, dp]
        _lib.oracle_cg.argtypes = [pp, dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), dp]
        _lib.oracle_cg.restype = ctypes.c_int
        _lib.oracle_schur_solve.argtypes = [pp, dp, dp, dp, ctypes.c_double, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int), dp, dp]
        _lib.oracle_schur_solve.restype = ctypes.c_int
        _lib.oracle_import_physical.argtypes = [pp, dp, dp, dp]
        _lib.oracle_export_physical.argtypes = [pp, dp, dp]
        _lib.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib.oracle_num_threads.restype = ctypes.c_int
    return _lib


def _dp(a):
    assert a.dtype == np.complex128 or a.dtype == np.float64
    assert a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One operator: kind 'wilson' (Ls copies of 4D Wilson; Ls = Nrhs for multi-RHS) or 'mobius'."""

    def __init__(self, U, dims, Ls=1, kind='wilson', mass=0.1, M5=1.8, b=1.5, c=0.5, phases=(1, 1, 1, 1)):
        self.dims = tuple(int(d) for d in dims)
        self.Ls = int(Ls)
        p = Params()
        for i in range(4):
            p.dims[i] = self.dims[i]
            p.phase_re[i] = complex(phases[i]).real
            p.phase_im[i] = complex(phases[i]).imag
        p.Ls = self.Ls
        p.kind = 0 if kind == 'wilson' else 1
        p.mass, p.M5, p.b, p.c = mass, M5, b, c
        self.p = p
        lx, ly, lz, lt = self.dims
        self.full_shape = (lt, lz, ly, lx, self.Ls, 4, 3)
        self.half_shape = (lt, lz, ly, lx // 2, self.Ls, 4, 3)
        U = np.ascontiguousarray(U, dtype=np.complex128)
        self.Uds = np.zeros((lt, lz, ly, lx, 8, 3, 3), dtype=np.complex128)
        lib().oracle_double_store(ctypes.byref(p), _dp(U), _dp(self.Uds))

    def _in(self, a, shape):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == shape, (a.shape, shape)
        return a

    def dhop(self, psi, cb_out=-1, dag=False):
        """cb_out = -1 full lattice; 0 (Even) = DhopEO from odd input; 1 = DhopOE."""
        shape = self.full_shape if cb_out < 0 else self.half_shape
        psi = self._in(psi, shape)
        out = np.zeros(shape, dtype=np.complex128)
        lib().oracle_dhop(ctypes.byref(self.p), _dp(self.Uds), _dp(psi), _dp(out), cb_out, int(dag))
        return out

    def apply(self, op, psi, cb=ODD):
        full = op in ('M', 'Mdag')
        shape = self.full_shape if full else self.half_shape
        psi = self._in(psi, shape)
        out = np.zeros(shape, dtype=np.complex128)
        rc = lib().oracle_apply(ctypes.byref(self.p), _dp(self.Uds), _dp(psi), _dp(out), OP[op], cb)
        assert rc == 0
        return out

    def schur(self, which, psi, cb=ODD):
        psi = self._in(psi, self.half_shape)
        out = np.zeros(self.half_shape, dtype=np.complex128)
        lib().oracle_schur(ctypes.byref(self.p), _dp(self.Uds), _dp(psi), _dp(out),
                           dict(Mpc=0, MpcDag=1, MpcDagMpc=2)[which], cb)
        return out

    def import_physical(self, in4d):
        lx, ly, lz, lt = self.dims
        in4d = self._in(in4d, (lt, lz, ly, lx, 1, 4, 3))
        out = np.zeros(self.full_shape, dtype=np.complex128)
        lib().oracle_import_physical(ctypes.byref(self.p), _dp(self.Uds), _dp(in4d), _dp(out))
        return out

    def export_physical(self, sol5d):
        lx, ly, lz, lt = self.dims
        sol5d = self._in(sol5d, self.full_shape)
        out = np.zeros((lt, lz, ly, lx, 1, 4, 3), dtype=np.complex128)
        lib().oracle_export_physical(ctypes.byref(self.p), _dp(sol5d), _dp(out))
        return out

    def coeffs(self):
        out = np.zeros(9 * self.Ls)
        lib().oracle_cayley_coeffs(ctypes.byref(self.p), _dp(out))
        names = ['bs', 'cs', 'bee', 'cee', 'dee', 'lee', 'leem', 'uee', 'ueem']
        return {n: out[k * self.Ls:(k + 1) * self.Ls].copy() for k, n in enumerate(names)}

    def cg(self, src, tol=1e-10, maxit=10000, cb=ODD, guess=None):
        src = self._in(src, self.half_shape)
        psi = np.zeros(self.half_shape, dtype=np.complex128) if guess is None else self._in(guess, self.half_shape).copy()
        it = ctypes.c_int(0)
        tr = ctypes.c_double(0)
        rc = lib().oracle_cg(ctypes.byref(self.p), _dp(self.Uds), _dp(src), _dp(psi), cb, tol, maxit,
                             ctypes.byref(it), ctypes.byref(tr))
        return psi, dict(converged=(rc == 0), iterations=it.value, true_residual=tr.value)

    def schur_solve(self, src, tol=1e-10, maxit=10000):
        src = self._in(src, self.full_shape)
        sol = np.zeros(self.full_shape, dtype=np.complex128)
        it = ctypes.c_int(0)
        tr = ctypes.c_double(0)
        ur = ctypes.c_double(0)
        rc = lib().oracle_schur_solve(ctypes.byref(self.p), _dp(self.Uds), _dp(src), _dp(sol), tol, maxit,
                                      ctypes.byref(it), ctypes.byref(tr), ctypes.byref(ur))
        return sol, dict(converged=(rc == 0), iterations=it.value, true_residual=tr.value, unprec_residual=ur.value)


def set_threads(n):
    lib().oracle_set_threads(int(n))


def num_threads():
    return lib().oracle_num_threads()
