"""grid_amd: Python binding of libgx, the MI355X (gfx950) Wilson / Mobius domain-wall Dirac
operator and even-odd CG library (C ABI in include/gx.h).

The binding is a thin ctypes layer: every call goes to the HIP library.  There is no CPU
fallback; importing works without a GPU, but creating a context needs one and raises
otherwise.

Host arrays use Grid's lexicographic order with s innermost, complex128:
  full field (Lt, Lz, Ly, Lx, Ls, 4, 3), half field (Lt, Lz, Ly, Lx/2, Ls, 4, 3),
  gauge (Lt, Lz, Ly, Lx, 4, 3, 3).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libgx.so')

F64, F32, F16 = 0, 1, 2
FULL, EVEN, ODD = -1, 0, 1
WILSON, MOBIUS = 0, 1
HERM_SCHUR, HERM_MDAGM = 0, 1
STATUS = {0: 'GX_OK', 1: 'GX_EINVAL', 2: 'GX_ENOMEM', 3: 'GX_ENOCONV', 4: 'GX_EHIP', 5: 'GX_ECB', 6: 'GX_ENAN'}


class GxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'{STATUS.get(code, code)}: {msg}')
        self.code = code


class SolverStats(ctypes.Structure):
    _fields_ = [('iterations', ctypes.c_int), ('converged', ctypes.c_int), ('true_residual', ctypes.c_double),
                ('computed_residual', ctypes.c_double), ('seconds', ctypes.c_double),
                ('inner_iterations', ctypes.c_int), ('outer_iterations', ctypes.c_int),
                ('final_iterations', ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class NerscInfo(ctypes.Structure):
    _fields_ = [('dims', ctypes.c_int * 4), ('two_row', ctypes.c_int), ('checksum_header', ctypes.c_uint32),
                ('checksum_data', ctypes.c_uint32), ('plaquette_header', ctypes.c_double),
                ('plaquette', ctypes.c_double), ('link_trace_header', ctypes.c_double),
                ('link_trace', ctypes.c_double), ('sequence_number', ctypes.c_int),
                ('floating_point', ctypes.c_char * 16), ('ensemble_id', ctypes.c_char * 64),
                ('ensemble_label', ctypes.c_char * 64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d['dims'] = tuple(self.dims)
        for k in ('floating_point', 'ensemble_id', 'ensemble_label'):
            d[k] = d[k].decode()
        return d


class IldgInfo(ctypes.Structure):
    _fields_ = [('dims', ctypes.c_int * 4), ('precision', ctypes.c_int),
                ('scidac_checksuma', ctypes.c_uint32), ('scidac_checksumb', ctypes.c_uint32),
                ('scidac_checksuma_file', ctypes.c_uint32), ('scidac_checksumb_file', ctypes.c_uint32),
                ('has_plaquette', ctypes.c_int), ('plaquette_file', ctypes.c_double),
                ('link_trace_file', ctypes.c_double), ('plaquette', ctypes.c_double), ('link_trace', ctypes.c_double),
                ('has_norm', ctypes.c_int), ('norm2_file', ctypes.c_double), ('norm2', ctypes.c_double),
                ('lfn', ctypes.c_char * 256)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d['dims'] = tuple(self.dims)
        d['lfn'] = d['lfn'].decode()
        return d


class Request(ctypes.Structure):
    _fields_ = [('dims', ctypes.c_int * 4), ('nsrc', ctypes.c_int), ('mass', ctypes.c_double),
                ('gauge', ctypes.c_int), ('seed', ctypes.c_uint64), ('arrival', ctypes.c_double)]


class Response(ctypes.Structure):
    _fields_ = [('status', ctypes.c_int), ('slot', ctypes.c_int), ('iterations', ctypes.c_int),
                ('outer_iterations', ctypes.c_int), ('residual', ctypes.c_double),
                ('solver_residual', ctypes.c_double), ('t_admit', ctypes.c_double),
                ('t_done', ctypes.c_double), ('footprint', ctypes.c_double), ('digest', ctypes.c_uint64),
                ('batch', ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeParams(ctypes.Structure):
    _fields_ = [('device', ctypes.c_int), ('slots', ctypes.c_int), ('budget', ctypes.c_double),
                ('n_gauge', ctypes.c_int), ('gauge_seed', ctypes.c_uint64), ('gauge_eps', ctypes.c_double),
                ('tol', ctypes.c_double), ('mixed', ctypes.c_int), ('maxit', ctypes.c_int),
                ('realtime', ctypes.c_int), ('progress', ctypes.c_int), ('gauge_dir', ctypes.c_char_p),
                ('fuse_rhs', ctypes.c_int)]


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_SIG = {
    'gx_last_error': ([], ctypes.c_char_p),
    'gx_version': ([], ctypes.c_char_p),
    'gx_ctx_create': ([_I, ctypes.POINTER(_P)], _I),
    'gx_ctx_destroy': ([_P], _I),
    'gx_ctx_sync': ([_P], _I),
    'gx_ctx_event_record': ([_P, _I], _I),
    'gx_ctx_event_elapsed': ([_P, _I, _I, _DP], _I),
    'gx_field_create': ([_P, _IP, _I, _I, _I, ctypes.POINTER(_P)], _I),
    'gx_field_destroy': ([_P], _I),
    'gx_field_import': ([_P, _DP], _I),
    'gx_field_export': ([_P, _DP], _I),
    'gx_field_checkerboard': ([_P, _IP], _I),
    'gx_field_set_checkerboard_tag': ([_P, _I], _I),
    'gx_field_zero': ([_P], _I),
    'gx_field_copy': ([_P, _P], _I),
    'gx_pick_checkerboard': ([_I, _P, _P], _I),
    'gx_set_checkerboard': ([_P, _P], _I),
    'gx_precision_change': ([_P, _P], _I),
    'gx_axpy': ([_P, _D, _P, _P], _I),
    'gx_axpby': ([_P, _D, _D, _P, _P], _I),
    'gx_axpy_norm': ([_P, _D, _P, _P, _DP], _I),
    'gx_scale': ([_P, _D, _P], _I),
    'gx_norm2': ([_P, _DP], _I),
    'gx_inner_product': ([_P, _P, _DP, _DP], _I),
    'gx_inner_product_norm': ([_P, _P, _DP, _DP, _DP], _I),
    'gx_wilson_create': ([_P, _IP, _I, _D, _DP, _DP, ctypes.POINTER(_P)], _I),
    'gx_mobius_create': ([_P, _IP, _I, _D, _D, _D, _D, _DP, _DP, ctypes.POINTER(_P)], _I),
    'gx_op_destroy': ([_P], _I),
    'gx_op_import_gauge': ([_P, _DP], _I),
    'gx_op_set_mass': ([_P, _D], _I),
    'gx_op_set_slice_masses': ([_P, _I, _DP], _I),
    'gx_op_slice_masses': ([_P, _I, _DP], _I),
    'gx_M': ([_P, _P, _P], _I),
    'gx_Mdag': ([_P, _P, _P], _I),
    'gx_Meooe': ([_P, _P, _P], _I),
    'gx_MeooeDag': ([_P, _P, _P], _I),
    'gx_Mooee': ([_P, _P, _P], _I),
    'gx_MooeeDag': ([_P, _P, _P], _I),
    'gx_MooeeInv': ([_P, _P, _P], _I),
    'gx_MooeeInvDag': ([_P, _P, _P], _I),
    'gx_Dhop': ([_P, _P, _P, _I], _I),
    'gx_DhopEO': ([_P, _P, _P, _I], _I),
    'gx_DhopOE': ([_P, _P, _P, _I], _I),
    'gx_DhopDir': ([_P, _P, _P, _I, _I], _I),
    'gx_DhopDeriv': ([_P, _DP, _P, _P, _I], _I),
    'gx_DhopDerivEO': ([_P, _DP, _P, _P, _I], _I),
    'gx_DhopDerivOE': ([_P, _DP, _P, _P, _I], _I),
    'gx_M5D': ([_P, _P, _P], _I),
    'gx_M5Ddag': ([_P, _P, _P], _I),
    'gx_Meooe5D': ([_P, _P, _P], _I),
    'gx_MeooeDag5D': ([_P, _P, _P], _I),
    'gx_import_physical_fermion_source': ([_P, _P, _P], _I),
    'gx_export_physical_fermion_solution': ([_P, _P, _P], _I),
    'gx_Mpc': ([_P, _P, _P], _I),
    'gx_MpcDag': ([_P, _P, _P], _I),
    'gx_MpcDagMpc': ([_P, _P, _P], _I),
    'gx_HermOpAndNorm': ([_P, _P, _P, _DP, _DP], _I),
    'gx_cg': ([_P, _P, _P, _D, _I, ctypes.POINTER(SolverStats)], _I),
    'gx_mixed_cg': ([_P, _P, _P, _D, _D, _I, _I, _I, ctypes.POINTER(SolverStats)], _I),
    'gx_schur_solve': ([_P, _P, _P, _D, _I, _I, ctypes.POINTER(SolverStats), _DP], _I),
    'gx_block_cg': ([_P, _P, _P, _D, _I, _I, ctypes.POINTER(SolverStats)], _I),
    'gx_block_cgrq': ([_P, _P, _P, _D, _I, ctypes.POINTER(SolverStats)], _I),
    'gx_mixed_cg_batched': ([_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_P), _D, _D, _I, _I, _I, _I,
                             ctypes.POINTER(SolverStats)], _I),
    'gx_slice_inner_product_matrix_noalloc': ([_P, _P, _DP], _I),
    'gx_inner_product_matrix': ([_I, ctypes.POINTER(_P), ctypes.POINTER(_P), _DP], _I),
    'gx_madd_matrix': ([_I, ctypes.POINTER(_P), _DP, ctypes.POINTER(_P), ctypes.POINTER(_P), _D], _I),
    'gx_block_cgrq_vec': ([_P, _I, _I, ctypes.POINTER(_P), ctypes.POINTER(_P), _D, _I, ctypes.POINTER(SolverStats)],
                          _I),
    'gx_gauge_statistics': ([_P, _IP, _DP, _DP, _DP], _I),
    'gx_nersc_read': ([_P, ctypes.c_char_p, _IP, _DP, ctypes.POINTER(NerscInfo)], _I),
    'gx_ildg_read': ([_P, ctypes.c_char_p, _IP, _DP, ctypes.POINTER(IldgInfo)], _I),
    'gx_ildg_write': ([_P, ctypes.c_char_p, _IP, _DP, _I, ctypes.c_char_p, ctypes.c_char_p], _I),
    'gx_nersc_write': ([_P, ctypes.c_char_p, _IP, _DP, _I, ctypes.c_char_p, ctypes.c_char_p, _I], _I),
    'gx_cg_multishift': ([_P, _P, _I, _DP, _DP, ctypes.POINTER(_P), _I, ctypes.POINTER(SolverStats)], _I),
    'gx_slice_inner_product_matrix': ([_P, _P, _DP], _I),
    'gx_slice_madd_matrix': ([_P, _DP, _P, _P, _D], _I),
    'gx_reliable_update_cg': ([_P, _P, _P, _D, _I, _D, _I, ctypes.POINTER(SolverStats)], _I),
    'gx_reliable_update_cg_fallback': ([_P, _P, _P, _D, _I, _D, _I, _I, _D, ctypes.POINTER(SolverStats)], _I),
    'gx_tune_set': ([_I, _I], _I),
    'gx_tune_get': ([_I, _IP], _I),
    'gx_tune_set_prec': ([_I, _I, _I], _I),
    'gx_tune_get_prec': ([_I, _I, _IP], _I),
    'gx_random_gauge': ([_IP, ctypes.c_uint64, _D, _DP], _I),
    'gx_gaussian_fermion': ([_IP, _I, ctypes.c_uint64, _DP], _I),
    'gx_point_source': ([_IP, _I, _IP, _I, _I, _DP], _I),
    'gx_wilson_clone': ([_P, _P, _I, _D, ctypes.POINTER(_P)], _I),
    'gx_op_share_links': ([_P, _P], _I),
    'gx_op_footprint': ([_P, ctypes.POINTER(ctypes.c_size_t)], _I),
    'gx_op_same_action': ([_P, _P, _IP], _I),
    'gx_field_gaussian': ([_P, ctypes.c_uint64], _I),
    'gx_field_digest': ([_P, ctypes.POINTER(ctypes.c_uint64)], _I),
    'gx_field_gaussian_slices': ([_P, ctypes.c_uint64, _I, _I], _I),
    'gx_field_digest_slices': ([_P, _I, _I, ctypes.POINTER(ctypes.c_uint64)], _I),
    'gx_field_point_source': ([_P, _IP, _I, _I], _I),
    'gx_extract_slice': ([_P, _P, _I], _I),
    'gx_insert_slice': ([_P, _P, _I], _I),
    'gx_serve_footprint': ([_IP, _I, _I, _DP], _I),
    'gx_serve': ([ctypes.POINTER(ServeParams), ctypes.POINTER(Request), _I, ctypes.POINTER(Response),
                  ctypes.POINTER(_DP), _DP, _DP], _I),
}


def lib():
    """Load libgx.so (built in-tree by __graft_entry__.build()); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'{LIB_PATH} is missing: build it first (python -c "import __graft_entry__ as g; g.build()")')
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIG.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIG.keys())


def _check(rc):
    if rc != 0:
        raise GxError(rc, lib().gx_last_error().decode())


def _dims(d):
    return (ctypes.c_int * 4)(*[int(x) for x in d])


def _dp(a):
    assert a.flags['C_CONTIGUOUS'] and a.dtype in (np.complex128, np.float64)
    return a.ctypes.data_as(_DP)


class Context:
    def __init__(self, device=0):
        h = _P()
        _check(lib().gx_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def sync(self):
        _check(lib().gx_ctx_sync(self.h))

    def record(self, slot):
        _check(lib().gx_ctx_event_record(self.h, slot))

    def elapsed_ms(self, s0, s1):
        v = _D()
        _check(lib().gx_ctx_event_elapsed(self.h, s0, s1, ctypes.byref(v)))
        return v.value

    def close(self):
        if self.h:
            lib().gx_ctx_destroy(self.h)
            self.h = None


class Field:
    def __init__(self, ctx, dims, Ls=1, prec=F64, subset=FULL):
        self.ctx, self.dims, self.Ls, self.prec, self.subset = ctx, tuple(int(d) for d in dims), int(Ls), prec, subset
        h = _P()
        _check(lib().gx_field_create(ctx.h, _dims(dims), self.Ls, prec, subset, ctypes.byref(h)))
        self.h = h

    @property
    def shape(self):
        lx, ly, lz, lt = self.dims
        x = lx if self.subset == FULL else lx // 2
        return (lt, lz, ly, x, self.Ls, 4, 3)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == self.shape, (a.shape, self.shape)
        _check(lib().gx_field_import(self.h, _dp(a)))
        return self

    def download(self):
        out = np.zeros(self.shape, dtype=np.complex128)
        _check(lib().gx_field_export(self.h, _dp(out)))
        return out

    @property
    def checkerboard(self):
        v = _I()
        _check(lib().gx_field_checkerboard(self.h, ctypes.byref(v)))
        return v.value

    def set_checkerboard_tag(self, cb):
        _check(lib().gx_field_set_checkerboard_tag(self.h, cb))

    def zero(self):
        _check(lib().gx_field_zero(self.h))

    def norm2(self):
        v = _D()
        _check(lib().gx_norm2(self.h, ctypes.byref(v)))
        return v.value

    def gaussian(self, seed):
        """Fill a full field on the device with gaussian_fermion(dims, Ls, seed)."""
        _check(lib().gx_field_gaussian(self.h, ctypes.c_uint64(seed)))
        return self

    def point_source(self, x, spin=-1, colour=-1):
        """Device point source at site x = (x, y, z, t): spin = colour = -1 puts spin-colour
        component s in slice s (the 12 point sources of a site as one multi-RHS field)."""
        _check(lib().gx_field_point_source(self.h, _dims(x), int(spin), int(colour)))
        return self

    def digest(self):
        v = ctypes.c_uint64()
        _check(lib().gx_field_digest(self.h, ctypes.byref(v)))
        return v.value

    def gaussian_slices(self, seed, s0, ns):
        """Slices s0 .. s0+ns-1 get the ns-slice field gaussian(seed) would give."""
        _check(lib().gx_field_gaussian_slices(self.h, ctypes.c_uint64(seed), int(s0), int(ns)))
        return self

    def digest_slices(self, s0, ns):
        """Digest of slices s0 .. s0+ns-1 by logical position (full F64 fields)."""
        v = ctypes.c_uint64()
        _check(lib().gx_field_digest_slices(self.h, int(s0), int(ns), ctypes.byref(v)))
        return v.value

    def close(self):
        if self.h:
            lib().gx_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def inner_product(x, y):
    re, im = _D(), _D()
    _check(lib().gx_inner_product(x.h, y.h, ctypes.byref(re), ctypes.byref(im)))
    return complex(re.value, im.value)


def axpy_norm(z, a, x, y):
    v = _D()
    _check(lib().gx_axpy_norm(z.h, a, x.h, y.h, ctypes.byref(v)))
    return v.value


def axpby(z, a, b, x, y):
    _check(lib().gx_axpby(z.h, a, b, x.h, y.h))


def extract_slice(out4d, in5d, s):
    _check(lib().gx_extract_slice(out4d.h, in5d.h, int(s)))


def insert_slice(in4d, out5d, s):
    _check(lib().gx_insert_slice(in4d.h, out5d.h, int(s)))


def field_copy(dst, src):
    _check(lib().gx_field_copy(dst.h, src.h))


(TUNE_DHOP_MAP, TUNE_DHOP_PF, TUNE_DHOP_NT, TUNE_DHOP_ORDER, TUNE_CG_GRAPH, TUNE_FUSED, TUNE_DHOP_SHAPE,
 TUNE_DHOP_PERSIST, TUNE_DHOP_MFMA, TUNE_DHOP_TILE) = 0, 1, 2, 3, 4, 5, 6, 7, 8, 9


def tune_set(knob, value, prec=None):
    """Set a launch knob for every precision, or for one storage precision (F64, F32, F16)."""
    if prec is None:
        _check(lib().gx_tune_set(int(knob), int(value)))
    else:
        _check(lib().gx_tune_set_prec(int(knob), int(prec), int(value)))


def tune_get(knob, prec=None):
    """A launch knob's value (of the FP64 set, or of precision prec)."""
    v = _I()
    if prec is None:
        _check(lib().gx_tune_get(int(knob), ctypes.byref(v)))
    else:
        _check(lib().gx_tune_get_prec(int(knob), int(prec), ctypes.byref(v)))
    return v.value


def precision_change(out, inp):
    _check(lib().gx_precision_change(out.h, inp.h))


def gauge_statistics(ctx, U, dims):
    """(avgPlaquette, linkTrace) of a host gauge field, computed on the device."""
    U = np.ascontiguousarray(U, dtype=np.complex128)
    p, t = _D(), _D()
    _check(lib().gx_gauge_statistics(ctx.h, _dims(dims), _dp(U), ctypes.byref(p), ctypes.byref(t)))
    return p.value, t.value


def nersc_read(ctx, path, dims, check=True):
    """NerscIO::readConfiguration: returns (U, header info).  With check=False a failed checksum,
    plaquette or link-trace comparison returns (U, info) with info['ok'] = False instead of raising."""
    lx, ly, lz, lt = dims
    U = np.zeros((lt, lz, ly, lx, 4, 3, 3), dtype=np.complex128)
    info = NerscInfo()
    rc = lib().gx_nersc_read(ctx.h, os.fsencode(path), _dims(dims), _dp(U), ctypes.byref(info))
    if check:
        _check(rc)
    d = info.as_dict()
    d['ok'] = rc == 0
    return U, d


def nersc_write(ctx, path, U, dims, two_row=False, ensemble_label='DWF', ensemble_id='UKQCD', sequence_number=1):
    """NerscIO::writeConfiguration (IEEE64BIG)."""
    U = np.ascontiguousarray(U, dtype=np.complex128)
    _check(lib().gx_nersc_write(ctx.h, os.fsencode(path), _dims(dims), _dp(U), int(two_row),
                                ensemble_label.encode(), ensemble_id.encode(), sequence_number))


def ildg_read(ctx, path, dims, check=True):
    """IldgReader::readConfiguration: returns (U, info).  With check=False a failed check returns
    (U, info) with info['ok'] = False instead of raising."""
    lx, ly, lz, lt = dims
    U = np.zeros((lt, lz, ly, lx, 4, 3, 3), dtype=np.complex128)
    info = IldgInfo()
    rc = lib().gx_ildg_read(ctx.h, os.fsencode(path), _dims(dims), _dp(U), ctypes.byref(info))
    if check:
        _check(rc)
    d = info.as_dict()
    d['ok'] = rc == 0
    return U, d


def ildg_write(ctx, path, U, dims, precision=64, lfn='', description=''):
    """IldgWriter::writeConfiguration with big-endian words of `precision` bits."""
    U = np.ascontiguousarray(U, dtype=np.complex128)
    _check(lib().gx_ildg_write(ctx.h, os.fsencode(path), _dims(dims), _dp(U), int(precision), lfn.encode(),
                               description.encode()))


def slice_inner_product_matrix(lhs, rhs):
    """sliceInnerProductMatrix (Orthog = the fifth dimension): m[s, t] = <lhs_s, rhs_t>."""
    n = lhs.Ls
    m = np.zeros((n, n), dtype=np.complex128)
    _check(lib().gx_slice_inner_product_matrix(lhs.h, rhs.h, _dp(m)))
    return m


def slice_madd_matrix(r, aa, x, y=None, scale=1.0):
    """sliceMaddMatrix: r_t = y_t + scale sum_s x_s aa[s, t] (y = None: sliceMulMatrix)."""
    a = np.ascontiguousarray(aa, dtype=np.complex128)
    assert a.shape == (x.Ls, x.Ls), a.shape
    _check(lib().gx_slice_madd_matrix(r.h, _dp(a), x.h, y.h if y is not None else None, scale))


def inner_product_matrix(xs, ys):
    """InnerProductMatrix over lists of fields: m[b, bp] = <xs[b], ys[bp]>."""
    n = len(xs)
    m = np.zeros((n, n), dtype=np.complex128)
    _check(lib().gx_inner_product_matrix(n, (_P * n)(*[f.h for f in xs]), (_P * n)(*[f.h for f in ys]), _dp(m)))
    return m


def madd_matrix(rs, aa, xs, ys=None, scale=1.0):
    """MaddMatrix over lists of fields: rs[b] = ys[b] + scale sum_bp xs[bp] aa[bp, b] (ys None: MulMatrix)."""
    n = len(xs)
    a = np.ascontiguousarray(aa, dtype=np.complex128)
    yp = (_P * n)(*[f.h for f in ys]) if ys is not None else None
    _check(lib().gx_madd_matrix(n, (_P * n)(*[f.h for f in rs]), _dp(a), (_P * n)(*[f.h for f in xs]), yp, scale))


def pick_checkerboard(cb, half, full):
    _check(lib().gx_pick_checkerboard(cb, half.h, full.h))


def set_checkerboard(full, half):
    _check(lib().gx_set_checkerboard(full.h, half.h))


class FermionOp:
    """WilsonFermion (kind='wilson'; Ls>1 = multi-RHS) or MobiusFermion (kind='mobius')."""

    def __init__(self, ctx, U, dims, Ls=1, kind='wilson', mass=0.1, M5=1.8, b=1.5, c=0.5, phases=None):
        self.ctx, self.dims, self.Ls, self.kind = ctx, tuple(int(d) for d in dims), int(Ls), kind
        U = np.ascontiguousarray(U, dtype=np.complex128)
        ph = None
        if phases is not None:
            ph = np.array([[complex(p).real, complex(p).imag] for p in phases], dtype=np.float64).reshape(-1)
        h = _P()
        phptr = _dp(ph) if ph is not None else None
        if kind == 'wilson':
            _check(lib().gx_wilson_create(ctx.h, _dims(dims), self.Ls, mass, _dp(U), phptr, ctypes.byref(h)))
        else:
            _check(lib().gx_mobius_create(ctx.h, _dims(dims), self.Ls, mass, M5, b, c, _dp(U), phptr, ctypes.byref(h)))
        self.h = h

    def field(self, prec=F64, subset=FULL):
        return Field(self.ctx, self.dims, self.Ls, prec, subset)

    def set_mass(self, mass):
        self._call('op_set_mass', ctypes.c_double(mass))

    def set_slice_masses(self, masses):
        """Wilson multi-RHS operator: mass masses[s] on slice s (None: back to the single mass)."""
        if masses is None:
            self._call('op_set_slice_masses', 0, None)
            return
        m = np.ascontiguousarray(masses, dtype=np.float64)
        self._call('op_set_slice_masses', len(m), _dp(m))

    def slice_masses(self):
        m = np.zeros(self.Ls)
        self._call('op_slice_masses', self.Ls, _dp(m))
        return m

    def clone(self, ctx, Ls, mass):
        """Wilson operator reading this operator's resident links, with its own context, Ls and mass."""
        c = FermionOp.__new__(FermionOp)
        c.ctx, c.dims, c.Ls, c.kind = ctx, self.dims, int(Ls), 'wilson'
        h = _P()
        _check(lib().gx_wilson_clone(self.h, ctx.h, int(Ls), mass, ctypes.byref(h)))
        c.h = h
        c._base = self  # the links must outlive the clone
        return c

    def share_links(self, base):
        self._call('op_share_links', base.h)
        self._base = base

    def footprint(self):
        v = ctypes.c_size_t()
        self._call('op_footprint', ctypes.byref(v))
        return v.value

    def same_action(self, other):
        v = ctypes.c_int()
        self._call('op_same_action', other.h, ctypes.byref(v))
        return bool(v.value)

    def import_gauge(self, U):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        self._call('op_import_gauge', _dp(U))

    def _call(self, name, *args):
        _check(getattr(lib(), 'gx_' + name)(self.h, *args))

    def M(self, i, o): self._call('M', i.h, o.h)
    def Mdag(self, i, o): self._call('Mdag', i.h, o.h)
    def Meooe(self, i, o): self._call('Meooe', i.h, o.h)
    def MeooeDag(self, i, o): self._call('MeooeDag', i.h, o.h)
    def Mooee(self, i, o): self._call('Mooee', i.h, o.h)
    def MooeeDag(self, i, o): self._call('MooeeDag', i.h, o.h)
    def MooeeInv(self, i, o): self._call('MooeeInv', i.h, o.h)
    def MooeeInvDag(self, i, o): self._call('MooeeInvDag', i.h, o.h)
    def Dhop(self, i, o, dag=0): self._call('Dhop', i.h, o.h, int(dag))
    def DhopEO(self, i, o, dag=0): self._call('DhopEO', i.h, o.h, int(dag))
    def DhopOE(self, i, o, dag=0): self._call('DhopOE', i.h, o.h, int(dag))
    def DhopDir(self, i, o, dir, disp): self._call('DhopDir', i.h, o.h, int(dir), int(disp))

    def DhopDeriv(self, A, B, dag=0, which=''):
        """DhopDeriv / DhopDerivEO / DhopDerivOE (which = '', 'EO', 'OE'): the force matrices
        mat[t, z, y, x, mu, a, b] (zeros on the sites the EO/OE forms do not fill)."""
        lx, ly, lz, lt = self.dims
        mat = np.zeros((lt, lz, ly, lx, 4, 3, 3), dtype=np.complex128)
        self._call('DhopDeriv' + which, _dp(mat), A.h, B.h, int(dag))
        return mat
    def M5D(self, i, o): self._call('M5D', i.h, o.h)
    def M5Ddag(self, i, o): self._call('M5Ddag', i.h, o.h)
    def Meooe5D(self, i, o): self._call('Meooe5D', i.h, o.h)
    def MeooeDag5D(self, i, o): self._call('MeooeDag5D', i.h, o.h)
    def ImportPhysicalFermionSource(self, in4d, out5d): self._call('import_physical_fermion_source', in4d.h, out5d.h)
    def ExportPhysicalFermionSolution(self, sol5d, out4d): self._call('export_physical_fermion_solution', sol5d.h, out4d.h)
    def Mpc(self, i, o): self._call('Mpc', i.h, o.h)
    def MpcDag(self, i, o): self._call('MpcDag', i.h, o.h)
    def MpcDagMpc(self, i, o): self._call('MpcDagMpc', i.h, o.h)

    def cg(self, src, psi, tol=1e-10, maxit=10000):
        st = SolverStats()
        rc = lib().gx_cg(self.h, src.h, psi.h, tol, maxit, ctypes.byref(st))
        if rc not in (0, 3):
            _check(rc)
        return st.as_dict()

    def mixed_cg(self, src, psi, tol=1e-10, inner_tol=1e-5, inner_prec=F32, maxinner=10000, maxouter=100):
        st = SolverStats()
        rc = lib().gx_mixed_cg(self.h, src.h, psi.h, tol, inner_tol, inner_prec, maxinner, maxouter, ctypes.byref(st))
        if rc not in (0, 3):
            _check(rc)
        return st.as_dict()

    def reliable_update_cg(self, src, psi, tol=1e-10, maxit=10000, delta=0.1, sloppy_prec=F32):
        """ConjugateGradientReliableUpdate: sloppy-precision CG with FP64 reliable updates."""
        st = SolverStats()
        rc = lib().gx_reliable_update_cg(self.h, src.h, psi.h, tol, maxit, delta, sloppy_prec, ctypes.byref(st))
        if rc not in (0, 3):
            _check(rc)
        return st.as_dict()

    def reliable_update_cg_fallback(self, src, psi, tol=1e-10, maxit=10000, delta=0.1, sloppy_prec=F16,
                                    fallback_prec=F32, fallback_transition_tol=0.0):
        """ConjugateGradientReliableUpdate with setFallbackLinop: FP16 sloppy vectors switch to FP32 once
        the sloppy |r|^2 drops below fallback_transition_tol (absolute)."""
        st = SolverStats()
        rc = lib().gx_reliable_update_cg_fallback(self.h, src.h, psi.h, tol, maxit, delta, sloppy_prec, fallback_prec,
                                                  fallback_transition_tol, ctypes.byref(st))
        if rc not in (0, 3):
            _check(rc)
        return st.as_dict()

    def block_cg(self, src, psi, tol=1e-10, maxit=10000, inner_prec=F64):
        """Multi-RHS CG (right-hand side k = slice s = k); returns one stats dict per RHS."""
        st = (SolverStats * self.Ls)()
        rc = lib().gx_block_cg(self.h, src.h, psi.h, tol, maxit, inner_prec, st)
        if rc not in (0, 3):
            _check(rc)
        return [s.as_dict() for s in st]

    def cg_multishift(self, src, psis, shifts, tolerances, maxit=10000):
        """ConjugateGradientMultiShift: (MpcDagMpc + shifts[s]) psis[s] = src for every shift at once;
        shifts[0] must be the smallest.  Returns one stats dict per shift."""
        n = len(shifts)
        assert len(psis) == n and len(tolerances) == n
        sh = np.ascontiguousarray(shifts, dtype=np.float64)
        tl = np.ascontiguousarray(tolerances, dtype=np.float64)
        hs = (_P * n)(*[f.h for f in psis])
        st = (SolverStats * n)()
        rc = lib().gx_cg_multishift(self.h, src.h, n, _dp(sh), _dp(tl), hs, maxit, st)
        if rc not in (0, 3):
            _check(rc)
        return [s.as_dict() for s in st]

    def block_cgrq(self, src, psi, tol=1e-10, maxit=10000):
        """BlockConjugateGradient(BlockCGrQ): one block Krylov space for the Ls right-hand sides;
        returns one stats dict per RHS."""
        st = (SolverStats * self.Ls)()
        rc = lib().gx_block_cgrq(self.h, src.h, psi.h, tol, maxit, st)
        if rc not in (0, 3):
            _check(rc)
        return [s.as_dict() for s in st]

    def mixed_cg_batched(self, srcs, sols, tol=1e-10, inner_tol=1e-5, inner_prec=F32, maxinner=10000, maxouter=100,
                         maxpatch=10000):
        """MixedPrecisionConjugateGradientBatched over a list of half fields; one stats dict per source."""
        n = len(srcs)
        assert len(sols) == n
        hs = (_P * n)(*[f.h for f in srcs])
        hx = (_P * n)(*[f.h for f in sols])
        st = (SolverStats * n)()
        rc = lib().gx_mixed_cg_batched(self.h, n, hs, hx, tol, inner_tol, int(inner_prec), maxinner, maxouter, maxpatch,
                                       st)
        if rc not in (0, 3):
            _check(rc)
        return [s.as_dict() for s in st]

    def block_cgrq_vec(self, srcs, psis, tol=1e-10, maxit=10000, herm=None):
        """BlockConjugateGradient(BlockCGrQVec) over a list of fields: MpcDagMpc on half fields
        (herm HERM_SCHUR, the default for half fields) or Mdag M on full fields (HERM_MDAGM); returns
        one stats dict per field."""
        n = len(srcs)
        assert len(psis) == n
        if herm is None:
            herm = HERM_MDAGM if srcs[0].subset == FULL else HERM_SCHUR
        hs = (_P * n)(*[f.h for f in srcs])
        hp = (_P * n)(*[f.h for f in psis])
        st = (SolverStats * n)()
        rc = lib().gx_block_cgrq_vec(self.h, int(herm), n, hs, hp, tol, maxit, st)
        if rc not in (0, 3):
            _check(rc)
        return [s.as_dict() for s in st]

    def schur_solve(self, src, sol, tol=1e-10, maxit=10000, mixed=False):
        st = SolverStats()
        ur = _D()
        rc = lib().gx_schur_solve(self.h, src.h, sol.h, tol, maxit, int(mixed), ctypes.byref(st), ctypes.byref(ur))
        if rc not in (0, 3):
            _check(rc)
        d = st.as_dict()
        d['unprec_residual'] = ur.value
        return d

    def close(self):
        """Releases the operator.  A base whose links live clones still read raises (GX_EINVAL) and
        stays alive: close its clones first."""
        if self.h:
            _check(lib().gx_op_destroy(self.h))
            self.h = None


def random_gauge(dims, seed, eps):
    lx, ly, lz, lt = dims
    U = np.zeros((lt, lz, ly, lx, 4, 3, 3), dtype=np.complex128)
    _check(lib().gx_random_gauge(_dims(dims), seed, eps, _dp(U)))
    return U


def gaussian_fermion(dims, Ls, seed):
    lx, ly, lz, lt = dims
    psi = np.zeros((lt, lz, ly, lx, Ls, 4, 3), dtype=np.complex128)
    _check(lib().gx_gaussian_fermion(_dims(dims), int(Ls), seed, _dp(psi)))
    return psi


def point_source(dims, Ls, x, spin, colour):
    lx, ly, lz, lt = dims
    psi = np.zeros((lt, lz, ly, lx, Ls, 4, 3), dtype=np.complex128)
    _check(lib().gx_point_source(_dims(dims), int(Ls), _dims(x), spin, colour, _dp(psi)))
    return psi
