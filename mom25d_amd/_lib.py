"""ctypes binding of libm25x.so (include/m25x.h). Loading fails loudly: there is no CPU fallback."""
import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libm25x.so')

C128, C64 = 0, 1
EINVAL, EHIP, ENOMEM, ESINGULAR = -1, -2, -3, -4
PIVOT_PARTIAL, PIVOT_TOURNAMENT = 0, 1
GETRS_SCRATCH_BYTES = 16 << 20


class Tables(C.Structure):
    _fields_ = [('ntab', C.c_int), ('nfine', C.c_int), ('L', C.c_int), ('AQ', C.c_void_p),
                ('dr', C.c_void_p), ('d', C.c_void_p), ('dimg', C.c_void_p),
                ('cA', C.c_void_p), ('cq', C.c_void_p), ('cAi', C.c_void_p), ('cqi', C.c_void_p),
                ('tab', C.c_void_p), ('kind', C.c_void_p)]


MAX_LAYERS = 16
SPEC_TABLE, SPEC_VIA_D, SPEC_VIA_R = 0, 1, 2
SPEC_NI, SPEC_ND = 8, 10


class Stack(C.Structure):
    _fields_ = [('N', C.c_int), ('bottom_pec', C.c_int), ('top_pec', C.c_int), ('pad', C.c_int),
                ('eps', (C.c_double * 2) * MAX_LAYERS), ('h', C.c_double * MAX_LAYERS),
                ('eps_bot', C.c_double * 2), ('eps_top', C.c_double * 2),
                ('k0', C.c_double), ('wmu', C.c_double), ('weps', C.c_double)]


class LuItem(C.Structure):
    _fields_ = [('A', C.c_void_p), ('n', C.c_long), ('lda', C.c_long), ('ipiv', C.c_void_p), ('perm', C.c_void_p),
                ('info', C.c_void_p), ('work', C.c_void_p), ('work_bytes', C.c_size_t)]


class RsItem(C.Structure):
    _fields_ = [('LU', C.c_void_p), ('n', C.c_long), ('lda', C.c_long), ('perm', C.c_void_p), ('B', C.c_void_p),
                ('ldb', C.c_long), ('nrhs', C.c_int), ('pad', C.c_int), ('work', C.c_void_p),
                ('work_bytes', C.c_size_t)]


class MvItem(C.Structure):
    _fields_ = [('Z', C.c_void_p), ('n', C.c_long), ('ld', C.c_long), ('X', C.c_void_p), ('Y', C.c_void_p),
                ('ldx', C.c_long), ('ldy', C.c_long), ('k', C.c_int), ('pad', C.c_int)]


class Mesh(C.Structure):
    _fields_ = [('Ne', C.c_int), ('Nt', C.c_int), ('ep', C.c_void_p), ('em', C.c_void_p), ('lt', C.c_void_p),
                ('te', C.c_void_p), ('ap', C.c_void_p), ('am', C.c_void_p), ('vp', C.c_void_p), ('vm', C.c_void_p),
                ('cent', C.c_void_p), ('area', C.c_void_p), ('ta', C.c_void_p), ('tv', C.c_void_p)]


class Quad(C.Structure):
    _fields_ = [('nodes', C.c_void_p), ('tris', C.c_void_p), ('P3', C.c_void_p), ('W3', C.c_void_p),
                ('P6', C.c_void_p), ('W6', C.c_void_p)]


P, I, L, D, S = C.c_void_p, C.c_int, C.c_long, C.c_double, C.c_size_t
_SIGS = {
    'm25x_version': ([], I),
    'm25x_init': ([I], I),
    'm25x_fill_far': ([P, I, L, P, P, D, P], I),
    'm25x_far_block': ([P, I, P, I, P, P, D, P, P], I),
    'm25x_near_moments': ([P, P, P, P, P, L, D, P, I, P, P, P], I),
    'm25x_near_scatter': ([P, I, L, P, P, P, P, L, P, P, P, P, D, P], I),
    'm25x_near_coo': ([P, P, P, P, L, D, P, P, P, P], I),
    'm25x_near_pairs': ([P, P, P, I, D, P, C.POINTER(C.c_long), P, P, P], I),
    'm25x_matvec_far_work_bytes': ([I, I], S),
    'm25x_matvec_far': ([P, I, L, P, P, D, P, L, P, S, P], I),
    'm25x_near_values': ([P, P, P, P, L, P, P, D, P, P], I),
    'm25x_csr_matvec': ([P, P, P, L, P, I, L, P, L, I, P], I),
    'm25x_fan_q': ([P, P, P, I, P, P], I),
    'm25x_radial_table_sum': ([P, P, P, L, L, P, P, L, P, L, D, P, P], I),
    'm25x_hankel': ([P, I, P, I, P, I, P, P], I),
    'm25x_bessel_jn_complex': ([I, P, L, P, P], I),
    'm25x_spectra': ([P, D, P, P, I, P, P, I, P, P, L, P], I),
    'm25x_spectral_point': ([P, D, D, D, I, I, P], I),
    'm25x_scale_c64': ([P, L, L, P, P, L, P], I),
    'm25x_diag_scales': ([P, L, L, P, P], I),
    'm25x_getrf_work_bytes': ([I, L], S),
    'm25x_getrf': ([I, I, P, L, L, P, P, P, P, S, P], I),
    'm25x_ipiv_perm': ([P, L, P, P], I),
    'm25x_getrs_work_bytes': ([I, L, L], S),
    'm25x_getrs': ([I, P, L, L, P, P, P, L, I, I, P, S, P], I),
    'm25x_getrs_noalloc': ([I, P, L, L, P, P, P, L, I, I, P, P], I),
    'm25x_lu_set_gemm_stage': ([I], I),
    'm25x_lu_gemm_stage': ([], I),
    'm25x_lu_set_update_precision': ([I], I),
    'm25x_lu_set_lookahead': ([I], I),
    'm25x_lu_lookahead': ([], I),
    'm25x_getrf_batched': ([I, I, P, P, P], I),
    'm25x_getrs_batched': ([I, I, P, P, I, P, P], I),
    'm25x_zgemm_nk_batched': ([I, P, P, P], I),
    'm25x_timing': ([C.c_uint], I),
    'm25x_timing_records': ([P, P, P, P, P, L], L),
    'm25x_kernel_name': ([I], C.c_char_p),
    'm25x_zgemm_nk': ([P, L, L, P, I, P, P], I),
    'm25x_cgs2_work_bytes': ([], S),
    'm25x_cgs2': ([P, L, I, I, P, P, P, P, S, P], I),
}

_lib = None
_lock = threading.Lock()
_inited = set()


class M25xError(RuntimeError):
    pass


def lib():
    """The loaded library (RuntimeError when it is not built: `python -c 'import __graft_entry__ as g; g.build()'`
    or `make -C mom25d_amd/csrc`)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(f"mom25d_amd: {LIB_PATH} is not built (make -C mom25d_amd/csrc); the HIP "
                                       f"kernels are the only implementation")
                h = C.CDLL(LIB_PATH)
                h.m25x_last_error.restype = C.c_char_p
                h.m25x_last_error.argtypes = []
                for name, (args, res) in _SIGS.items():
                    f = getattr(h, name)
                    f.argtypes, f.restype = args, res
                _lib = h
    return _lib


def check(rc, what=''):
    """Raises the reference's exception types for the C ABI's codes."""
    if rc == 0:
        return
    msg = lib().m25x_last_error().decode() or what
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise M25xError(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def init(device_index):
    """m25x_init once per device (checks that it is a gfx950)."""
    if device_index not in _inited:
        call('m25x_init', device_index)
        _inited.add(device_index)
