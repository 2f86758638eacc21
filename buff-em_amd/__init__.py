"""Python host bindings of the MI355X dense VIE hot path.

Thin ctypes layer over the C ABI in include/buffem_amd.h (libbuffem_amd.so,
built by buff-em_amd/build.py).  It mirrors the reference's scattering API
(HomerReid/buff-em, src/libs/libbuff/libbuff.h:200-283):

    SWGGeometry          -> Geometry (mesh, SWG topology, device upload)
    AssembleVIEMatrix    -> Geometry.assemble_vie
    AssembleRHSVector    -> Geometry.assemble_rhs
    HMatrix::LUFactorize -> lu_factor
    HMatrix::LUSolve     -> lu_solve
    GetPFTMatrix (EMT)   -> Geometry.emt_power

Device buffers are passed as raw pointers; torch tensors are accepted for
convenience (complex128, contiguous).  A dense N x N matrix in the library's
column-major layout with ld = N is a torch tensor T of shape (N, N) with
T[c, r] = M(r, c), i.e. M = T.T.

There is no CPU fallback: every compute call goes to the HIP library and
raises if it is missing or fails.
"""
import atexit
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbuffem_amd.so")
_lib = None


class BFAError(RuntimeError):
    pass


class bfa_material(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("p", ctypes.c_double * 6), ("expr", ctypes.c_void_p)]


class bfa_planewave(ctypes.Structure):
    _fields_ = [("E0", ctypes.c_double * 6), ("dir", ctypes.c_double * 3)]


P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
D = ctypes.c_double

_SIGS = {
    "bfa_last_error": (ctypes.c_char_p, []),
    "bfa_status_string": (ctypes.c_char_p, [I]),
    "bfa_version": (I, []),
    "bfa_mesh_kuhn_sphere": (I, [I, D, D, ctypes.c_uint64, P, P, P, P]),
    "bfa_mesh_read_vmsh": (I, [ctypes.c_char_p, P, P, P, P]),
    "bfa_geometry_create": (I, [I, P, I, P, P]),
    "bfa_geometry_destroy": (None, [P]),
    "bfa_geometry_info": (I, [P, P, P, P, P]),
    "bfa_geometry_faces": (I, [P, P, P, P, P, P, P, P, P]),
    "bfa_geometry_touch_count": (I, [P, P]),
    "bfa_geometry_touch_pairs": (I, [P, P, P]),
    "bfa_geometry_upload": (I, [P, P]),
    "bfa_static_build": (I, [P, I, P]),
    "bfa_static_download": (I, [P, P]),
    "bfa_static_upload": (I, [P, P]),
    "bfa_get_fields": (I, [P, D, P, I, P, I, P, P, P]),
    "bfa_assemble_g": (I, [P, D, D, P, I64, P]),
    "bfa_overlap_blocks": (I, [P, D, P, P, D, P, P, P, P, P]),
    "bfa_gram_inverse": (I, [P, P, I64, P]),
    "bfa_dressed_rytov_work_bytes": (ctypes.c_size_t, [I, I64]),
    "bfa_dressed_rytov": (I, [P, D, P, P, D, P, P, I64, P, P]),
    "bfa_emt_power_dr": (I, [P, D, P, I64, P, P]),
    "bfa_material_svtensor": (I, [ctypes.c_char_p, I, P]),
    "bfa_material_expr": (I, [ctypes.c_char_p, P]),
    "bfa_material_release": (I, [P]),
    "bfa_material_check": (I, [ctypes.c_char_p, I]),
    "bfa_material_expr_check": (I, [ctypes.c_char_p]),
    "bfa_material_eval": (I, [P, D, I, P, P]),
    "bfa_buffgeo_load": (I, [ctypes.c_char_p, P, P, P, P, P]),
    "bfa_fibbi_keys": (I, [P, P]),
    "bfa_fibbi_read": (I, [ctypes.c_char_p, P, P, P]),
    "bfa_fibbi_write": (I, [ctypes.c_char_p, I64, P, P]),
    "bfa_fibbi_store": (I, [P, ctypes.c_char_p, P]),
    "bfa_fibbi_preload": (I, [P, ctypes.c_char_p, I, P, P, P]),
    "bfa_assemble_vie": (I, [P, D, D, P, P, I64, P]),
    "bfa_assemble_rhs": (I, [P, D, P, I, P, I64, P]),
    "bfa_assemble_vie_far32": (I, [P, D, D, P, P, I64, P]),
    "bfa_lu_factor_z": (I, [P, I, I64, P, P, P]),
    "bfa_lu_solve_z": (I, [P, I, I64, P, P, I, I64, P]),
    "bfa_lu_solve_z_noalloc": (I, [P, I, I64, P, P, I, I64, P]),
    "bfa_lu_factor_c_local": (I, [P, I, I64, P, P, P]),
    "bfa_shutdown": (I, []),
    "bfa_cgemm_sub_split": (I, [I, I, I, P, I64, P, I64, P, I64, P]),
    "bfa_emt_power": (I, [P, D, P, I, P, I64, P, P]),
    "bfa_zgemm_sub": (I, [I, I, I, P, I64, P, I64, P, I64, P]),
    "bfa_cgemm_sub": (I, [I, I, I, P, I64, P, I64, P, I64, P]),
    "bfa_convert_z2c": (I, [P, I, I, I64, P, I64, P]),
    "bfa_lu_factor_c": (I, [P, I, I64, P, P, P]),
    "bfa_lu_solve_c": (I, [P, I, I64, P, P, I, I64, P]),
    "bfa_solve_refine": (I, [P, I64, P, I64, P, I, P, I, I64, D, I, P, P, P]),
    "bfa_profile_enable": (I, [I]),
    "bfa_profile_read": (I, [I, P, P, P]),
    "bfa_lu_partition": (I, [I]),
    "bfa_device_malloc": (I, [P, ctypes.c_size_t]),
    "bfa_device_free": (I, [P]),
    "bfa_memcpy_htod": (I, [P, P, ctypes.c_size_t, P]),
    "bfa_memcpy_dtoh": (I, [P, P, ctypes.c_size_t, P]),
    "bfa_stream_synchronize": (I, [P]),
    "bfa_batch_create": (I, [I, P, P, P, I, P]),
    "bfa_batch_destroy": (I, [P]),
    "bfa_batch_info": (I, [P, P, P, P, P]),
    "bfa_batch_tune": (I, [P, I]),
    "bfa_batch_tuning": (I, [P, P, P]),
    "bfa_batch_set_tiles": (I, [P, P]),
    "bfa_batch_run": (I, [P, P, I, I, D, I, P, P, P, P]),
    "bfa_batch_timing": (I, [P, P]),
    "bfa_batch_solution": (I, [P, I, P]),
    "bfa_batch_matrix": (I, [P, I, P]),
    "bfa_mlp_param_count": (I, [I, I, I, P, P]),
    "bfa_mlp_grad": (I, [I, I, I, I, P, I, P, P, P, P, P]),
    "bfa_surr_create": (I, [P, D, I, I, ctypes.c_uint64, P]),
    "bfa_surr_destroy": (I, [P]),
    "bfa_surr_pairs": (I, [P, P, P, P]),
    "bfa_surr_cubature": (I, [P, P, P, I, I, I, P]),
    "bfa_surr_features": (I, [P, P, P, I, P]),
    "bfa_surr_dataset": (I, [P, I, D, D, I, I]),
    "bfa_surr_train": (I, [P, I, I, D, P]),
    "bfa_surr_table": (I, [P, D, I, D, I, I, P, P]),
    "bfa_surr_table_download": (I, [P, P]),
    "bfa_surr_table_upload": (I, [P, P]),
    "bfa_assemble_vie_table": (I, [P, D, P, P, I64, P]),
    "bfa_surr_info": (I, [P, P, P, P, P]),
    "bfa_surr_params": (I, [P, P, P]),
}


def lib():
    """Load libbuffem_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BFAError(f"{LIB_PATH} missing: run python buff-em_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        # release the library's device objects before the GPU runtime's own
        # teardown (atexit handlers run last-registered first)
        atexit.register(L.bfa_shutdown)
    return _lib


def declared_symbols():
    return sorted(_SIGS)


def _check(rc, what):
    if rc != 0:
        msg = lib().bfa_last_error().decode(errors="replace")
        raise BFAError(f"{what} failed ({rc}): {msg}")


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    raise TypeError(type(x))


def _stream_ptr(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return int(stream)


def material(kind="iso", **kw):
    """Material descriptor: iso(eps), diag(eps=(ex,ey,ez)), sincos(base, amp, imag), janus(up, down)."""
    m = bfa_material()
    p = [0.0] * 6
    if kind == "iso":
        e = complex(kw["eps"])
        m.kind, p[0], p[1] = 0, e.real, e.imag
    elif kind == "diag":
        es = [complex(x) for x in kw["eps"]]
        m.kind = 1
        p = [es[0].real, es[0].imag, es[1].real, es[1].imag, es[2].real, es[2].imag]
    elif kind == "sincos":
        m.kind = 2
        p = [kw.get("base", 4.0), kw.get("imag", 0.1), kw.get("amp", 2.0), 0.0, 0.0, 0.0]
    elif kind == "janus":
        up, dn = complex(kw["up"]), complex(kw["down"])
        m.kind = 3
        p = [up.real, up.imag, dn.real, dn.imag, 0.0, 0.0]
    else:
        raise ValueError(kind)
    for i in range(6):
        m.p[i] = p[i]
    return m


def _owned(m):
    """Release the device program of a compiled material when it is collected."""
    import weakref
    keep = bfa_material()
    ctypes.pointer(keep)[0] = m
    weakref.finalize(m, lambda k=keep: lib().bfa_material_release(ctypes.byref(k)))
    return m


def material_svtensor(name, matprop=False):
    """Compiled material from an SVTensor file or a MatProp name (CONST_EPS_<value>, MATERIAL sections)."""
    m = bfa_material()
    _check(lib().bfa_material_svtensor(os.fsencode(name), int(bool(matprop)), ctypes.byref(m)),
           "bfa_material_svtensor")
    return _owned(m)


def material_expr(eps):
    """Compiled isotropic material eps(w, x, y, z, r, Theta, Phi) from one expression."""
    m = bfa_material()
    _check(lib().bfa_material_expr(eps.encode(), ctypes.byref(m)), "bfa_material_expr")
    return _owned(m)


def material_check(name=None, matprop=False, expr=None):
    """Parse and compile a material without a device (raises BFAError on errors)."""
    if expr is not None:
        _check(lib().bfa_material_expr_check(expr.encode()), "bfa_material_expr_check")
    else:
        _check(lib().bfa_material_check(os.fsencode(name), int(bool(matprop))), "bfa_material_check")


def material_eval(mat, omega, pts):
    """Permittivity tensors at points pts [n, 3] (device evaluation) -> [n, 3, 3] complex."""
    x = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    q = np.zeros((len(x), 9, 2))
    _check(lib().bfa_material_eval(ctypes.byref(mat), float(omega), len(x), x.ctypes.data, q.ctypes.data),
           "bfa_material_eval")
    return (q[..., 0] + 1j * q[..., 1]).reshape(-1, 3, 3)


def buffgeo_load(path):
    """.buffgeo file -> (V [nv, 3] transformed, T [nt, 4], compiled material)."""
    L = lib()
    nv, nt = I(), I()
    b = os.fsencode(path)
    _check(L.bfa_buffgeo_load(b, ctypes.byref(nv), ctypes.byref(nt), None, None, None), "bfa_buffgeo_load")
    V = np.zeros((nv.value, 3))
    T = np.zeros((nt.value, 4), dtype=np.int32)
    m = bfa_material()
    _check(L.bfa_buffgeo_load(b, ctypes.byref(nv), ctypes.byref(nt), V.ctypes.data, T.ctypes.data, ctypes.byref(m)),
           "bfa_buffgeo_load")
    return V, T, _owned(m)


def planewaves(specs):
    """specs: list of (E0 3-vector (complex), direction 3-vector)."""
    arr = (bfa_planewave * len(specs))()
    for j, (e0, d) in enumerate(specs):
        e0 = [complex(x) for x in e0]
        for c in range(3):
            arr[j].E0[2 * c] = e0[c].real
            arr[j].E0[2 * c + 1] = e0[c].imag
            arr[j].dir[c] = float(d[c])
    return arr


def kuhn_sphere(n, radius=1.0, jitter=0.05, seed=0x5EED):
    """Kuhn-subdivided sphere mesh: returns (V[nv,3], tets[nt,4])."""
    L = lib()
    nv, nt = I(), I()
    _check(L.bfa_mesh_kuhn_sphere(n, radius, jitter, seed, ctypes.byref(nv), ctypes.byref(nt), None, None),
           "bfa_mesh_kuhn_sphere")
    V = np.zeros((nv.value, 3))
    T = np.zeros((nt.value, 4), dtype=np.int32)
    _check(L.bfa_mesh_kuhn_sphere(n, radius, jitter, seed, ctypes.byref(nv), ctypes.byref(nt), V.ctypes.data,
                                  T.ctypes.data), "bfa_mesh_kuhn_sphere")
    return V, T


def read_vmsh(path):
    L = lib()
    nv, nt = I(), I()
    b = os.fsencode(path)
    _check(L.bfa_mesh_read_vmsh(b, ctypes.byref(nv), ctypes.byref(nt), None, None), "bfa_mesh_read_vmsh")
    V = np.zeros((nv.value, 3))
    T = np.zeros((nt.value, 4), dtype=np.int32)
    _check(L.bfa_mesh_read_vmsh(b, ctypes.byref(nv), ctypes.byref(nt), V.ctypes.data, T.ctypes.data),
           "bfa_mesh_read_vmsh")
    return V, T


class Geometry:
    """SWGGeometry counterpart for one object (src/libs/libbuff/SWGGeometry.cc:214-383)."""

    def __init__(self, V, T):
        L = lib()
        self.V = np.ascontiguousarray(V, dtype=np.float64)
        self.T = np.ascontiguousarray(T, dtype=np.int32)
        h = P()
        _check(L.bfa_geometry_create(len(self.V), self.V.ctypes.data, len(self.T), self.T.ctypes.data,
                                     ctypes.byref(h)), "bfa_geometry_create")
        self.h = h
        nv, nt, nbf, nft = I(), I(), I(), I()
        _check(L.bfa_geometry_info(h, ctypes.byref(nv), ctypes.byref(nt), ctypes.byref(nbf), ctypes.byref(nft)),
               "bfa_geometry_info")
        self.nv, self.nt, self.N, self.nfaces = nv.value, nt.value, nbf.value, nft.value

    def close(self):
        if getattr(self, "h", None):
            lib().bfa_geometry_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def faces(self):
        n = self.nfaces
        out = {k: np.zeros(n, dtype=np.int32) for k in ("QP", "QM", "PT", "MT", "PI", "MI")}
        fV = np.zeros((n, 3), dtype=np.int32)
        area = np.zeros(n)
        _check(lib().bfa_geometry_faces(self.h, fV.ctypes.data, out["QP"].ctypes.data, out["QM"].ctypes.data,
                                        out["PT"].ctypes.data, out["MT"].ctypes.data, out["PI"].ctypes.data,
                                        out["MI"].ctypes.data, area.ctypes.data), "bfa_geometry_faces")
        out["V"] = fV
        out["Area"] = area
        return out

    def touch_pairs(self):
        nnz = I64()
        _check(lib().bfa_geometry_touch_count(self.h, ctypes.byref(nnz)), "bfa_geometry_touch_count")
        rowptr = np.zeros(self.N + 1, dtype=np.int32)
        cols = np.zeros(nnz.value, dtype=np.int32)
        _check(lib().bfa_geometry_touch_pairs(self.h, rowptr.ctypes.data, cols.ctypes.data),
               "bfa_geometry_touch_pairs")
        return rowptr, cols

    def upload(self, stream=None):
        _check(lib().bfa_geometry_upload(self.h, _stream_ptr(stream)), "bfa_geometry_upload")

    def build_static(self, td_order=0, stream=None):
        _check(lib().bfa_static_build(self.h, int(td_order), _stream_ptr(stream)), "bfa_static_build")

    def static_table(self):
        nnz = I64()
        _check(lib().bfa_geometry_touch_count(self.h, ctypes.byref(nnz)), "bfa_geometry_touch_count")
        t = np.zeros((nnz.value, 6))
        _check(lib().bfa_static_download(self.h, t.ctypes.data), "bfa_static_download")
        return t

    def get_fields(self, omega, X, J=None, pws=None, stream=None):
        """Total fields at points X [n, 3]: (E [n, 3], H [n, 3]) complex; J: N complex128 on the device."""
        x = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
        F = np.zeros((len(x), 6, 2))
        npw = len(pws) if pws is not None else 0
        _check(lib().bfa_get_fields(self.h, float(omega), pws if npw else None, npw,
                                    _ptr(J) if J is not None else None, len(x), x.ctypes.data, F.ctypes.data,
                                    _stream_ptr(stream)), "bfa_get_fields")
        Fc = F[..., 0] + 1j * F[..., 1]
        return Fc[:, :3], Fc[:, 3:]

    def assemble_g(self, omega, d_M, ld=None, stream=None):
        """AssembleGBlock (self block): the VIE matrix without the chi^-1 overlap terms."""
        omega = complex(omega)
        _check(lib().bfa_assemble_g(self.h, omega.real, omega.imag, _ptr(d_M), int(ld or self.N), _stream_ptr(stream)),
               "bfa_assemble_g")

    def overlap_blocks(self, omega, mat, temperature=None, t_env=0.0):
        """AssembleOverlapBlocks: per row <= 7 entries -> (cnt [N], cols [N, 7], V, VInv [N, 7] complex, Rytov [N, 7])."""
        N = self.N
        cnt = np.zeros(N, dtype=np.int32)
        cols = np.zeros((N, 7), dtype=np.int32)
        V = np.zeros((N, 7, 2))
        VI = np.zeros((N, 7, 2))
        R = np.zeros((N, 7))
        _check(lib().bfa_overlap_blocks(self.h, float(omega), ctypes.byref(mat),
                                        ctypes.byref(temperature) if temperature is not None else None, float(t_env),
                                        cnt.ctypes.data, cols.ctypes.data, V.ctypes.data, VI.ctypes.data,
                                        R.ctypes.data), "bfa_overlap_blocks")
        return cnt, cols, V[..., 0] + 1j * V[..., 1], VI[..., 0] + 1j * VI[..., 1], R

    def gram_inverse(self, d_Sinv, ld=None, stream=None):
        """S^-1 of the SWG overlap (Gram) matrix into d_Sinv (N x N complex128), as buff-neq's CreateBNEQData."""
        _check(lib().bfa_gram_inverse(self.h, _ptr(d_Sinv), int(ld or self.N), _stream_ptr(stream)), "bfa_gram_inverse")

    def dressed_rytov(self, omega, mat, d_Sinv, d_DR, d_work, temperature=None, t_env=0.0, ld=None, stream=None):
        """ComputeDressedRytovMatrix: d_DR <- W (S^-1 R S^-1)^H; d_work of dressed_rytov_work_bytes() bytes."""
        _check(lib().bfa_dressed_rytov(self.h, float(omega), ctypes.byref(mat),
                                       ctypes.byref(temperature) if temperature is not None else None, float(t_env),
                                       _ptr(d_Sinv), _ptr(d_DR), int(ld or self.N), _ptr(d_work),
                                       _stream_ptr(stream)), "bfa_dressed_rytov")

    def dressed_rytov_work_bytes(self, ld=None):
        return int(lib().bfa_dressed_rytov_work_bytes(self.N, int(ld or self.N)))

    def emt_power_dr(self, omega, d_DR, ld=None, stream=None):
        """EMT power of a dressed Rytov matrix (no incident field): (P_abs, P_scat)."""
        out = np.zeros(2)
        _check(lib().bfa_emt_power_dr(self.h, float(omega), _ptr(d_DR), int(ld or self.N), out.ctypes.data,
                                      _stream_ptr(stream)), "bfa_emt_power_dr")
        return out

    def static_upload(self, table):
        """Static records (nnz x 6, touching-pair CSR order) from the caller."""
        t = np.ascontiguousarray(table, dtype=np.float64)
        _check(lib().bfa_static_upload(self.h, t.ctypes.data), "bfa_static_upload")

    def fibbi_keys(self):
        """FIBBI cache search keys (nnz x 27 float32) of the touching pairs."""
        nnz = I64()
        _check(lib().bfa_geometry_touch_count(self.h, ctypes.byref(nnz)), "bfa_geometry_touch_count")
        k = np.zeros((nnz.value, 27), dtype=np.float32)
        _check(lib().bfa_fibbi_keys(self.h, k.ctypes.data), "bfa_fibbi_keys")
        return k

    def fibbi_store(self, path):
        """Write the static table as a FIBBI cache file; returns the record count."""
        n = I64()
        _check(lib().bfa_fibbi_store(self.h, os.fsencode(path), ctypes.byref(n)), "bfa_fibbi_store")
        return n.value

    def fibbi_preload(self, path, td_order=0, stream=None):
        """Static table from a FIBBI cache file (misses computed); returns (hits, misses)."""
        h, m = I64(), I64()
        _check(lib().bfa_fibbi_preload(self.h, os.fsencode(path), int(td_order), ctypes.byref(h), ctypes.byref(m),
                                       _stream_ptr(stream)), "bfa_fibbi_preload")
        return h.value, m.value

    def assemble_vie(self, omega, mat, d_M, ld=None, stream=None, far32=False):
        """VIE matrix at omega; far32 evaluates the far-field kernel with its
        remainder beyond second order in FP32 (the precision form)."""
        omega = complex(omega)
        fn = lib().bfa_assemble_vie_far32 if far32 else lib().bfa_assemble_vie
        _check(fn(self.h, omega.real, omega.imag, ctypes.byref(mat), _ptr(d_M), int(ld or self.N),
                  _stream_ptr(stream)), "bfa_assemble_vie")

    def assemble_rhs(self, omega, pws, d_B, ldb=None, stream=None):
        _check(lib().bfa_assemble_rhs(self.h, float(omega), pws, len(pws), _ptr(d_B), int(ldb or self.N),
                                      _stream_ptr(stream)), "bfa_assemble_rhs")

    def emt_power(self, omega, pws, d_J, ldj=None, stream=None):
        out = np.zeros(2 * len(pws))
        _check(lib().bfa_emt_power(self.h, float(omega), pws, len(pws), _ptr(d_J), int(ldj or self.N),
                                   out.ctypes.data, _stream_ptr(stream)), "bfa_emt_power")
        return out.reshape(-1, 2)


def lu_factor(d_A, n, d_ipiv, lda=None, stream=None):
    info = I()
    _check(lib().bfa_lu_factor_z(_ptr(d_A), int(n), int(lda or n), _ptr(d_ipiv), ctypes.byref(info),
                                 _stream_ptr(stream)), "bfa_lu_factor_z")
    return info.value


def lu_solve(d_A, n, d_ipiv, d_B, nrhs, lda=None, ldb=None, stream=None):
    _check(lib().bfa_lu_solve_z(_ptr(d_A), int(n), int(lda or n), _ptr(d_ipiv), _ptr(d_B), int(nrhs),
                                int(ldb or n), _stream_ptr(stream)), "bfa_lu_solve_z")


def lu_solve_noalloc(d_A, n, d_ipiv, d_B, nrhs, lda=None, ldb=None, stream=None):
    """lu_solve with the library's static scratch array: no device allocation
    (n <= 20480, nrhs <= 8; concurrent calls on one GPU race)."""
    _check(lib().bfa_lu_solve_z_noalloc(_ptr(d_A), int(n), int(lda or n), _ptr(d_ipiv), _ptr(d_B), int(nrhs),
                                        int(ldb or n), _stream_ptr(stream)), "bfa_lu_solve_z_noalloc")


def zgemm_sub(m, n, k, d_A, lda, d_B, ldb, d_C, ldc, stream=None):
    """C -= A @ B on the device (column-major complex128)."""
    _check(lib().bfa_zgemm_sub(int(m), int(n), int(k), _ptr(d_A), int(lda), _ptr(d_B), int(ldb), _ptr(d_C),
                               int(ldc), _stream_ptr(stream)), "bfa_zgemm_sub")


def cgemm_sub(m, n, k, d_A, lda, d_B, ldb, d_C, ldc, stream=None):
    """C -= A @ B on the device (column-major complex64)."""
    _check(lib().bfa_cgemm_sub(int(m), int(n), int(k), _ptr(d_A), int(lda), _ptr(d_B), int(ldb), _ptr(d_C),
                               int(ldc), _stream_ptr(stream)), "bfa_cgemm_sub")


def lu_partition(on=True):
    """CU-partitioned look-ahead LU on this device (default on); off keeps every stream on every CU."""
    _check(lib().bfa_lu_partition(1 if on else 0), "bfa_lu_partition")


def profile_enable(on=True):
    _check(lib().bfa_profile_enable(1 if on else 0), "bfa_profile_enable")


def profile_read(which=0):
    """(kernel ms, algorithmic flops, launches) of the GEMM kernel since the last read."""
    ms, fl, nl = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    _check(lib().bfa_profile_read(int(which), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(nl)),
           "bfa_profile_read")
    return ms.value, fl.value, nl.value


def convert_z2c(d_A64, m, n, d_A32, lda=None, lda32=None, stream=None):
    _check(lib().bfa_convert_z2c(_ptr(d_A64), int(m), int(n), int(lda or m), _ptr(d_A32), int(lda32 or m),
                                 _stream_ptr(stream)), "bfa_convert_z2c")


def lu_factor_c(d_A, n, d_ipiv, lda=None, stream=None):
    info = I()
    _check(lib().bfa_lu_factor_c(_ptr(d_A), int(n), int(lda or n), _ptr(d_ipiv), ctypes.byref(info),
                                 _stream_ptr(stream)), "bfa_lu_factor_c")
    return info.value


def lu_factor_c_local(d_A, n, d_ipiv, lda=None, stream=None):
    """complex64 LU, pivoting within 256-row diagonal blocks, split-FP16 trailing updates."""
    info = I()
    _check(lib().bfa_lu_factor_c_local(_ptr(d_A), int(n), int(lda or n), _ptr(d_ipiv), ctypes.byref(info),
                                       _stream_ptr(stream)), "bfa_lu_factor_c_local")
    return info.value


def cgemm_sub_split(m, n, k, d_A, lda, d_B, ldb, d_C, ldc, stream=None):
    """C -= A @ B (complex64, k <= 256) by split-FP16 MFMA products."""
    _check(lib().bfa_cgemm_sub_split(int(m), int(n), int(k), _ptr(d_A), int(lda), _ptr(d_B), int(ldb), _ptr(d_C),
                                     int(ldc), _stream_ptr(stream)), "bfa_cgemm_sub_split")


def lu_solve_c(d_A, n, d_ipiv, d_B, nrhs, lda=None, ldb=None, stream=None):
    _check(lib().bfa_lu_solve_c(_ptr(d_A), int(n), int(lda or n), _ptr(d_ipiv), _ptr(d_B), int(nrhs),
                                int(ldb or n), _stream_ptr(stream)), "bfa_lu_solve_c")


def solve_refine(d_M64, d_LU32, d_ipiv, n, d_B, nrhs, tol=1e-10, maxit=10, ldm=None, ldlu=None, ldb=None,
                 stream=None):
    """Mixed-precision solve; returns (iterations, final relative residual)."""
    it, res = I(), D()
    _check(lib().bfa_solve_refine(_ptr(d_M64), int(ldm or n), _ptr(d_LU32), int(ldlu or n), _ptr(d_ipiv), int(n),
                                  _ptr(d_B), int(nrhs), int(ldb or n), float(tol), int(maxit), ctypes.byref(it),
                                  ctypes.byref(res), _stream_ptr(stream)), "bfa_solve_refine")
    return it.value, res.value


class Batch:
    """The batching form: many instances (Geometry, material) swept over
    frequencies in grouped launches, FP32 LU + FP64 refinement, EMT power.

    geoms must be uploaded with their static tables built; they must outlive
    the batch.  pws: planewaves(...) result (at most 8)."""

    def __init__(self, geoms, mats, pws):
        self._geoms = list(geoms)
        n = len(self._geoms)
        gh = (P * n)(*[g.h for g in self._geoms])
        ma = (bfa_material * n)(*mats)
        self.npw = len(pws)
        h = P()
        _check(lib().bfa_batch_create(n, gh, ma, pws, self.npw, ctypes.byref(h)), "bfa_batch_create")
        self._h = h
        self.count = n
        self.N = np.zeros(n, np.int32)
        self.Np = np.zeros(n, np.int32)
        _check(lib().bfa_batch_info(h, None, None, self.N.ctypes.data, self.Np.ctypes.data), "bfa_batch_info")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.bfa_batch_destroy(h)
            self._h = None

    def tune(self, reps=3):
        _check(lib().bfa_batch_tune(self._h, int(reps)), "bfa_batch_tune")
        tile = np.zeros(16, np.int32)
        ms = np.zeros((16, 3))
        _check(lib().bfa_batch_tuning(self._h, tile.ctypes.data, ms.ctypes.data), "bfa_batch_tuning")
        return tile, ms

    def set_tiles(self, tile):
        """tile variant per size bucket (16 values, -1 = the fixed rule)"""
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(tile, np.int32), (16,)), np.int32)
        _check(lib().bfa_batch_set_tiles(self._h, t.ctypes.data), "bfa_batch_set_tiles")

    def run(self, omegas, use_graph=True, tol=1e-10, maxit=6, with_power=True):
        """Returns dict with power [nf, count, npw, 2] (None without the EMT pass),
        iters/resid/info [nf, count]."""
        w = np.ascontiguousarray(omegas, dtype=np.float64)
        nf = len(w)
        power = np.zeros((nf, self.count, self.npw, 2)) if with_power else None
        iters = np.zeros((nf, self.count), np.int32)
        resid = np.zeros((nf, self.count))
        info = np.zeros((nf, self.count), np.int32)
        _check(lib().bfa_batch_run(self._h, w.ctypes.data, nf, int(bool(use_graph)), float(tol), int(maxit),
                                   power.ctypes.data if with_power else None, iters.ctypes.data,
                                   resid.ctypes.data, info.ctypes.data), "bfa_batch_run")
        return {"power": power, "iters": iters, "resid": resid, "info": info}

    def timing(self):
        ms = np.zeros(6)
        _check(lib().bfa_batch_timing(self._h, ms.ctypes.data), "bfa_batch_timing")
        return {"solve_path_ms": ms[0], "assembly_ms": ms[1], "factor_ms": ms[2], "solve_ms": ms[3],
                "power_phase_ms": ms[4], "power_ms": ms[5]}

    def solution(self, i):
        J = np.zeros((self.npw, int(self.N[i])), np.complex128)
        _check(lib().bfa_batch_solution(self._h, int(i), J.ctypes.data), "bfa_batch_solution")
        return J.T

    def matrix(self, i):
        n = int(self.N[i])
        T = np.zeros((n, n), np.complex128)
        _check(lib().bfa_batch_matrix(self._h, int(i), T.ctypes.data), "bfa_batch_matrix")
        return T.T


def fibbi_read(path):
    """Records of a FIBBI cache file: (keys [n, 27] float32, data [n, 6] float64)."""
    n = I64(0)
    _check(lib().bfa_fibbi_read(os.fsencode(path), ctypes.byref(n), None, None), "bfa_fibbi_read")
    k = np.zeros((n.value, 27), dtype=np.float32)
    d = np.zeros((n.value, 6))
    _check(lib().bfa_fibbi_read(os.fsencode(path), ctypes.byref(n), k.ctypes.data, d.ctypes.data), "bfa_fibbi_read")
    return k, d


def fibbi_write(path, keys, data):
    """Write records as a FIBBI cache file (the reference's format)."""
    k = np.ascontiguousarray(keys, dtype=np.float32).reshape(-1, 27)
    d = np.ascontiguousarray(data, dtype=np.float64).reshape(-1, 6)
    if len(k) != len(d):
        raise ValueError("keys and data differ in length")
    _check(lib().bfa_fibbi_write(os.fsencode(path), len(k), k.ctypes.data, d.ctypes.data), "bfa_fibbi_write")


def mlp_param_count(nin, hidden=256, nout=32):
    n, kp = I64(), I()
    _check(lib().bfa_mlp_param_count(int(nin), int(hidden), int(nout), ctypes.byref(n), ctypes.byref(kp)),
           "bfa_mlp_param_count")
    return n.value, kp.value


def mlp_grad(params, X, T, nin, hidden=256, nout=32):
    """GPU forward + MSE + backward of E ensemble members on one batch.
    params [E][nparam] float32, X [E][B][nin], T [E][B][nout] -> (loss [E], grad [E][nparam], Y [E][B][nout])."""
    params = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    E, B = X.shape[0], X.shape[1]
    loss = np.zeros(E)
    grad = np.zeros_like(params)
    Y = np.zeros((E, B, nout), np.float32)
    _check(lib().bfa_mlp_grad(int(nin), int(hidden), int(nout), E, params.ctypes.data, B, X.ctypes.data,
                              T.ctypes.data, loss.ctypes.data, grad.ctypes.data, Y.ctypes.data), "bfa_mlp_grad")
    return loss, grad, Y


class Surrogate:
    """Learned near-field surrogate over the non-touching tetrahedron pairs of
    one geometry closer than `factor` diameters (4x4 complex blocks)."""

    RULE, CUBATURE, SURROGATE = 0, 1, 2
    DEFAULT_TOL = 0.3

    def __init__(self, geom, factor=3.0, ensemble=4, hidden=256, seed=1):
        self.geom = geom
        h = P()
        _check(lib().bfa_surr_create(geom.h, float(factor), int(ensemble), int(hidden), int(seed),
                                     ctypes.byref(h)), "bfa_surr_create")
        self._h = h
        n = I()
        _check(lib().bfa_surr_pairs(h, ctypes.byref(n), None, None), "bfa_surr_pairs")
        self.npairs = n.value
        self.ensemble = int(ensemble)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.bfa_surr_destroy(h)
            self._h = None

    def pairs(self):
        ps = np.zeros(self.npairs, np.int32)
        pt = np.zeros(self.npairs, np.int32)
        _check(lib().bfa_surr_pairs(self._h, None, ps.ctypes.data, pt.ctypes.data), "bfa_surr_pairs")
        return ps, pt

    def cubature(self, pair, omega, npts=33, level=1):
        pair = np.ascontiguousarray(pair, np.int32)
        omega = np.ascontiguousarray(np.broadcast_to(omega, pair.shape), np.float64)
        out = np.zeros((len(pair), 32))
        _check(lib().bfa_surr_cubature(self._h, pair.ctypes.data, omega.ctypes.data, len(pair), int(npts),
                                       int(level), out.ctypes.data), "bfa_surr_cubature")
        return (out[:, 0::2] + 1j * out[:, 1::2]).reshape(-1, 4, 4)

    def features(self, pair, omega):
        pair = np.ascontiguousarray(pair, np.int32)
        omega = np.ascontiguousarray(np.broadcast_to(omega, pair.shape), np.float64)
        out = np.zeros((len(pair), 19), np.float32)
        _check(lib().bfa_surr_features(self._h, pair.ctypes.data, omega.ctypes.data, len(pair), out.ctypes.data),
               "bfa_surr_features")
        return out

    def dataset(self, per_pair=2, wmin=0.1, wmax=10.0, label_npts=33, label_level=1):
        _check(lib().bfa_surr_dataset(self._h, int(per_pair), float(wmin), float(wmax), int(label_npts),
                                      int(label_level)), "bfa_surr_dataset")

    def train(self, epochs=10, batch=4096, lr=1e-3):
        hist = np.zeros((epochs, self.ensemble))
        _check(lib().bfa_surr_train(self._h, int(epochs), int(batch), float(lr), hist.ctypes.data), "bfa_surr_train")
        return hist

    def table(self, omega, mode=2, tol=None, npts=33, level=1):
        """Build the block table at omega; returns (cubature fallbacks, max ensemble disagreement).
        tol: largest accepted disagreement of the members, in units of
        ||R16 - R4|| (default DEFAULT_TOL)."""
        if tol is None:
            tol = self.DEFAULT_TOL
        nf, dm = I(), D()
        _check(lib().bfa_surr_table(self._h, float(omega), int(mode), float(tol), int(npts), int(level),
                                    ctypes.byref(nf), ctypes.byref(dm)), "bfa_surr_table")
        return nf.value, dm.value

    def table_download(self):
        out = np.zeros((self.npairs, 32))
        _check(lib().bfa_surr_table_download(self._h, out.ctypes.data), "bfa_surr_table_download")
        return (out[:, 0::2] + 1j * out[:, 1::2]).reshape(-1, 4, 4)

    def table_upload(self, blocks):
        """blocks: complex (npairs, 4, 4), the layout of table_download"""
        b = np.asarray(blocks, np.complex128).reshape(self.npairs, 16)
        out = np.ascontiguousarray(np.stack([b.real, b.imag], axis=-1).reshape(self.npairs, 32))
        _check(lib().bfa_surr_table_upload(self._h, out.ctypes.data), "bfa_surr_table_upload")

    def assemble(self, omega, mat, d_M, ld=None, stream=None):
        _check(lib().bfa_assemble_vie_table(self._h, float(omega), ctypes.byref(mat), _ptr(d_M),
                                            int(ld or self.geom.N), _stream_ptr(stream)), "bfa_assemble_vie_table")

    def params(self, new=None):
        E, n = I(), I64()
        _check(lib().bfa_surr_info(self._h, ctypes.byref(E), ctypes.byref(n), None, None), "bfa_surr_info")
        out = np.zeros((E.value, n.value), np.float32)
        src = None if new is None else np.ascontiguousarray(new, np.float32)
        _check(lib().bfa_surr_params(self._h, out.ctypes.data, None if src is None else src.ctypes.data),
               "bfa_surr_params")
        return out
