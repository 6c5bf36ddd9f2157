"""Python binding of ``libcvchip.so`` (ctypes over the C ABI in include/cvchip.h).

The HIP library is the only compute path: importing this module loads
``cvc_amd/libcvchip.so`` and raises if it is missing.  Arrays use the
reference's host layouts: spinor ``float64[V, 24]`` (or ``[n, V, 24]`` for n
right-hand sides), gauge ``float64[V, 4, 18]``, conn ``float64[V, 16, 2]``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcvchip.so")

PREC_F64, PREC_F32, PREC_F16 = 0, 1, 2
SOLVE_LAGGED, SOLVE_VERBOSE, SOLVE_EO = 1, 2, 4
ERR_NULL, ERR_NOCONV = -2, -3
ERR_ARG, ERR_HIP, ERR_OOM, ERR_STATE = -100, -101, -102, -103

# every entry point declared in include/cvchip.h
EXPORTS = [
    "cvchip_ctx_create", "cvchip_ctx_destroy", "cvchip_ctx_set_params", "cvchip_ctx_set_link_recon",
    "cvchip_ctx_set_solver_flags", "cvchip_piq_orbits",
    "cvchip_lime_write_contraction", "cvchip_lime_read_contraction", "cvchip_lime_read_gauge",
    "cvchip_lime_write_gauge", "cvchip_lime_read_spinor", "cvchip_lime_write_spinor",
    "cvchip_lime_write_propagator", "cvchip_cvc_propagators", "cvchip_field_point_sources", "cvchip_field_zero",
    "cvchip_field_copy_columns", "cvchip_gamma_tables", "cvchip_contract_twopoint",
    "cvchip_field_contract_twopoint", "cvchip_field_contract_twopoint_noalloc",
    "cvchip_field_sequential_sources", "cvchip_p2gg_source",
    "cvchip_ctx_set_dslash_policy",
    "cvchip_last_error", "cvchip_set_gauge", "cvchip_Q_phi", "cvchip_scalar_product_co",
    "cvchip_scalar_product_re", "cvchip_invert_Qtm", "cvchip_invert_Qtm_her", "cvchip_field_create",
    "cvchip_field_destroy", "cvchip_field_upload", "cvchip_field_download", "cvchip_field_copy",
    "cvchip_field_fill_random",
    "cvchip_field_dslash",
    "cvchip_field_dslash_timed", "cvchip_field_invert", "cvchip_field_invert_her", "cvchip_cvc_contract",
    "cvchip_field_cvc_contract", "cvchip_wi", "cvchip_pi_q", "cvchip_cvc_source", "cvchip_cvc_gradient",
    "cvchip_gauge_random",
]


class CvchipError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"libcvchip error {rc}: {msg}")
        self.rc = rc


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("reliable_updates", C.c_int),
                ("dslash_calls", C.c_int), ("true_res2_max", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


def lib():
    """Load libcvchip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not found: build it with `python -m cvc_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.cvchip_last_error.restype = C.c_char_p
        L.cvchip_ctx_create.argtypes = [C.POINTER(_vp), C.POINTER(C.c_int), _dp, C.c_double, C.c_double]
        L.cvchip_ctx_destroy.argtypes = [_vp]
        L.cvchip_ctx_set_params.argtypes = [_vp, C.c_double, C.c_double]
        L.cvchip_ctx_set_link_recon.argtypes = [_vp, C.c_int]
        L.cvchip_ctx_set_solver_flags.argtypes = [_vp, C.c_uint]
        L.cvchip_ctx_set_dslash_policy.argtypes = [_vp, C.c_int]
        L.cvchip_set_gauge.argtypes = [_vp, _dp]
        L.cvchip_Q_phi.argtypes = [_vp, _dp, _dp, C.c_double, C.c_int]
        L.cvchip_scalar_product_co.argtypes = [_vp, _dp, _dp, _dp]
        L.cvchip_scalar_product_re.argtypes = [_vp, _dp, _dp, _dp]
        L.cvchip_invert_Qtm.argtypes = [_vp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int,
                                        C.c_double, C.c_uint, C.POINTER(Stats)]
        L.cvchip_invert_Qtm_her.argtypes = [_vp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int,
                                            C.c_double, C.c_uint, C.POINTER(Stats)]
        L.cvchip_field_create.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(_vp)]
        L.cvchip_field_destroy.argtypes = [_vp]
        L.cvchip_field_upload.argtypes = [_vp, _vp, _dp, C.c_int, C.c_int]
        L.cvchip_field_download.argtypes = [_vp, _vp, _dp, C.c_int, C.c_int]
        L.cvchip_field_copy.argtypes = [_vp, _vp, _vp]
        L.cvchip_field_fill_random.argtypes = [_vp, _vp, C.c_uint64]
        L.cvchip_field_dslash.argtypes = [_vp, _vp, _vp, C.c_double, C.c_int]
        L.cvchip_field_dslash_timed.argtypes = [_vp, _vp, _vp, C.c_double, C.c_int, _dp]
        L.cvchip_field_invert.argtypes = [_vp, _vp, _vp, C.c_double, C.c_int, C.c_int, C.c_double,
                                          C.c_uint, C.POINTER(Stats)]
        L.cvchip_field_invert_her.argtypes = [_vp, _vp, _vp, C.c_double, C.c_int, C.c_int, C.c_double,
                                              C.c_uint, C.POINTER(Stats)]
        L.cvchip_cvc_contract.argtypes = [_vp, _dp, _dp, C.POINTER(_dp), C.POINTER(C.c_int)]
        L.cvchip_field_cvc_contract.argtypes = [_vp, _dp, _dp, _vp, _vp, C.POINTER(C.c_int)]
        L.cvchip_wi.argtypes = [_vp, _dp, _dp]
        L.cvchip_pi_q.argtypes = [_vp, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.cvchip_cvc_source.argtypes = [_vp, _dp, _dp, C.POINTER(C.c_int), C.c_double, C.c_int, C.c_int,
                                        C.c_double, C.POINTER(Stats)]
        L.cvchip_cvc_gradient.argtypes = [_vp, C.POINTER(C.c_int), C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_double, _dp, _dp, _dp, C.POINTER(Stats)]
        L.cvchip_gauge_random.argtypes = [C.POINTER(C.c_int), C.c_uint64, C.c_double, _dp]
        L.cvchip_piq_orbits.argtypes = [C.POINTER(C.c_int), _dp, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]
        _u32p = C.POINTER(C.c_uint32)
        L.cvchip_lime_write_contraction.argtypes = [C.c_char_p, _dp, C.c_int, C.c_int, C.POINTER(C.c_int),
                                                    C.c_char_p, C.c_int, C.c_int, _u32p]
        L.cvchip_lime_read_contraction.argtypes = [C.c_char_p, _dp, C.c_int, C.c_int, C.POINTER(C.c_int), _u32p]
        L.cvchip_lime_read_gauge.argtypes = [C.c_char_p, C.POINTER(C.c_int), _dp, _u32p]
        L.cvchip_lime_write_gauge.argtypes = [C.c_char_p, C.POINTER(C.c_int), _dp, C.c_int, _u32p]
        L.cvchip_lime_read_spinor.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_int, C.c_double, _dp, _u32p]
        L.cvchip_lime_write_spinor.argtypes = [C.c_char_p, C.POINTER(C.c_int), _dp, C.c_int, C.c_int, _u32p]
        L.cvchip_lime_write_propagator.argtypes = [C.c_char_p, C.POINTER(C.c_int), _dp, C.c_int, _u32p]
        L.cvchip_field_zero.argtypes = [_vp, _vp]
        L.cvchip_field_point_sources.argtypes = [_vp, _vp, C.c_int, C.POINTER(C.c_int)]
        L.cvchip_field_copy_columns.argtypes = [_vp, _vp, C.c_int, _vp, C.c_int, C.c_int]
        L.cvchip_cvc_propagators.argtypes = [_vp, C.POINTER(_dp), C.POINTER(C.c_int), C.c_double, C.c_int, C.c_int,
                                             C.c_double, C.POINTER(Stats)]
        L.cvchip_gamma_tables.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.cvchip_contract_twopoint.argtypes = [_vp, _dp, C.c_int, C.c_int, C.POINTER(_dp), C.POINTER(_dp), C.c_int,
                                               C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.cvchip_field_contract_twopoint.argtypes = [_vp, _dp, C.c_int, C.c_int, _vp, _vp, C.c_int,
                                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.cvchip_field_contract_twopoint_noalloc.argtypes = [_vp, C.c_void_p, C.c_int, C.c_int, _vp, _vp, C.c_int,
                                                             C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.cvchip_field_sequential_sources.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                                      C.POINTER(C.c_int), C.c_double]
        L.cvchip_p2gg_source.argtypes = [_vp, _dp, _dp, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_int),
                                         C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.POINTER(Stats)]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().cvchip_last_error().decode(errors="replace")


def _check(rc):
    if rc < 0:
        raise CvchipError(rc, last_error())
    return rc


def _ptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def gauge_random(dims, seed, eps=0.3) -> np.ndarray:
    """U_mu(x) = exp(i eps H) with Philox-seeded Gaussian H (host, deterministic)."""
    dims = [int(d) for d in dims]
    V = int(np.prod(dims))
    g = np.empty((V, 4, 18))
    _check(lib().cvchip_gauge_random((C.c_int * 4)(*dims), C.c_uint64(int(seed)), float(eps), _ptr(g)))
    return g


def piq_orbits(dims, piq, max_orbits=1 << 20):
    """Scalar vacuum polarisation Pi(qhat^2) = -sum_mu Re Pi_mumu(q) / (3 qhat^2) averaged on
    orbits of equal qhat^2 (q = 0 left out): returns qhat2, pi, pi_sd, count (arrays)."""
    dims = [int(d) for d in dims]
    piq = np.ascontiguousarray(piq, dtype=np.float64)
    assert piq.shape == (int(np.prod(dims)), 16, 2)
    m = min(int(max_orbits), int(np.prod(dims)))
    q2, pi, sd = np.empty(m), np.empty(m), np.empty(m)
    cnt = np.empty(m, dtype=np.int32)
    no = C.c_int(0)
    _check(lib().cvchip_piq_orbits((C.c_int * 4)(*dims), _ptr(piq), m, _ptr(q2), _ptr(pi), _ptr(sd),
                                   cnt.ctypes.data_as(C.POINTER(C.c_int)), C.byref(no)))
    k = no.value
    return q2[:k], pi[:k], sd[:k], cnt[:k]


def lime_write_contraction(filename, s, dims, prec=64, type="cvc - cvc in position space, all 16 components",
                           nconf=0, source=0):
    """write_lime_contraction: s float64[V, N, 2] -> LIME file; returns the (suma, sumb) checksum."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    ck = (C.c_uint32 * 2)()
    _check(lib().cvchip_lime_write_contraction(os.fsencode(filename), _ptr(s), int(s.shape[1]), int(prec),
                                               (C.c_int * 4)(*dims), type.encode(), int(nconf), int(source), ck))
    return int(ck[0]), int(ck[1])


def lime_read_contraction(filename, dims, N=16, position=0):
    """read_lime_contraction: -> (float64[V, N, 2], (suma, sumb))."""
    s = np.empty((int(np.prod(dims)), N, 2))
    ck = (C.c_uint32 * 2)()
    _check(lib().cvchip_lime_read_contraction(os.fsencode(filename), _ptr(s), int(N), int(position),
                                              (C.c_int * 4)(*dims), ck))
    return s, (int(ck[0]), int(ck[1]))


def lime_read_gauge(filename, dims):
    """read_lime_gauge_field_doubleprec: ILDG file -> (float64[V, 4, 18], (suma, sumb))."""
    g = np.empty((int(np.prod(dims)), 4, 18))
    ck = (C.c_uint32 * 2)()
    _check(lib().cvchip_lime_read_gauge(os.fsencode(filename), (C.c_int * 4)(*dims), _ptr(g), ck))
    return g, (int(ck[0]), int(ck[1]))


def lime_write_gauge(filename, dims, gauge, prec=64):
    g = np.ascontiguousarray(gauge, dtype=np.float64)
    ck = (C.c_uint32 * 2)()
    _check(lib().cvchip_lime_write_gauge(os.fsencode(filename), (C.c_int * 4)(*dims), _ptr(g), int(prec), ck))
    return int(ck[0]), int(ck[1])


def lime_read_spinor(filename, dims, position=0, kappa=0.0):
    """read_lime_spinor: record `position` (0 up, 1 down) -> (float64[V, 24], (suma, sumb));
    a file without such a record is read as CMI (scaled by 2 kappa)."""
    s = np.empty((int(np.prod(dims)), 24))
    ck = (C.c_uint32 * 2)()
    _check(lib().cvchip_lime_read_spinor(os.fsencode(filename), (C.c_int * 4)(*dims), int(position), float(kappa),
                                         _ptr(s), ck))
    return s, (int(ck[0]), int(ck[1]))


def lime_write_spinor(filename, dims, s, append=False, prec=64):
    """write_lime_spinor: s float64[V, 24] -> binary-data + checksum records; returns (suma, sumb)."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    ck = (C.c_uint32 * 2)()
    _check(lib().cvchip_lime_write_spinor(os.fsencode(filename), (C.c_int * 4)(*dims), _ptr(s), int(bool(append)),
                                          int(prec), ck))
    return int(ck[0]), int(ck[1])


def lime_write_propagator(filename, dims, s, prec=64):
    """write_propagator: new file, etmc-propagator-format record, then write_lime_spinor."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    ck = (C.c_uint32 * 2)()
    _check(lib().cvchip_lime_write_propagator(os.fsencode(filename), (C.c_int * 4)(*dims), _ptr(s), int(prec), ck))
    return int(ck[0]), int(ck[1])


def gamma_tables():
    """gamma_permutation, gamma_sign: int[16][24] each (the library's own tables)."""
    perm = np.zeros((16, 24), dtype=np.int32)
    sign = np.zeros((16, 24), dtype=np.int32)
    _check(lib().cvchip_gamma_tables(perm.ctypes.data_as(C.POINTER(C.c_int)), sign.ctypes.data_as(C.POINTER(C.c_int))))
    return perm, sign


def point_source(V, site, spin, colour) -> np.ndarray:
    """Unit source at one site, spin-colour component (cvc_exact2_xspace.c:521-524)."""
    b = np.zeros((V, 24))
    b[site, 2 * (3 * spin + colour)] = 1.0
    return b


class Field:
    """nrhs spinor fields resident in HBM (library layout)."""

    def __init__(self, ctx: "Context", nrhs: int, prec: int = PREC_F64):
        self.ctx, self.nrhs, self.prec = ctx, int(nrhs), int(prec)
        h = _vp()
        _check(lib().cvchip_field_create(ctx._h, self.nrhs, self.prec, C.byref(h)))
        self._h = h

    def upload(self, host: np.ndarray, rhs0: int = 0):
        host = np.ascontiguousarray(host, dtype=np.float64).reshape(-1, self.ctx.V, 24)
        _check(lib().cvchip_field_upload(self.ctx._h, self._h, _ptr(host), int(rhs0), host.shape[0]))

    def download(self, rhs0: int = 0, n: int | None = None) -> np.ndarray:
        n = self.nrhs - rhs0 if n is None else n
        out = np.empty((n, self.ctx.V, 24))
        _check(lib().cvchip_field_download(self.ctx._h, self._h, _ptr(out), int(rhs0), int(n)))
        return out

    def fill_random(self, seed: int):
        _check(lib().cvchip_field_fill_random(self.ctx._h, self._h, C.c_uint64(int(seed))))

    def copy_from(self, src: "Field"):
        _check(lib().cvchip_field_copy(self.ctx._h, self._h, src._h))

    def zero(self):
        _check(lib().cvchip_field_zero(self.ctx._h, self._h))

    def point_sources(self, coords):
        """Point sources of len(coords) source points (t, x, y, z): column 12 k + 3 spin + colour."""
        flat = [int(v) for c in coords for v in c]
        _check(lib().cvchip_field_point_sources(self.ctx._h, self._h, len(coords), (C.c_int * len(flat))(*flat)))

    def copy_columns(self, dst_col0, src: "Field", src_col0, n):
        """self[:, dst_col0:dst_col0+n] = src[:, src_col0:src_col0+n] (FP64 sets)."""
        _check(lib().cvchip_field_copy_columns(self.ctx._h, self._h, int(dst_col0), src._h, int(src_col0), int(n)))

    def close(self):
        if getattr(self, "_h", None):
            lib().cvchip_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One lattice on the current GPU: geometry, kappa, mu, boundary phases, gauge."""

    def __init__(self, dims, kappa, mu, bc=(1.0, 0.0, 0.0, 0.0), gauge=None):
        self.dims = tuple(int(d) for d in dims)
        self.V = int(np.prod(self.dims))
        self.kappa, self.mu = float(kappa), float(mu)
        h = _vp()
        _check(lib().cvchip_ctx_create(C.byref(h), (C.c_int * 4)(*self.dims), (C.c_double * 4)(*bc),
                                       self.kappa, self.mu))
        self._h = h
        if gauge is not None:
            self.set_gauge(gauge)

    def close(self):
        if getattr(self, "_h", None):
            lib().cvchip_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, kappa, mu):
        self.kappa, self.mu = float(kappa), float(mu)
        _check(lib().cvchip_ctx_set_params(self._h, self.kappa, self.mu))

    def set_link_recon(self, recon: int):
        _check(lib().cvchip_ctx_set_link_recon(self._h, int(recon)))

    def set_solver_flags(self, flags: int):
        """Flags OR-ed into every solve of this context (e.g. SOLVE_EO)."""
        _check(lib().cvchip_ctx_set_solver_flags(self._h, int(flags)))

    def set_dslash_policy(self, policy: int):
        _check(lib().cvchip_ctx_set_dslash_policy(self._h, int(policy)))

    def set_gauge(self, gauge: np.ndarray):
        g = np.ascontiguousarray(gauge, dtype=np.float64)
        assert g.size == self.V * 72
        _check(lib().cvchip_set_gauge(self._h, _ptr(g)))

    # -- host-array drop-ins ---------------------------------------------------
    def Q_phi(self, phi, mutm=None, gamma5_out=False, out=None):
        """xi = D(mutm) phi on host arrays [V][24] (Q_phi_tbc / Q_phi / Qf5); out, if given, is a
        C-contiguous float64 [V][24] array written in place (as the reference's xi)."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        if out is None:
            out = np.empty((self.V, 24))
        elif out.dtype != np.float64 or not out.flags.c_contiguous or out.size != 24 * self.V:
            raise ValueError("out must be a C-contiguous float64 array of V x 24 entries")
        _check(lib().cvchip_Q_phi(self._h, _ptr(out), _ptr(phi), self.mu if mutm is None else float(mutm),
                                  int(bool(gamma5_out))))
        return out

    def scalar_product_co(self, a, b):
        w = np.zeros(2)
        _check(lib().cvchip_scalar_product_co(self._h, _ptr(w), _ptr(np.ascontiguousarray(a)),
                                              _ptr(np.ascontiguousarray(b))))
        return complex(w[0], w[1])

    def scalar_product_re(self, a, b):
        w = np.zeros(1)
        _check(lib().cvchip_scalar_product_re(self._h, _ptr(w), _ptr(np.ascontiguousarray(a)),
                                              _ptr(np.ascontiguousarray(b))))
        return float(w[0])

    def invert_Qtm(self, b, sp, niter_max=1000, x0=None, prec=PREC_F64, delta=0.1, lagged=False,
                   verbose=False, check=True, eo=False):
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, self.V, 24)
        n = b.shape[0]
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, order="C").reshape(b.shape)
        st = Stats()
        flags = (SOLVE_LAGGED if lagged else 0) | (SOLVE_VERBOSE if verbose else 0) | (SOLVE_EO if eo else 0)
        rc = lib().cvchip_invert_Qtm(self._h, _ptr(x), _ptr(b), n, float(sp), int(niter_max), int(prec),
                                     float(delta), flags, C.byref(st))
        if check and rc < 0:
            raise CvchipError(rc, last_error())
        return rc, x, st.as_dict()

    def invert_Qtm_her(self, b, sp, niter_max=1000, x0=None, prec=PREC_F64, delta=0.1, lagged=False,
                       check=True, eo=False):
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, self.V, 24)
        n = b.shape[0]
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, order="C").reshape(b.shape)
        st = Stats()
        rc = lib().cvchip_invert_Qtm_her(self._h, _ptr(x), _ptr(b), n, float(sp), int(niter_max), int(prec),
                                         float(delta), (SOLVE_LAGGED if lagged else 0) | (SOLVE_EO if eo else 0),
                                         C.byref(st))
        if check and rc < 0:
            raise CvchipError(rc, last_error())
        return rc, x, st.as_dict()

    # -- device-resident fields ------------------------------------------------
    def field(self, nrhs, prec=PREC_F64) -> Field:
        return Field(self, nrhs, prec)

    def dslash(self, out: Field, inp: Field, mutm=None, gamma5_out=False):
        _check(lib().cvchip_field_dslash(self._h, out._h, inp._h, self.mu if mutm is None else float(mutm),
                                         int(bool(gamma5_out))))

    def dslash_timed(self, out: Field, inp: Field, reps: int, mutm=None) -> float:
        ms = np.zeros(1)
        _check(lib().cvchip_field_dslash_timed(self._h, out._h, inp._h, self.mu if mutm is None else float(mutm),
                                               int(reps), _ptr(ms)))
        return float(ms[0])

    def invert(self, x: Field, b: Field, sp, niter_max=1000, prec=PREC_F64, delta=0.1, lagged=False,
               check=True, eo=False):
        st = Stats()
        rc = lib().cvchip_field_invert(self._h, x._h, b._h, float(sp), int(niter_max), int(prec), float(delta),
                                       (SOLVE_LAGGED if lagged else 0) | (SOLVE_EO if eo else 0), C.byref(st))
        if check and rc < 0:
            raise CvchipError(rc, last_error())
        return rc, st.as_dict()

    def invert_her(self, x: Field, b: Field, sp, niter_max=1000, prec=PREC_F64, delta=0.1, check=True,
                   verbose=False, eo=False):
        st = Stats()
        rc = lib().cvchip_field_invert_her(self._h, x._h, b._h, float(sp), int(niter_max), int(prec),
                                           float(delta), (SOLVE_VERBOSE if verbose else 0) | (SOLVE_EO if eo else 0),
                                           C.byref(st))
        if check and rc < 0:
            raise CvchipError(rc, last_error())
        return rc, st.as_dict()

    # -- observables -----------------------------------------------------------
    def contract_twopoint(self, idsource, idsink, chi, phi, n_c=3, snk_mom=None, src=(0, 0, 0, 0)):
        """Meson two-point function C[T] (complex) of host propagator columns (4 n_c arrays [V,24])."""
        ca = [np.ascontiguousarray(a, dtype=np.float64) for a in chi]
        pa = [np.ascontiguousarray(a, dtype=np.float64) for a in phi]
        out = np.zeros((self.dims[0], 2))
        _check(lib().cvchip_contract_twopoint(self._h, _ptr(out), int(idsource), int(idsink),
                                              (_dp * len(ca))(*[_ptr(a) for a in ca]),
                                              (_dp * len(pa))(*[_ptr(a) for a in pa]), int(n_c),
                                              (C.c_int * 3)(*snk_mom) if snk_mom is not None else None,
                                              (C.c_int * 4)(*src)))
        return out[:, 0] + 1j * out[:, 1]

    def field_contract_twopoint(self, idsource, idsink, chi: Field, phi: Field, n_c=3, snk_mom=None,
                                src=(0, 0, 0, 0)):
        out = np.zeros((self.dims[0], 2))
        _check(lib().cvchip_field_contract_twopoint(self._h, _ptr(out), int(idsource), int(idsink), chi._h, phi._h,
                                                    int(n_c), (C.c_int * 3)(*snk_mom) if snk_mom is not None else None,
                                                    (C.c_int * 4)(*src)))
        return out[:, 0] + 1j * out[:, 1]

    def field_contract_twopoint_noalloc(self, contr_dev: int, idsource, idsink, chi: Field, phi: Field, n_c=3,
                                        snk_mom=None, src=(0, 0, 0, 0)):
        """Adds C(t) to the device buffer at address contr_dev (2 T doubles) without any
        allocation (see include/cvchip.h)."""
        _check(lib().cvchip_field_contract_twopoint_noalloc(
            self._h, C.c_void_p(int(contr_dev)), int(idsource), int(idsink), chi._h, phi._h, int(n_c),
            (C.c_int * 3)(*snk_mom) if snk_mom is not None else None, (C.c_int * 4)(*src)))

    def cvc_contract(self, props, src):
        assert len(props) == 120
        arrs = [np.ascontiguousarray(p, dtype=np.float64) for p in props]
        ptrs = (_dp * 120)(*[_ptr(a) for a in arrs])
        conn = np.empty((self.V, 16, 2))
        contact = np.zeros(8)
        _check(lib().cvchip_cvc_contract(self._h, _ptr(conn), _ptr(contact), ptrs, (C.c_int * 4)(*src)))
        return conn, contact[0::2] + 1j * contact[1::2]

    def field_cvc_contract(self, up: Field, dn: Field, src):
        conn = np.empty((self.V, 16, 2))
        contact = np.zeros(8)
        _check(lib().cvchip_field_cvc_contract(self._h, _ptr(conn), _ptr(contact), up._h, dn._h,
                                               (C.c_int * 4)(*src)))
        return conn, contact[0::2] + 1j * contact[1::2]

    def wi(self, conn):
        out = np.empty((self.V, 4, 2))
        _check(lib().cvchip_wi(self._h, _ptr(out), _ptr(np.ascontiguousarray(conn))))
        return out

    def pi_q(self, conn, contact, src):
        ct = np.zeros(8)
        ct[0::2] = np.real(contact)
        ct[1::2] = np.imag(contact)
        out = np.empty((self.V, 16, 2))
        _check(lib().cvchip_pi_q(self._h, _ptr(out), _ptr(np.ascontiguousarray(conn)), _ptr(ct),
                                 (C.c_int * 4)(*src)))
        return out

    def cvc_gradient(self, src, qmax=2.0, sp=1e-24, niter_max=5000, prec=PREC_F64, delta=0.1, force=True):
        """L = sum_mu sum_{|q|<=qmax} Re Pi_mumu(q), dL/dmu and the link force F[V,4,3,3]
        (dL/deps = tr(T F) for U -> exp(i eps T) U)."""
        Lv, dmu = np.zeros(1), np.zeros(1)
        F = np.empty((self.V, 4, 18)) if force else None
        st = Stats()
        _check(lib().cvchip_cvc_gradient(self._h, (C.c_int * 4)(*src), float(qmax), float(sp), int(niter_max),
                                         int(prec), float(delta), _ptr(Lv), _ptr(dmu),
                                         _ptr(F) if force else None, C.byref(st)))
        Fc = None
        if force:
            Fc = (F[..., 0::2] + 1j * F[..., 1::2]).reshape(self.V, 4, 3, 3)
        return float(Lv[0]), float(dmu[0]), Fc, st.as_dict()

    def cvc_propagators(self, src, sp, niter_max=1000, prec=PREC_F64, delta=0.1):
        """The 120 propagators of one source point, float64[120, V, 24] in g_spinor_field order."""
        props = np.empty((120, self.V, 24))
        ptrs = (_dp * 120)(*[_ptr(props[k]) for k in range(120)])
        st = Stats()
        _check(lib().cvchip_cvc_propagators(self._h, ptrs, (C.c_int * 4)(*src), float(sp), int(niter_max), int(prec),
                                            float(delta), C.byref(st)))
        return props, st.as_dict()

    def field_sequential_sources(self, out: Field, inp: Field, col0, tseq, gamma_id, mom=None, sgn=1.0):
        _check(lib().cvchip_field_sequential_sources(self._h, out._h, inp._h, int(col0), int(tseq), int(gamma_id),
                                                     (C.c_int * 3)(*mom) if mom is not None else None,
                                                     float(sgn)))

    def p2gg_source(self, src, seq_timeslice, seq_gamma_id, seq_mom=(0, 0, 0), propagator_position=0, sp=1e-20,
                    niter_max=1000, prec=PREC_F64, delta=0.1):
        """P -> gamma gamma for one source point: (conn[V,16,2], contact[4] complex, stats)."""
        conn = np.empty((self.V, 16, 2))
        contact = np.zeros(8)
        st = Stats()
        _check(lib().cvchip_p2gg_source(self._h, _ptr(conn), _ptr(contact), (C.c_int * 4)(*src), int(seq_timeslice),
                                        int(seq_gamma_id), (C.c_int * 3)(*seq_mom), int(propagator_position),
                                        float(sp), int(niter_max), int(prec), float(delta), C.byref(st)))
        return conn, contact[0::2] + 1j * contact[1::2], st.as_dict()

    def cvc_source(self, src, sp, niter_max=1000, prec=PREC_F64, delta=0.1):
        conn = np.empty((self.V, 16, 2))
        contact = np.zeros(8)
        st = Stats()
        _check(lib().cvchip_cvc_source(self._h, _ptr(conn), _ptr(contact), (C.c_int * 4)(*src), float(sp),
                                       int(niter_max), int(prec), float(delta), C.byref(st)))
        return conn, contact[0::2] + 1j * contact[1::2], st.as_dict()
