"""ctypes binding of the CPU restatement (``libcvcoracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
library never imports it.

Arrays use the reference's host layouts (reference global.h:55-61):
spinor ``float64[V, 24]``, gauge ``float64[V, 4, 18]``, conn ``float64[V, 16, 2]``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcvcoracle.so")


def build() -> str:
    """Compile the restatement with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


class _Lat(C.Structure):
    _fields_ = [
        ("T", C.c_int), ("LX", C.c_int), ("LY", C.c_int), ("LZ", C.c_int),
        ("V", C.c_long), ("kappa", C.c_double), ("mu", C.c_double),
        ("bc", C.c_double * 4), ("phase", (C.c_double * 2) * 4),
        ("perm", (C.c_int * 24) * 16), ("sign", (C.c_int * 24) * 16),
        ("iup", C.POINTER(C.c_int)), ("idn", C.POINTER(C.c_int)),
        ("nthreads", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(_Lat)
        dp = C.POINTER(C.c_double)
        L.orc_lat_init.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp]
        L.orc_lat_free.argtypes = [P]
        L.orc_Q_phi.argtypes = [P, dp, dp, dp, C.c_double]
        L.orc_Qf5.argtypes = [P, dp, dp, dp, C.c_double]
        L.orc_g5_phi.argtypes = [P, dp]
        L.orc_dot_co.argtypes = [P, dp, dp, dp]
        L.orc_dot_re.argtypes = [P, dp, dp]
        L.orc_dot_re.restype = C.c_double
        L.orc_invert_Qtm.argtypes = [P, dp, dp, dp, C.c_double, C.c_int, C.c_int, dp]
        L.orc_invert_Qtm_her.argtypes = [P, dp, dp, dp, C.c_double, C.c_int, dp]
        L.orc_cvc_contract.argtypes = [P, dp, C.POINTER(dp), C.POINTER(C.c_int), dp, dp]
        L.orc_wi.argtypes = [P, dp, dp]
        L.orc_contract_twopoint.argtypes = [P, dp, C.c_int, C.c_int, C.POINTER(dp), C.POINTER(dp), C.c_int,
                                            C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_plaquette.argtypes = [P, dp]
        L.orc_plaquette.restype = C.c_double
        L.orc_gamma_tables.argtypes = [C.POINTER((C.c_int * 24) * 16), C.POINTER((C.c_int * 24) * 16)]
        L.orc_seq_source.argtypes = [P, dp, dp, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_double]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def gamma_tables():
    perm = ((C.c_int * 24) * 16)()
    sign = ((C.c_int * 24) * 16)()
    lib().orc_gamma_tables(C.byref(perm), C.byref(sign))
    return np.array(perm, dtype=np.int64), np.array(sign, dtype=np.int64)


class Lattice:
    """Geometry, twist phases and gamma tables of one lattice (cvc_geometry.c)."""

    def __init__(self, dims, kappa, mu, bc=(1.0, 0.0, 0.0, 0.0)):
        self.dims = tuple(int(d) for d in dims)
        self.kappa, self.mu = float(kappa), float(mu)
        self.bc = tuple(float(b) for b in bc)
        self._s = _Lat()
        bca = (C.c_double * 4)(*self.bc)
        rc = lib().orc_lat_init(C.byref(self._s), *self.dims, self.kappa, self.mu, bca)
        if rc != 0:
            raise MemoryError("orc_lat_init failed")
        self.V = int(self._s.V)

    def __del__(self):
        try:
            lib().orc_lat_free(C.byref(self._s))
        except Exception:
            pass

    @property
    def phase(self):
        return np.array([[self._s.phase[m][0], self._s.phase[m][1]] for m in range(4)])

    def iup(self):
        return np.ctypeslib.as_array(self._s.iup, shape=(self.V, 4)).copy()

    def idn(self):
        return np.ctypeslib.as_array(self._s.idn, shape=(self.V, 4)).copy()

    # ---- operators -------------------------------------------------------
    def Q_phi(self, gauge, phi, mutm=None):
        """xi = D(mutm) phi (Q_phi.c:1092); mutm=None gives Q_phi_tbc (g_mu)."""
        out = np.empty((self.V, 24))
        lib().orc_Q_phi(C.byref(self._s), _dp(gauge), _dp(out), _dp(phi),
                        self.mu if mutm is None else float(mutm))
        return out

    def Qf5(self, gauge, phi, mutm):
        out = np.empty((self.V, 24))
        lib().orc_Qf5(C.byref(self._s), _dp(gauge), _dp(out), _dp(phi), float(mutm))
        return out

    def g5_phi(self, phi):
        out = np.ascontiguousarray(phi, dtype=np.float64).copy()
        lib().orc_g5_phi(C.byref(self._s), _dp(out))
        return out

    def dot_co(self, a, b):
        w = np.zeros(2)
        lib().orc_dot_co(C.byref(self._s), _dp(w), _dp(a), _dp(b))
        return complex(w[0], w[1])

    def dot_re(self, a, b):
        return lib().orc_dot_re(C.byref(self._s), _dp(a), _dp(b))

    def invert_Qtm(self, gauge, b, sp, niter_max=1000, x0=None, lagged=True):
        x = np.zeros((self.V, 24)) if x0 is None else np.array(x0, dtype=np.float64, order="C")
        tr = np.zeros(1)
        it = lib().orc_invert_Qtm(C.byref(self._s), _dp(gauge), _dp(x), _dp(b), float(sp),
                                  int(niter_max), int(bool(lagged)), _dp(tr))
        return it, x, float(tr[0])

    def invert_Qtm_her(self, gauge, b, sp, niter_max=1000, x0=None):
        x = np.zeros((self.V, 24)) if x0 is None else np.array(x0, dtype=np.float64, order="C")
        tr = np.zeros(1)
        it = lib().orc_invert_Qtm_her(C.byref(self._s), _dp(gauge), _dp(x), _dp(b), float(sp),
                                      int(niter_max), _dp(tr))
        return it, x, float(tr[0])

    def cvc_contract(self, gauge, props, src):
        """props: 120 arrays [V,24] (up 0..59, down 60..119); returns (conn[V,16,2], contact[4] complex)."""
        assert len(props) == 120
        arrs = [np.ascontiguousarray(p, dtype=np.float64) for p in props]
        ptrs = (C.POINTER(C.c_double) * 120)(*[_dp(a) for a in arrs])
        conn = np.empty((self.V, 16, 2))
        contact = np.zeros(8)
        s = (C.c_int * 4)(*[int(v) for v in src])
        lib().orc_cvc_contract(C.byref(self._s), _dp(gauge), ptrs, s, _dp(conn), _dp(contact))
        return conn, contact[0::2] + 1j * contact[1::2]

    def wi(self, conn):
        out = np.empty((self.V, 4, 2))
        lib().orc_wi(C.byref(self._s), _dp(np.ascontiguousarray(conn)), _dp(out))
        return out

    def contract_twopoint(self, idsource, idsink, chi, phi, n_c=3, snk_mom=None, src=(0, 0, 0, 0)):
        """contract_twopoint(_snk_momentum) (cvc_utils.c:3075-3235): complex C[T]."""
        assert len(chi) == 4 * n_c and len(phi) == 4 * n_c
        ca = [np.ascontiguousarray(a, dtype=np.float64) for a in chi]
        pa = [np.ascontiguousarray(a, dtype=np.float64) for a in phi]
        cp = (C.POINTER(C.c_double) * len(ca))(*[_dp(a) for a in ca])
        pp = (C.POINTER(C.c_double) * len(pa))(*[_dp(a) for a in pa])
        out = np.zeros((self.dims[0], 2))
        mom = (C.c_int * 3)(*snk_mom) if snk_mom is not None else None
        lib().orc_contract_twopoint(C.byref(self._s), _dp(out), int(idsource), int(idsink), cp, pp, int(n_c), mom,
                                    (C.c_int * 4)(*src))
        return out[:, 0] + 1j * out[:, 1]

    def seq_source(self, prop, tseq, gid, mom=(0, 0, 0), sgn=1.0):
        """Sequential source of p2gg_xspace.c:517-548 from one propagator column [V,24]."""
        out = np.empty((self.V, 24))
        lib().orc_seq_source(C.byref(self._s), _dp(np.ascontiguousarray(prop, dtype=np.float64)), _dp(out),
                             int(tseq), int(gid), (C.c_int * 3)(*[int(m) for m in mom]), C.c_double(sgn))
        return out

    def set_threads(self, n):
        """Threads of the operator loop (the reference's OpenMP Q_phi_tbc, Q_phi.c:190-206)."""
        lib().orc_set_threads(C.byref(self._s), int(n))

    def plaquette(self, gauge):
        return lib().orc_plaquette(C.byref(self._s), _dp(gauge))
