"""ctypes wrapper of the C restatement (oracle/cyx_oracle.c) -- test
infrastructure and CPU baseline only.  Build: ``make -C oracle``."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "libcyx_oracle.so"


class Grid(C.Structure):
    _fields_ = [("nr", C.c_int), ("nt", C.c_int), ("nz", C.c_int), ("fdn", C.c_int), ("cell_typ", C.c_int),
                ("xin", C.c_double), ("dr", C.c_double), ("dt", C.c_double), ("dz", C.c_double),
                ("twist", C.c_double)]


class Proj(C.Structure):
    _fields_ = [("n_atom", C.c_int), ("n_image", C.c_int), ("angnum", C.c_void_p), ("gamma", C.c_void_p),
                ("image_atom", C.c_void_p), ("image_nnodes", C.c_void_p), ("image_disp", C.c_void_p),
                ("rc_pos", C.c_void_p), ("chi", C.c_void_p)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        _lib = C.CDLL(str(LIB))
    return _lib


class COracle:
    """H-apply, Chebyshev filter and Jacobi-AAR of one channel in plain C."""

    def __init__(self, g, kpt, veff, W, atoms=()):
        self.lib = load()
        self.g = Grid(g.nr, g.nt, g.nz, g.fdn, g.cell_typ, g.xin, g.dr, g.dt, g.dz, g.twist)
        self.N = g.nr * g.nt * g.nz
        self.kpt = np.ascontiguousarray(kpt, dtype=float)
        self.veff = np.ascontiguousarray(veff, dtype=float)
        self.W = np.ascontiguousarray(W, dtype=float)
        ang, gam, ia, nn, disp, pos, chi = [], [], [], [], [], [], []
        for J, a in enumerate(atoms):
            ang.append(a.gamma.size)
            gam.extend(a.gamma.tolist())
            for im in a.images:
                ia.append(J)
                nn.append(im.rc_pos.size)
                disp.extend((a.coords - im.coords).tolist())
                pos.append(im.rc_pos.astype(np.int32))
                chi.append(np.asarray(im.chi, float).ravel(order="F"))
        self._keep = [np.array(ang, np.int32), np.array(gam, float), np.array(ia, np.int32),
                      np.array(nn, np.int32), np.array(disp, float),
                      np.concatenate(pos) if pos else np.zeros(0, np.int32),
                      np.concatenate(chi) if chi else np.zeros(0)]
        k = self._keep
        self.proj = Proj(len(ang), len(ia), *(a.ctypes.data for a in k))

    def apply(self, X):
        X = np.asfortranarray(X, dtype=complex)
        out = np.zeros_like(X, order="F")
        fn = self.lib.cyxo_h_apply
        fn.argtypes = [C.c_void_p] * 7 + [C.c_int]
        rc = fn(C.addressof(self.g), self.kpt.ctypes.data, self.veff.ctypes.data, self.W.ctypes.data,
                C.addressof(self.proj), X.ctypes.data, out.ctypes.data, X.shape[1])
        assert rc == 0
        return out

    def chebyshev_filter(self, X, m, a, b, a0):
        Xw = np.array(X, dtype=complex, order="F", copy=True)
        Y = np.zeros_like(Xw, order="F")
        fn = self.lib.cyxo_chebyshev_filter
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        rc = fn(C.addressof(self.g), self.kpt.ctypes.data, self.veff.ctypes.data, self.W.ctypes.data,
                C.addressof(self.proj), Xw.ctypes.data, Y.ctypes.data, Xw.shape[1], int(m), a, b, a0)
        assert rc == 0
        return Y

    def aar_jacobi(self, f, x0=None, tol=1e-8, max_iter=1000, omega=0.6, beta=0.6, m=7, p=6):
        f = np.ascontiguousarray(f, dtype=float)
        x = np.ascontiguousarray(x0 if x0 is not None else np.zeros(self.N), dtype=float).copy()
        rel = C.c_double()
        fn = self.lib.cyxo_aar_jacobi
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                       C.c_int, C.c_int, C.POINTER(C.c_double)]
        it = fn(C.addressof(self.g), f.ctypes.data, x.ctypes.data, 1 if x0 is not None else 0, tol, max_iter,
                omega, beta, m, p, C.byref(rel))
        return x, rel.value, it
