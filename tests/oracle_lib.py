"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker for the HIP path and is
never imported by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

QUAD, TRI, HEX, TET = 0, 1, 2, 3
STENCIL_ELEMENT, STENCIL_INTERSECTION, STENCIL_ELEMENT_AND_INTERSECTION = 0, 1, 2

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def _build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        _build()
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "or_last_error": (ctypes.c_char_p, []),
        "or_mesh_structured": (_vp, [ctypes.c_int, _ip, _dp, _dp]),
        "or_mesh_create": (_vp, [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_int64, _ip]),
        "or_mesh_mixed": (_vp, [ctypes.c_int64, _dp, ctypes.c_int64, _ip, _lp, _ip]),
        "or_mesh_is_mixed": (ctypes.c_int, [_vp]),
        "or_mesh_quadrature_offsets": (ctypes.c_int, [_vp, ctypes.c_int, _lp]),
        "or_mesh_destroy": (None, [_vp]),
        "or_mesh_num_vertices": (ctypes.c_int64, [_vp]),
        "or_mesh_num_elements": (ctypes.c_int64, [_vp]),
        "or_mesh_dim": (ctypes.c_int, [_vp]),
        "or_mesh_verts_per_element": (ctypes.c_int, [_vp]),
        "or_mesh_coords": (None, [_vp, _dp]),
        "or_mesh_conn": (None, [_vp, _ip]),
        "or_mesh_faces_per_element": (ctypes.c_int, [_vp]),
        "or_mesh_neighbors": (None, [_vp, _ip]),
        "or_space_create": (_vp, [_vp, ctypes.c_int, ctypes.c_int]),
        "or_space_destroy": (None, [_vp]),
        "or_space_size": (ctypes.c_int64, [_vp]),
        "or_space_local_size": (ctypes.c_int, [_vp]),
        "or_space_global_indices": (None, [_vp, _ip]),
        "or_operator_create": (_vp, [_vp, ctypes.c_int]),
        "or_operator_create_between": (_vp, [_vp, _vp, ctypes.c_int]),
        "or_operator_cols": (ctypes.c_int64, [_vp]),
        "or_operator_destroy": (None, [_vp]),
        "or_operator_nnz": (ctypes.c_int64, [_vp]),
        "or_operator_pattern": (None, [_vp, _lp, _ip]),
        "or_operator_append_laplace": (ctypes.c_int, [_vp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
        "or_operator_append_product": (ctypes.c_int, [_vp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
        "or_operator_append_inner_coupling": (
            ctypes.c_int,
            [_vp, ctypes.c_double, _dp, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double],
        ),
        "or_operator_append_inner_penalty": (
            ctypes.c_int,
            [_vp, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double],
        ),
        "or_operator_append_dirichlet_coupling": (
            ctypes.c_int,
            [_vp, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double],
        ),
        "or_operator_append_boundary_penalty": (
            ctypes.c_int,
            [_vp, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double],
        ),
        "or_operator_append_source": (
            ctypes.c_int,
            [_vp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int],
        ),
        "or_operator_set_dirichlet": (ctypes.c_int, [_vp, ctypes.c_int]),
        "or_operator_set_boundary_source_tensor": (ctypes.c_int, [_vp, _dp]),
        "or_operator_set_face_tensors": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
        "or_operator_set_dirichlet_mode": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
        "or_operator_assemble": (ctypes.c_int, [_vp, ctypes.c_int]),
        "or_operator_values": (None, [_vp, _dp]),
        "or_operator_rhs": (None, [_vp, _dp]),
        "or_operator_num_dirichlet": (ctypes.c_int64, [_vp]),
        "or_operator_dirichlet_dofs": (None, [_vp, _ip]),
        "or_spmv": (None, [ctypes.c_int64, _lp, _ip, _dp, _dp, _dp, ctypes.c_int]),
        "or_pcg_jacobi": (
            ctypes.c_int,
            [ctypes.c_int64, _lp, _ip, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int64, ctypes.c_int,
             ctypes.c_int64, _lp, _dp],
        ),
        "or_laplace_evaluate": (None, [ctypes.c_int] * 4 + [_dp, _dp, _dp, _dp]),
        "or_product_evaluate": (None, [ctypes.c_int] * 3 + [_dp, _dp, _dp, _dp]),
        "or_quadrature": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]),
        "or_esv2007_h1_semi_error": (ctypes.c_double, [_vp, _dp, ctypes.c_int]),
        "or_adjoint_log_kappa_gradient": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
        "or_error_norms": (ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, _dp]),
        "or_mesh_quadrature_points": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "or_operator_append_laplace_tensor": (ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "or_operator_append_laplace_points": (
            ctypes.c_int, [_vp, _dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "or_operator_append_product_points": (
            ctypes.c_int, [_vp, _dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "or_operator_append_boundary_source": (
            ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_double, _dp, ctypes.c_double,
                           ctypes.c_int, ctypes.c_int]),
        "or_mesh_boundary_quadrature_points": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
        "or_operator_set_boundary_filter": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int64]),
        "or_operator_set_dirichlet_faces": (ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_int64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _err():
    return lib().or_last_error().decode()


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def lptr(a):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_lp)


class Mesh:
    def __init__(self, handle):
        if not handle:
            raise RuntimeError(_err())
        self.h = handle
        L = lib()
        self.nv = L.or_mesh_num_vertices(handle)
        self.ne = L.or_mesh_num_elements(handle)
        self.dim = L.or_mesh_dim(handle)
        self.k = L.or_mesh_verts_per_element(handle)

    @classmethod
    def structured(cls, geom, cells, lo, hi):
        c = np.asarray(cells, np.int32)
        l = np.asarray(lo, np.float64)
        h = np.asarray(hi, np.float64)
        return cls(lib().or_mesh_structured(geom, iptr(c), dptr(l), dptr(h)))

    @classmethod
    def explicit(cls, geom, coords, conn):
        coords = np.ascontiguousarray(coords, np.float64)
        conn = np.ascontiguousarray(conn, np.int32)
        return cls(lib().or_mesh_create(geom, coords.shape[0], dptr(coords), conn.shape[0], iptr(conn)))

    @classmethod
    def mixed(cls, coords, elements):
        """triangles (3 vertices) and quadrilaterals (4) in one mesh, element by element"""
        coords = np.ascontiguousarray(coords, np.float64)
        sizes = np.array([len(v) for v in elements], np.int64)
        geometry = np.where(sizes == 4, QUAD, TRI).astype(np.int32)
        offsets = np.zeros(len(elements) + 1, np.int64)
        offsets[1:] = np.cumsum(sizes)
        verts = np.ascontiguousarray(np.concatenate([np.asarray(v, np.int32) for v in elements]), np.int32)
        return cls(lib().or_mesh_mixed(coords.shape[0], dptr(coords), len(elements), iptr(geometry), lptr(offsets),
                                       iptr(verts)))

    @property
    def is_mixed(self):
        return bool(lib().or_mesh_is_mixed(self.h))

    def quadrature_points(self, order):
        """physical coordinates of the element-rule points, [ne, npts, dim] ([P, dim] on a mixed mesh)"""
        n = lib().or_mesh_quadrature_points(self.h, order, None)
        if n < 0:
            raise RuntimeError(_err())
        if self.is_mixed:
            out = np.empty((int(self.quadrature_offsets(order)[-1]), self.dim))
        else:
            out = np.empty((self.ne, n, self.dim))
        lib().or_mesh_quadrature_points(self.h, order, dptr(out))
        return out

    def quadrature_offsets(self, order):
        out = np.empty(self.ne + 1, np.int64)
        if lib().or_mesh_quadrature_offsets(self.h, order, lptr(out)) != 0:
            raise RuntimeError(_err())
        return out

    def boundary_quadrature_points(self, order):
        """physical coordinates of the face-rule points on the boundary intersections (walk order), [nb, npts, dim]"""
        n = lib().or_mesh_boundary_quadrature_points(self.h, order, None)
        if n < 0:
            raise RuntimeError(_err())
        nb = int(np.sum(self.neighbors() == -1))
        out = np.empty((nb, n, self.dim))
        lib().or_mesh_boundary_quadrature_points(self.h, order, dptr(out))
        return out

    def coords(self):
        out = np.empty((self.nv, self.dim))
        lib().or_mesh_coords(self.h, dptr(out))
        return out

    def conn(self):
        out = np.empty((self.ne, self.k), np.int32)
        lib().or_mesh_conn(self.h, iptr(out))
        return out

    def neighbors(self):
        nf = lib().or_mesh_faces_per_element(self.h)
        out = np.empty((self.ne, nf), np.int32)
        lib().or_mesh_neighbors(self.h, iptr(out))
        return out

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib().or_mesh_destroy(self.h)


class Space:
    def __init__(self, mesh, continuous=True, order=1):
        self.mesh = mesh
        self.h = lib().or_space_create(mesh.h, int(continuous), order)
        if not self.h:
            raise RuntimeError(_err())
        self.size = lib().or_space_size(self.h)
        self.local_size = lib().or_space_local_size(self.h)

    def global_indices(self):
        out = np.empty((self.mesh.ne, self.local_size), np.int32)
        lib().or_space_global_indices(self.h, iptr(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_space_destroy(self.h)


class Operator:
    """MatrixOperator restated on the CPU: pattern, appended local forms, functional, Dirichlet."""

    def __init__(self, space, stencil=STENCIL_ELEMENT, source=None):
        """rows of `space` (range: test functions), columns of `source` (ansatz functions; `space` if None)"""
        self.space = space
        self.source = space if source is None else source
        if source is None:
            self.h = lib().or_operator_create(space.h, stencil)
        else:
            self.h = lib().or_operator_create_between(source.h, space.h, stencil)
        if not self.h:
            raise RuntimeError(_err())
        self.n = space.size
        self.ncols = lib().or_operator_cols(self.h)
        self.nnz = lib().or_operator_nnz(self.h)
        self._keep = []

    def _coef(self, c):
        if c is None:
            return None, 1.0
        if np.isscalar(c):
            return None, float(c)
        a = np.ascontiguousarray(c, np.float64)
        assert a.shape == (self.space.mesh.ne,)
        self._keep.append(a)
        return dptr(a), 0.0

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(_err())

    def append_laplace(self, kappa=1.0, over_integrate=0, scaling=1.0):
        p, c = self._coef(kappa)
        self._check(lib().or_operator_append_laplace(self.h, p, c, over_integrate, scaling))

    def append_product(self, weight=1.0, over_integrate=0, scaling=1.0):
        p, c = self._coef(weight)
        self._check(lib().or_operator_append_product(self.h, p, c, over_integrate, scaling))

    def append_laplace_tensor(self, tensor, over_integrate=0, scaling=1.0):
        """Laplace form with a matrix-valued diffusion: [d, d] for all elements or [ne, d, d]"""
        t = np.ascontiguousarray(tensor, np.float64)
        self._keep = getattr(self, "_keep", []) + [t]
        self._check(lib().or_operator_append_laplace_tensor(self.h, dptr(t), int(t.ndim == 3), over_integrate,
                                                             scaling))

    def append_laplace_points(self, values, coeff_order, over_integrate=0, scaling=1.0):
        """Laplace form with the diffusion at the points of the element rules of order
        coeff_order + 2 k + over_integrate (Mesh.quadrature_points layout)"""
        v = np.ascontiguousarray(values, np.float64).reshape(-1)
        self._keep = getattr(self, "_keep", []) + [v]
        self._check(lib().or_operator_append_laplace_points(self.h, dptr(v), v.size, coeff_order, over_integrate,
                                                             scaling))

    def append_product_points(self, values, coeff_order, over_integrate=0, scaling=1.0):
        v = np.ascontiguousarray(values, np.float64).reshape(-1)
        self._keep = getattr(self, "_keep", []) + [v]
        self._check(lib().or_operator_append_product_points(self.h, dptr(v), v.size, coeff_order, over_integrate,
                                                             scaling))

    def append_inner_coupling(self, symmetry_prefactor, kappa=1.0, omega=1.0, over_integrate=0, scaling=1.0):
        kp, kc = self._coef(kappa)
        op, oc = self._coef(omega)
        self._check(lib().or_operator_append_inner_coupling(self.h, symmetry_prefactor, kp, kc, op, oc,
                                                              over_integrate, scaling))

    def append_inner_penalty(self, penalty, omega=1.0, over_integrate=0, scaling=1.0):
        op, oc = self._coef(omega)
        self._check(lib().or_operator_append_inner_penalty(self.h, penalty, op, oc, over_integrate, scaling))

    def append_dirichlet_coupling(self, symmetry_prefactor, kappa=1.0, over_integrate=0, scaling=1.0):
        kp, kc = self._coef(kappa)
        self._check(lib().or_operator_append_dirichlet_coupling(self.h, symmetry_prefactor, kp, kc,
                                                                  over_integrate, scaling))

    def append_boundary_penalty(self, penalty, omega=1.0, over_integrate=0, scaling=1.0):
        op, oc = self._coef(omega)
        self._check(lib().or_operator_append_boundary_penalty(self.h, penalty, op, oc, over_integrate, scaling))

    def append_source(self, f=1.0, f_order=0, over_integrate=0, esv2007=False):
        p, c = self._coef(f)
        self._check(lib().or_operator_append_source(self.h, 1 if esv2007 else 0, p, c, f_order, over_integrate))

    BS_DIRICHLET_COUPLING, BS_BOUNDARY_PENALTY, BS_PRODUCT = 0, 1, 2

    def append_boundary_source(self, kind, param=0.0, coeff=1.0, g=0.0, g_order=0, over_integrate=0):
        """boundary functional term; g: constant or values at the face quadrature points [nb, npts]"""
        p, c = self._coef(coeff)
        if np.isscalar(g):
            gp, gc = None, float(g)
        else:
            self._keep_g = getattr(self, "_keep_g", []) + [np.ascontiguousarray(g, np.float64).reshape(-1)]
            gp, gc = dptr(self._keep_g[-1]), 0.0
        self._check(lib().or_operator_append_boundary_source(self.h, kind, param, p, c, gp, gc, g_order,
                                                             over_integrate))

    def _mask(self, fn, mask):
        if mask is None:
            self._check(fn(self.h, None, 0))
            return
        m = np.ascontiguousarray(mask, np.uint8)
        self._check(fn(self.h, m.ctypes.data_as(ctypes.c_void_p), len(m)))

    def set_boundary_filter(self, mask):
        """boundary intersections (walk order) of the boundary forms and functionals appended next; None: all"""
        self._mask(lib().or_operator_set_boundary_filter, mask)

    def set_dirichlet_faces(self, mask):
        """the Dirichlet intersections of the constraints (walk order); None: every boundary intersection"""
        self._mask(lib().or_operator_set_dirichlet_faces, mask)

    def set_dirichlet(self, enabled=True):
        self._check(lib().or_operator_set_dirichlet(self.h, int(enabled)))

    def set_boundary_source_tensor(self, tensor):
        t = np.ascontiguousarray(tensor, np.float64)
        self._keep = getattr(self, "_keep", []) + [t]
        self._check(lib().or_operator_set_boundary_source_tensor(self.h, dptr(t)))

    def set_face_tensors(self, boundary, kappa=None, omega=None):
        """matrix-valued diffusion / weight ([ne, d, d]) of the last appended inner (boundary False) or boundary
        intersection form"""
        kt = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        wt = None if omega is None else np.ascontiguousarray(omega, np.float64)
        self._keep = getattr(self, "_keep", []) + [kt, wt]
        self._check(lib().or_operator_set_face_tensors(self.h, int(bool(boundary)), None if kt is None else dptr(kt),
                                                        None if wt is None else dptr(wt)))

    def set_dirichlet_mode(self, only_clear=False, ensure_symmetry=True):
        self._check(lib().or_operator_set_dirichlet_mode(self.h, int(only_clear), int(ensure_symmetry)))

    def assemble(self, nthreads=1):
        self._check(lib().or_operator_assemble(self.h, nthreads))

    def pattern(self):
        rp = np.empty(self.n + 1, np.int64)
        cols = np.empty(self.nnz, np.int32)
        lib().or_operator_pattern(self.h, lptr(rp), iptr(cols))
        return rp, cols

    def values(self):
        out = np.empty(self.nnz)
        lib().or_operator_values(self.h, dptr(out))
        return out

    def rhs(self):
        out = np.empty(self.n)
        lib().or_operator_rhs(self.h, dptr(out))
        return out

    def dirichlet_dofs(self):
        m = lib().or_operator_num_dirichlet(self.h)
        out = np.empty(m, np.int32)
        lib().or_operator_dirichlet_dofs(self.h, iptr(out))
        return out

    def csr(self):
        import scipy.sparse as sp

        rp, cols = self.pattern()
        return sp.csr_matrix((self.values(), cols, rp), shape=(self.n, self.ncols))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_operator_destroy(self.h)


def spmv(row_ptr, cols, vals, x, nthreads=1):
    y = np.empty_like(x)
    lib().or_spmv(len(row_ptr) - 1, lptr(row_ptr), iptr(cols), dptr(vals), dptr(x), dptr(y), nthreads)
    return y


def pcg_jacobi(row_ptr, cols, vals, b, rtol=1e-10, max_iter=100000, nthreads=1, x0=None, max_iters_run=0):
    n = len(row_ptr) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    it = ctypes.c_int64(0)
    rel = ctypes.c_double(0)
    rc = lib().or_pcg_jacobi(n, lptr(row_ptr), iptr(cols), dptr(vals), dptr(np.ascontiguousarray(b)), dptr(x),
                             rtol, max_iter, nthreads, max_iters_run, ctypes.byref(it), ctypes.byref(rel))
    return x, it.value, rel.value, rc


def laplace_evaluate(weight, test_grads, ansatz_grads):
    """weight: (d,d); grads: (n, r, d)"""
    w = np.ascontiguousarray(weight, np.float64)
    tg = np.ascontiguousarray(test_grads, np.float64)
    ag = np.ascontiguousarray(ansatz_grads, np.float64)
    rows, r, d = tg.shape
    cols = ag.shape[0]
    out = np.empty((rows, cols))
    lib().or_laplace_evaluate(d, r, rows, cols, dptr(w), dptr(tg), dptr(ag), dptr(out))
    return out


def product_evaluate(weight, test_values, ansatz_values):
    """weight: (r,r); values: (n, r)"""
    w = np.ascontiguousarray(weight, np.float64)
    tv = np.ascontiguousarray(test_values, np.float64)
    av = np.ascontiguousarray(ansatz_values, np.float64)
    rows, r = tv.shape
    cols = av.shape[0]
    out = np.empty((rows, cols))
    lib().or_product_evaluate(r, rows, cols, dptr(w), dptr(tv), dptr(av), dptr(out))
    return out


def quadrature(geom, order):
    dim = 2 if geom in (QUAD, TRI) else 3
    cap = 4096
    pts = np.empty(cap * dim)
    w = np.empty(cap)
    n = lib().or_quadrature(geom, order, dptr(pts), dptr(w), cap)
    if n < 0 or n > cap:
        raise RuntimeError(_err())
    return pts[: n * dim].reshape(n, dim), w[:n]


def esv2007_h1_semi_error(space, u, quad_order=8):
    return lib().or_esv2007_h1_semi_error(space.h, dptr(np.ascontiguousarray(u, np.float64)), quad_order)


def error_norms(space, u, order, gval=None, ggrad=None, kappa=1.0):
    """(||u_h - g||_L2^2, sum_e int kappa |grad(u_h - g)|^2) with g given at the element-rule points"""
    out = np.empty(2)
    keep = [np.ascontiguousarray(a, np.float64) if a is not None else None for a in (gval, ggrad)]
    kap = None if np.isscalar(kappa) else np.ascontiguousarray(kappa, np.float64)
    rc = lib().or_error_norms(space.h, dptr(np.ascontiguousarray(u, np.float64)), order,
                              dptr(keep[0]) if keep[0] is not None else None,
                              dptr(keep[1]) if keep[1] is not None else None,
                              dptr(kap) if kap is not None else None, float(kappa) if kap is None else 0.0, dptr(out))
    if rc != 0:
        raise RuntimeError(_err())
    return out


def adjoint_log_kappa_gradient(space, kappa, u, lam):
    g = np.empty(space.mesh.ne)
    rc = lib().or_adjoint_log_kappa_gradient(space.h, dptr(np.ascontiguousarray(kappa)), dptr(np.ascontiguousarray(u)),
                                             dptr(np.ascontiguousarray(lam)), dptr(g))
    if rc != 0:
        raise RuntimeError(_err())
    return g
