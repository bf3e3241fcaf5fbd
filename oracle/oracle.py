"""ctypes binding of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference used as the
checker by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py.  The product (ppm_legacy_amd) never imports this module.

Array conventions (C order, same bytes as the reference's Fortran arrays):
  xp     (np, dim)          -- xp(dim, Np)
  up     (np, lda)          -- up(lda, Np)
  field  (ez, ey, ex, lda)  -- field_up(lda, 1-g:Nm+g, 1-g:Nm+g, 1-g:Nm+g) (3D)
         (ey, ex, lda)      -- 2D
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KERNEL_BSP2 = 1
KERNEL_MP4 = 2
ERR_PART_UNASS = 405
ERR_ARGUMENT = 503
ERR_BOUNDS = 9001


class OrTopo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("Nm", ctypes.c_int * 3),
                ("min_phys", ctypes.c_double * 3), ("max_phys", ctypes.c_double * 3),
                ("bcdef", ctypes.c_int * 6)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def topo(dim, Nm, min_phys=None, max_phys=None, bcdef=None):
    """Single-sub topology descriptor; bcdef defaults to periodic on all faces."""
    t = OrTopo()
    t.dim = dim
    Nm = list(Nm) + [1] * (3 - len(Nm))
    mn = list(min_phys if min_phys is not None else [0.0] * dim) + [0.0] * (3 - dim)
    mx = list(max_phys if max_phys is not None else [1.0] * dim) + [1.0] * (3 - dim)
    bc = list(bcdef if bcdef is not None else [1] * (2 * dim)) + [1] * (6 - 2 * dim)
    for d in range(3):
        t.Nm[d] = int(Nm[d])
        t.min_phys[d] = float(mn[d])
        t.max_phys[d] = float(mx[d])
    for d in range(6):
        t.bcdef[d] = int(bc[d])
    return t


def field_shape(t, lda, ghost):
    dims = [t.Nm[d] + 2 * ghost[d] for d in range(t.dim)]
    return tuple(reversed(dims)) + (lda,)


def _sfx(dtype):
    return "_d" if np.dtype(dtype) == np.float64 else "_s"


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _ghost(g, dim):
    return (ctypes.c_int * 3)(*(list(g) + [0] * (3 - len(g))))


def p2m(t, kernel, xp, up, ghost, fold=True, dtype=np.float64):
    xp = np.ascontiguousarray(xp, dtype=dtype)
    up = np.ascontiguousarray(up, dtype=dtype)
    if up.ndim == 1:
        up = up[:, None]
    np_, lda = up.shape
    field = np.zeros(field_shape(t, lda, ghost), dtype=dtype)
    fn = getattr(lib(), "or_p2m" + _sfx(dtype))
    info = fn(ctypes.byref(t), kernel, np_, _ptr(xp), lda, _ptr(up), _ghost(ghost, t.dim),
              _ptr(field), int(fold))
    return info, field


def p2m_par(t, kernel, xp, up, ghost, nthreads):
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    up = np.ascontiguousarray(up, dtype=np.float64)
    if up.ndim == 1:
        up = up[:, None]
    np_, lda = up.shape
    field = np.zeros(field_shape(t, lda, ghost), dtype=np.float64)
    info = lib().or_p2m_par_d(ctypes.byref(t), kernel, np_, _ptr(xp), lda, _ptr(up),
                              _ghost(ghost, t.dim), _ptr(field), int(nthreads))
    return info, field


def m2p(t, kernel, xp, field, ghost, ghost_fill=True, dtype=np.float64):
    xp = np.ascontiguousarray(xp, dtype=dtype)
    field = np.ascontiguousarray(field, dtype=dtype)
    lda = field.shape[-1]
    np_ = xp.shape[0]
    up = np.zeros((np_, lda), dtype=dtype)
    fn = getattr(lib(), "or_m2p" + _sfx(dtype))
    info = fn(ctypes.byref(t), kernel, np_, _ptr(xp), lda, _ptr(field), _ghost(ghost, t.dim),
              _ptr(up), int(ghost_fill))
    return info, up


def m2p_par(t, kernel, xp, field, ghost, nthreads):
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    field = np.ascontiguousarray(field, dtype=np.float64)
    lda = field.shape[-1]
    up = np.zeros((xp.shape[0], lda), dtype=np.float64)
    info = lib().or_m2p_par_d(ctypes.byref(t), kernel, xp.shape[0], _ptr(xp), lda, _ptr(field),
                              _ghost(ghost, t.dim), _ptr(up), int(nthreads))
    return info, up


def ghost_put(t, field, ghost):
    field = np.ascontiguousarray(field).copy()
    fn = getattr(lib(), "or_ghost_put" + _sfx(field.dtype))
    info = fn(ctypes.byref(t), field.shape[-1], _ghost(ghost, t.dim), _ptr(field))
    return info, field


def ghost_get(t, field, ghost):
    field = np.ascontiguousarray(field).copy()
    fn = getattr(lib(), "or_ghost_get" + _sfx(field.dtype))
    info = fn(ctypes.byref(t), field.shape[-1], _ghost(ghost, t.dim), _ptr(field))
    return info, field


# element kinds of ghost_kind (OR_KIND_* of ppm_oracle.h)
KIND_D, KIND_S, KIND_I, KIND_L, KIND_DC, KIND_SC = 0, 2, 3, 4, 5, 6
_KIND_DTYPE = {KIND_D: np.float64, KIND_S: np.float32, KIND_I: np.int32, KIND_L: np.int32,
               KIND_DC: np.complex128, KIND_SC: np.complex64}


def ghost_kind(t, field, ghost, kind, get, mask=None):
    """ppm_map_field_ghost (ghost_get if get else ghost_put) for element kind
    KIND_*; field (field_shape, the kind's dtype; logicals as int32), mask
    None or one int32 logical per stored node (field.shape[:-1])."""
    field = np.ascontiguousarray(field, dtype=_KIND_DTYPE[kind]).copy()
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        assert mask.shape == field.shape[:-1]
    info = lib().or_ghost_kind(ctypes.byref(t), int(kind), field.shape[-1], _ghost(ghost, t.dim), _ptr(field),
                               _ptr(mask), int(bool(get)))
    return info, field


def rank(xp, xmin, xmax, Nm, Ngl, dtype=np.float64):
    """ppm_util_rank{2,3}d: returns (lpdx[:icount] 1-based, lhbx 1-based, icount, nout, ncorr)."""
    xp = np.ascontiguousarray(xp, dtype=dtype)
    dim = xp.shape[1]
    np_ = xp.shape[0]
    if dim == 3:
        nbox = (Nm[0] + Ngl[0] + Ngl[3]) * (Nm[1] + Ngl[1] + Ngl[4]) * (Nm[2] + Ngl[2] + Ngl[5])
    else:
        nbox = (Nm[0] + Ngl[0] + Ngl[2]) * (Nm[1] + Ngl[1] + Ngl[3])
    lpdx = np.zeros(max(np_, 1), dtype=np.int32)
    lhbx = np.zeros(nbox + 1, dtype=np.int32)
    xmin = np.ascontiguousarray(xmin, dtype=dtype)
    xmax = np.ascontiguousarray(xmax, dtype=dtype)
    nm = (ctypes.c_int * 3)(*list(Nm))
    ngl = (ctypes.c_int * 6)(*list(Ngl))
    ic, no, nc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    fn = getattr(lib(), ("or_rank3d" if dim == 3 else "or_rank2d") + _sfx(dtype))
    info = fn(np_, _ptr(xp), _ptr(xmin), _ptr(xmax), nm, ngl, _ptr(lpdx), _ptr(lhbx),
              ctypes.byref(ic), ctypes.byref(no), ctypes.byref(nc))
    assert info == 0
    return lpdx[:ic.value].copy(), lhbx, ic.value, no.value, nc.value


def impose_part_bc(t, xp, dtype=np.float64):
    xp = np.ascontiguousarray(xp, dtype=dtype).copy()
    fn = getattr(lib(), "or_impose_part_bc" + _sfx(dtype))
    fn(ctypes.byref(t), xp.shape[0], _ptr(xp))
    return xp


def comp_weights(t, kernel, xp, dtype=np.float64):
    xp = np.ascontiguousarray(xp, dtype=dtype)
    ks = 2 if kernel == KERNEL_MP4 else 1
    np_ = xp.shape[0]
    w = [np.zeros((np_, 2 * ks), dtype=dtype) for _ in range(3)]
    fn = getattr(lib(), "or_comp_weights" + _sfx(dtype))
    info = fn(ctypes.byref(t), kernel, np_, _ptr(xp), _ptr(w[0]), _ptr(w[1]), _ptr(w[2]))
    return info, w[: t.dim]


def remesh(t, kernel, xp, up, ghost, wx, dtype=np.float64):
    xp = np.ascontiguousarray(xp, dtype=dtype)
    up = np.ascontiguousarray(up, dtype=dtype)
    if up.ndim == 1:
        up = up[:, None]
    np_, lda = up.shape
    wx = [np.ascontiguousarray(w, dtype=dtype) for w in wx] + [None] * (3 - len(wx))
    field = np.zeros(field_shape(t, lda, ghost), dtype=dtype)
    fn = getattr(lib(), "or_remesh" + _sfx(dtype))
    info = fn(ctypes.byref(t), kernel, np_, _ptr(xp), lda, _ptr(up), _ghost(ghost, t.dim),
              _ptr(wx[0]), _ptr(wx[1]), _ptr(wx[2]), _ptr(field))
    return info, field


def create_part(t, field, ghost, cutoff, cutoff_weights=None, resetpos=True):
    field = np.ascontiguousarray(field)
    dtype = field.dtype
    lda = field.shape[-1]
    maxnp = int(np.prod([t.Nm[d] for d in range(t.dim)]))
    xp = np.zeros((maxnp, t.dim), dtype=dtype)
    up = np.zeros((maxnp, lda), dtype=dtype)
    cut = np.asarray(cutoff, dtype=dtype)
    cw = None if cutoff_weights is None else np.ascontiguousarray(cutoff_weights, dtype=dtype)
    npo = ctypes.c_int()
    fn = getattr(lib(), "or_create_part" + _sfx(dtype))
    info = fn(ctypes.byref(t), lda, _ptr(field), _ghost(ghost, t.dim), _ptr(cut), _ptr(cw),
              int(resetpos), maxnp, ctypes.byref(npo), _ptr(xp), _ptr(up))
    n = npo.value
    return info, xp[:n].copy(), up[:n].copy()


def p2m_adjoint(t, kernel, xp, up, g, dtype=np.float64):
    """g: compact cotangent on the unique nodes, shape (N_z, N_y, N_x, lda) (3D)."""
    xp = np.ascontiguousarray(xp, dtype=dtype)
    up = np.ascontiguousarray(up, dtype=dtype)
    if up.ndim == 1:
        up = up[:, None]
    g = np.ascontiguousarray(g, dtype=dtype)
    np_, lda = up.shape
    gu = np.zeros((np_, lda), dtype=dtype)
    gx = np.zeros((np_, t.dim), dtype=dtype)
    fn = getattr(lib(), "or_p2m_adjoint" + _sfx(dtype))
    info = fn(ctypes.byref(t), kernel, np_, _ptr(xp), lda, _ptr(up), _ptr(g), _ptr(gu), _ptr(gx))
    return info, gu, gx


def real_nodes(field, ghost, dim):
    """Slice the real nodes 1..Nm (Fortran numbering) out of a padded field."""
    sl = tuple(slice(ghost[d], field.shape[dim - 1 - d] - ghost[d]) for d in reversed(range(dim)))
    return field[sl]


def unique_nodes(field, ghost, dim):
    """Real nodes 1..Nm-1 (drops the duplicate periodic node Nm)."""
    sl = tuple(slice(ghost[d], field.shape[dim - 1 - d] - ghost[d] - 1) for d in reversed(range(dim)))
    return field[sl]


def _setup():
    L = lib()
    for name in ["or_p2m_d", "or_p2m_s", "or_m2p_d", "or_m2p_s", "or_p2m_par_d", "or_m2p_par_d",
                 "or_ghost_put_d", "or_ghost_put_s", "or_ghost_get_d", "or_ghost_get_s",
                 "or_rank3d_d", "or_rank3d_s", "or_rank2d_d", "or_rank2d_s",
                 "or_impose_part_bc_d", "or_impose_part_bc_s", "or_comp_weights_d",
                 "or_comp_weights_s", "or_remesh_d", "or_remesh_s", "or_create_part_d",
                 "or_create_part_s", "or_p2m_adjoint_d", "or_p2m_adjoint_s"]:
        getattr(L, name).restype = ctypes.c_int
