"""Core (re-entrant) interface over libppmg with torch device tensors.

Layouts are the reference's Fortran layouts seen from C order:
  xp     (np, dim)            xp(dim, Np)
  up     (np, lda)            up(lda, Np)
  field  (ez, ey, ex, lda)    field_up(lda, i, j, k) over a stored node box
         (ey, ex, lda)        2D
A stored box is described by its first node index per axis, ``lo`` (x, y[, z];
0-based, node 0 at min_phys): PPM's padded field_up is lo = -ghostsize, the
compact periodic layout (unique nodes only) is lo = 0 with Nm-1 nodes.

Precision forms: PREC_F64 (double everything), PREC_F32C (double positions,
FP64 index/offset, float weights and values with compensated sums),
PREC_F32 (float everything; the reference's single specifics).
"""
import ctypes

import torch

from . import _lib as L
from ._lib import KERNEL_BSP2, KERNEL_MP4, PREC_F32, PREC_F32C, PREC_F64, check, lib

POS_DTYPE = {PREC_F64: torch.float64, PREC_F32C: torch.float64, PREC_F32: torch.float32}
VAL_DTYPE = {PREC_F64: torch.float64, PREC_F32C: torch.float32, PREC_F32: torch.float32}


class Mesh:
    """One sub covering [min_phys, max_phys] with nm nodes per axis (PPM's Nm)."""

    def __init__(self, dim, nm, min_phys=None, max_phys=None, bcdef=None):
        self.dim = int(dim)
        nm = [int(nm)] * self.dim if isinstance(nm, int) else [int(v) for v in nm]
        self.nm = nm
        self.min_phys = [float(v) for v in (min_phys if min_phys is not None else [0.0] * self.dim)]
        self.max_phys = [float(v) for v in (max_phys if max_phys is not None else [1.0] * self.dim)]
        self.bcdef = [int(v) for v in (bcdef if bcdef is not None else [1] * (2 * self.dim))]
        c = L.PpmgMesh()
        c.dim = self.dim
        for d in range(3):
            c.nm[d] = nm[d] if d < self.dim else 2
            c.min_phys[d] = self.min_phys[d] if d < self.dim else 0.0
            c.max_phys[d] = self.max_phys[d] if d < self.dim else 1.0
        for f in range(6):
            c.bcdef[f] = self.bcdef[f] if f < 2 * self.dim else 1
        self._c = c

    @property
    def cells(self):
        return [n - 1 for n in self.nm]

    @property
    def periodic(self):
        return all(b == 1 for b in self.bcdef[: 2 * self.dim])

    def ref(self):
        return ctypes.byref(self._c)

    def compact_shape(self, lda):
        return tuple(reversed(self.cells)) + (lda,)

    def padded_shape(self, lda, ghost):
        return tuple(reversed([self.nm[d] + 2 * ghost[d] for d in range(self.dim)])) + (lda,)


class Plan:
    """Device scratch + the forward sort kept for the adjoint."""

    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().ppmg_plan_create(ctypes.byref(p)))
        self._p = p

    @property
    def handle(self):
        return self._p

    def set_timing(self, on=True):
        check(lib().ppmg_plan_set_timing(self._p, int(on)))

    def timings(self):
        """Phase durations (ms) of the last call on this plan (HIP events on its stream)."""
        arr = (ctypes.c_double * 8)()
        n = lib().ppmg_plan_timings(self._p, arr, 8)
        if n < 0:
            check(n)
        return list(arr)[:n]

    def sort_view(self):
        """(key, perm, start) of the last binning, as int32 device tensors (copies)."""
        k, p, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        n, nc = ctypes.c_int64(), ctypes.c_int64()
        check(lib().ppmg_plan_sort_view(self._p, ctypes.byref(k), ctypes.byref(p), ctypes.byref(s),
                                        ctypes.byref(n), ctypes.byref(nc)))
        return (_wrap_i32(k.value, n.value), _wrap_i32(p.value, n.value), _wrap_i32(s.value, nc.value + 1))

    def close(self):
        if self._p is not None and self._p.value:
            lib().ppmg_plan_destroy(self._p)
        self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _wrap_i32(addr, n):
    """Copy n int32 from a device address into a new tensor."""
    out = torch.empty(max(n, 0), dtype=torch.int32, device="cuda")
    if n > 0:
        from . import _lib
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.restype = ctypes.c_int
        rc = hip.hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(addr), ctypes.c_size_t(4 * n), 3)
        if rc != 0:
            raise _lib.PpmgError(-1, _lib.ERR_DEVICE, f"hipMemcpy failed ({rc})")
    return out


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_tensor(t, dtype, name, cols=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"{name} must have shape (n, {cols}), got {tuple(t.shape)}")


def field_desc(mesh, field, lo):
    """ppmg_field for a (ez, ey, ex, lda) / (ey, ex, lda) tensor whose first node is lo."""
    if field.dim() != mesh.dim + 1:
        raise ValueError(f"field must have {mesh.dim + 1} dims")
    f = L.PpmgField()
    f.data = field.data_ptr()
    f.lda = field.shape[-1]
    ext = list(reversed(field.shape[:-1]))
    lo = list(lo)
    for d in range(3):
        f.lo[d] = lo[d] if d < mesh.dim else 0
        f.ext[d] = ext[d] if d < mesh.dim else 1
    return f


def p2m(mesh, xp, up, field, lo, kernel=KERNEL_MP4, prec=PREC_F64, plan=None):
    """ppm_interp_p2m into `field` (a stored box starting at node `lo`)."""
    _check_tensor(xp, POS_DTYPE[prec], "xp", mesh.dim)
    _check_tensor(up, VAL_DTYPE[prec], "up")
    _check_tensor(field, VAL_DTYPE[prec], "field")
    if up.dim() != 2 or up.shape[0] != xp.shape[0]:
        raise ValueError("up must be (np, lda)")
    plan = plan or Plan()
    f = field_desc(mesh, field, lo)
    check(lib().ppmg_p2m(mesh.ref(), kernel, prec, ctypes.c_int64(xp.shape[0]), _ptr(xp), up.shape[1], _ptr(up),
                         ctypes.byref(f), plan.handle, _stream()))
    return plan


def m2p(mesh, xp, field, lo, up, kernel=KERNEL_MP4, prec=PREC_F64, plan=None):
    """ppm_interp_m2p from `field` into `up` (overwritten)."""
    _check_tensor(xp, POS_DTYPE[prec], "xp", mesh.dim)
    _check_tensor(up, VAL_DTYPE[prec], "up")
    _check_tensor(field, VAL_DTYPE[prec], "field")
    plan = plan or Plan()
    f = field_desc(mesh, field, lo)
    check(lib().ppmg_m2p(mesh.ref(), kernel, prec, ctypes.c_int64(xp.shape[0]), _ptr(xp), up.shape[1], ctypes.byref(f),
                         _ptr(up), plan.handle, _stream()))
    return plan


def m2p_prebin(mesh, xp, plan, kernel=KERNEL_MP4, prec=PREC_F64):
    """Bin the particles of the next m2p on `plan` now (on the current
    stream), so that binning can overlap other work; xp must not change
    before that m2p."""
    _check_tensor(xp, POS_DTYPE[prec], "xp", mesh.dim)
    check(lib().ppmg_m2p_prebin(mesh.ref(), kernel, prec, ctypes.c_int64(xp.shape[0]), _ptr(xp), plan.handle,
                                _stream()))
    return plan


def p2m_adjoint(mesh, xp, up, g, lo, plan, kernel=KERNEL_MP4, prec=PREC_F64, want_u=True, want_x=True):
    """Gradients of a loss of the P2M field: returns (dL/du, dL/dx)."""
    _check_tensor(g, VAL_DTYPE[prec], "g")
    gu = torch.empty_like(up) if want_u else None
    gx = torch.empty(xp.shape, dtype=VAL_DTYPE[prec], device=xp.device) if want_x else None
    f = field_desc(mesh, g, lo)
    check(lib().ppmg_p2m_adjoint(mesh.ref(), kernel, prec, ctypes.c_int64(xp.shape[0]), _ptr(xp), up.shape[1],
                                 _ptr(up), ctypes.byref(f), _ptr(gu), _ptr(gx), plan.handle, _stream()))
    return gu, gx


def map_field_ghost(mesh, field, lo, ghostsize, maptype, prec=PREC_F64, plan=None):
    """ppm_map_field_ghost ghost_get (L.MAP_GHOST_GET) or ghost_put
    (L.MAP_GHOST_PUT) + push/send/pop on a stored box holding nodes
    1-ghostsize .. Nm+ghostsize (lo <= -ghostsize), in place."""
    _check_tensor(field, VAL_DTYPE[prec], "field")
    plan = plan or Plan()
    f = field_desc(mesh, field, lo)
    g = (ctypes.c_int * 3)(*([int(v) for v in ghostsize] + [0] * 3)[:3])
    check(lib().ppmg_map_field_ghost(mesh.ref(), prec, ctypes.byref(f), g, int(maptype), plan.handle, _stream()))
    return plan


KIND_DTYPE = {L.KIND_D: torch.float64, L.KIND_S: torch.float32, L.KIND_I: torch.int32, L.KIND_L: torch.int32,
              L.KIND_DC: torch.complex128, L.KIND_SC: torch.complex64}


def map_field_ghost_kind(mesh, field, lo, ghostsize, maptype, kind, mask=None, plan=None):
    """map_field_ghost for every element kind of the reference's specifics
    (L.KIND_*: float64, float32, int32, logical as int32, complex128,
    complex64 tensors) with the optional mask of ppm_map_field_ghost: an int32
    tensor of the field's node shape (no component axis); the pop leaves a
    node whose mask is 0 unchanged.  In place."""
    _check_tensor(field, KIND_DTYPE[kind], "field")
    mptr = None
    if mask is not None:
        _check_tensor(mask, torch.int32, "mask")
        if tuple(mask.shape) != tuple(field.shape[:-1]):
            raise ValueError(f"mask must have shape {tuple(field.shape[:-1])}, got {tuple(mask.shape)}")
        mptr = ctypes.c_void_p(mask.data_ptr())
    plan = plan or Plan()
    f = field_desc(mesh, field, lo)
    g = (ctypes.c_int * 3)(*([int(v) for v in ghostsize] + [0] * 3)[:3])
    check(lib().ppmg_map_field_ghost_kind(mesh.ref(), int(kind), ctypes.byref(f), g, int(maptype), mptr,
                                          plan.handle, _stream()))
    return plan


def util_rank(xp, xmin, xmax, ncells, ngl, plan=None):
    """ppm_util_rank3d / rank2d on device: returns (lpdx[:icount] 1-based, lhbx 1-based)."""
    dim = xp.shape[1]
    prec = PREC_F32 if xp.dtype == torch.float32 else PREC_F64
    _check_tensor(xp, POS_DTYPE[prec], "xp", dim)
    np_ = xp.shape[0]
    nbox = 1
    for d in range(dim):
        nbox *= ncells[d] + ngl[d] + ngl[dim + d]
    lpdx = torch.empty(max(np_, 1), dtype=torch.int32, device=xp.device)
    lhbx = torch.empty(nbox + 1, dtype=torch.int32, device=xp.device)
    ctype = ctypes.c_float if prec == PREC_F32 else ctypes.c_double
    mn = (ctype * 3)(*list(xmin))
    mx = (ctype * 3)(*list(xmax))
    nc = (ctypes.c_int * 3)(*list(ncells))
    ng = (ctypes.c_int * 6)(*list(ngl))
    ic = ctypes.c_int64()
    plan = plan or Plan()
    check(lib().ppmg_util_rank(dim, prec, ctypes.c_int64(np_), _ptr(xp), mn, mx, nc, ng, _ptr(lpdx), _ptr(lhbx),
                               ctypes.byref(ic), plan.handle, _stream()))
    return lpdx[: ic.value], lhbx


def util_sort(xp, xmin, xmax, ncells, plan=None):
    """ppm_util_sort3d / sort2d on device: xp reordered in place into cell
    order (no ghost layers); returns npbx, the particles per box."""
    dim = xp.shape[1]
    prec = PREC_F32 if xp.dtype == torch.float32 else PREC_F64
    _check_tensor(xp, POS_DTYPE[prec], "xp", dim)
    nbox = 1
    for d in range(dim):
        nbox *= ncells[d]
    npbx = torch.empty(max(nbox, 1), dtype=torch.int32, device=xp.device)
    ctype = ctypes.c_float if prec == PREC_F32 else ctypes.c_double
    mn = (ctype * 3)(*list(xmin))
    mx = (ctype * 3)(*list(xmax))
    nc = (ctypes.c_int * 3)(*list(ncells))
    plan = plan or Plan()
    check(lib().ppmg_util_sort(dim, prec, ctypes.c_int64(xp.shape[0]), _ptr(xp), mn, mx, nc, _ptr(npbx), plan.handle,
                               _stream()))
    return npbx[:nbox]


def impose_part_bc(mesh, xp):
    prec = PREC_F32 if xp.dtype == torch.float32 else PREC_F64
    _check_tensor(xp, POS_DTYPE[prec], "xp", mesh.dim)
    check(lib().ppmg_impose_part_bc(mesh.ref(), prec, ctypes.c_int64(xp.shape[0]), _ptr(xp), _stream()))
    return xp


def comp_weights(mesh, xp, kernel=KERNEL_MP4, prec=PREC_F64):
    """ppm_rmsh_comp_weights: returns [wx1, wx2(, wx3)], each (np, 2*ks)."""
    _check_tensor(xp, POS_DTYPE[prec], "xp", mesh.dim)
    ks = 2 if kernel == KERNEL_MP4 else 1
    w = [torch.empty((xp.shape[0], 2 * ks), dtype=VAL_DTYPE[prec], device=xp.device) for _ in range(mesh.dim)]
    w3 = w[2] if mesh.dim == 3 else None
    check(lib().ppmg_rmsh_comp_weights(mesh.ref(), kernel, prec, ctypes.c_int64(xp.shape[0]), _ptr(xp), _ptr(w[0]),
                                       _ptr(w[1]), _ptr(w3), _stream()))
    return w


def remesh(mesh, xp, up, wx, field, lo, kernel=KERNEL_MP4, prec=PREC_F64, plan=None):
    """ppm_rmsh_remesh: scatter with precomputed weights + ghost fold into `field`."""
    _check_tensor(xp, POS_DTYPE[prec], "xp", mesh.dim)
    _check_tensor(up, VAL_DTYPE[prec], "up")
    _check_tensor(field, VAL_DTYPE[prec], "field")
    plan = plan or Plan()
    f = field_desc(mesh, field, lo)
    w3 = wx[2] if mesh.dim == 3 else None
    check(lib().ppmg_rmsh_remesh(mesh.ref(), kernel, prec, ctypes.c_int64(xp.shape[0]), _ptr(xp), up.shape[1],
                                 _ptr(up), _ptr(wx[0]), _ptr(wx[1]), _ptr(w3), ctypes.byref(f), plan.handle,
                                 _stream()))
    return plan


def create_part(mesh, field, lo, cutoff, cutoff_weights=None, resetpos=True, prec=PREC_F64, maxnp=None, plan=None,
                slave=None, slave_lo=None):
    """ppm_rmsh_create_part: returns (xp, up) of the nodes inside the strength band,
    or (xp, up, wp) when a slave field (the reference's field_wp) is given; wp
    holds the slave's values at the same nodes.  xp is None unless resetpos."""
    _check_tensor(field, VAL_DTYPE[prec], "field")
    lda = field.shape[-1]
    if maxnp is None:
        maxnp = 1
        for n in mesh.nm:
            maxnp *= n
    xp = torch.empty((max(maxnp, 1), mesh.dim), dtype=POS_DTYPE[prec], device=field.device)
    up = torch.empty((max(maxnp, 1), lda), dtype=VAL_DTYPE[prec], device=field.device)
    f = field_desc(mesh, field, lo)
    cut = (ctypes.c_double * 2)(*[float(c) for c in cutoff])
    cw = None if cutoff_weights is None else (ctypes.c_double * lda)(*[float(c) for c in cutoff_weights])
    npo = ctypes.c_int64()
    plan = plan or Plan()
    sref, wp = None, None
    if slave is not None:
        _check_tensor(slave, VAL_DTYPE[prec], "slave")
        sd = field_desc(mesh, slave, lo if slave_lo is None else slave_lo)
        sref = ctypes.byref(sd)
        wp = torch.empty((max(maxnp, 1), slave.shape[-1]), dtype=VAL_DTYPE[prec], device=field.device)
    check(lib().ppmg_rmsh_create_part(mesh.ref(), prec, ctypes.byref(f), cut, cw, int(resetpos),
                                      ctypes.c_int64(maxnp), ctypes.byref(npo), _ptr(xp), _ptr(up), sref, _ptr(wp),
                                      plan.handle, _stream()))
    n = npo.value
    xo = xp[:n] if resetpos else None
    if slave is not None:
        return xo, up[:n], wp[:n]
    return xo, up[:n]


def gen_jittered_lattice(n, seed, dim=3, jitter=0.3, device="cuda"):
    xp = torch.empty((n ** dim, dim), dtype=torch.float64, device=device)
    check(lib().ppmg_gen_jittered_lattice(dim, n, ctypes.c_double(jitter), ctypes.c_uint64(seed), _ptr(xp),
                                          _stream()))
    return xp


def gen_normals(count, lda, seed, dtype=torch.float64, device="cuda"):
    out = torch.empty((count, lda), dtype=dtype, device=device)
    prec = PREC_F64 if dtype == torch.float64 else PREC_F32C
    check(lib().ppmg_gen_normals(ctypes.c_int64(count), lda, ctypes.c_uint64(seed), prec, _ptr(out), _stream()))
    return out


def gen_divfree_velocity(xp, modes):
    k, A, phi = modes
    dev = xp.device
    kt = torch.as_tensor(k, dtype=torch.float64, device=dev).contiguous()
    At = torch.as_tensor(A, dtype=torch.float64, device=dev).contiguous()
    pt = torch.as_tensor(phi, dtype=torch.float64, device=dev).contiguous()
    u = torch.empty_like(xp)
    check(lib().ppmg_gen_divfree_velocity(ctypes.c_int64(xp.shape[0]), _ptr(xp), kt.shape[0], _ptr(kt), _ptr(At),
                                          _ptr(pt), _ptr(u), _stream()))
    return u
