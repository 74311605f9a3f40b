"""ctypes wrapper of the CPU oracle (``oracle/fv1_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg as the checker / CPU baseline.  The
product package ``plugin_navierstokes_amd`` never imports it.

The C library restates the reference's FV1 assembly
(``incompressible/fv1/navier_stokes_fv1.cpp:250-892``, ``upwind.cpp:52-575``,
``incompressible/fv1/stabilization.cpp:102-242,805-850``,
``incompressible/fv1/diffusion_length.h:48-199``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libfv1oracle.so")

TRI, QUAD, TET, HEX, PRISM = 0, 1, 2, 3, 4
UPW = {"no": 0, "full": 1, "skewed": 2, "lps": 3, "linearprofileskewed": 3, "positive": 4, "regular": 5}
STAB_FIELDS, STAB_WITHOUT, STAB_FLOW = 0, 1, 2
DIFFLEN = {"raw": 0, "fivepoint": 1, "cor": 2}
NSH = {TRI: 3, QUAD: 4, TET: 4, HEX: 8, PRISM: 6}
NIP = {TRI: 3, QUAD: 4, TET: 6, HEX: 12, PRISM: 9}
DIM = {TRI: 2, QUAD: 2, TET: 3, HEX: 3, PRISM: 3}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    src = os.path.join(_HERE, "fv1_oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "lib/libfv1oracle.so"])
    return _LIB_PATH


class _Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("conv_upwind", C.c_int), ("stab_type", C.c_int),
        ("stab_upwind", C.c_int), ("pac", C.c_int), ("difflen", C.c_int),
        ("stokes", C.c_int), ("laplace", C.c_int), ("peclet", C.c_int),
        ("exact", C.c_double), ("nu", C.c_double), ("rho", C.c_double),
        ("has_source", C.c_int), ("f", C.c_double * 3), ("force_matrix", C.c_int),
        ("nu_ip", C.POINTER(C.c_double)), ("f_scv", C.POINTER(C.c_double)), ("f_scvf", C.POINTER(C.c_double)),
        ("nu_bf", C.POINTER(C.c_double)), ("rho_ip", C.POINTER(C.c_double)), ("rho_scv", C.POINTER(C.c_double)),
    ]


class _BFGeo(C.Structure):
    _fields_ = [
        ("nbf", C.c_int), ("node", C.c_int * 4),
        ("n", (C.c_double * 3) * 4), ("ipl", (C.c_double * 3) * 4), ("ipg", (C.c_double * 3) * 4),
        ("phi", (C.c_double * 8) * 4), ("g", ((C.c_double * 3) * 8) * 4),
    ]


class _Geo(C.Structure):
    _fields_ = [
        ("et", C.c_int), ("dim", C.c_int), ("nsh", C.c_int), ("nip", C.c_int),
        ("co", (C.c_double * 3) * 8), ("frm", C.c_int * 12), ("to", C.c_int * 12),
        ("ipl", (C.c_double * 3) * 12), ("ipg", (C.c_double * 3) * 12), ("n", (C.c_double * 3) * 12),
        ("phi", (C.c_double * 8) * 12), ("g", ((C.c_double * 3) * 8) * 12),
        ("c0", (C.c_double * 3) * 12), ("c2", (C.c_double * 3) * 12), ("vol", C.c_double * 8),
    ]


@dataclass
class Params:
    """Oracle configuration, the same switches as the reference disc object."""
    dim: int = 2
    upwind: str | None = "full"          # convective upwind (m_spConvUpwind)
    stab: str | None = "fields"          # "fields" | "without" | None
    stab_upwind: str | None = None       # defaults to ``upwind`` (shared object)
    pac: bool = False
    difflen: str = "raw"
    stokes: bool = False
    laplace: bool = False
    peclet: bool = False
    exact: float = 0.0
    nu: float = 0.01
    rho: float = 1.0
    source: tuple | None = None
    force_matrix: bool = False           # pinning knob: matrix solve of the ip system without ip shapes
    # imports at the integration points (None: the constants above), indexed
    # by element for mesh-level calls: nu_ip [ne, nip] (SCVF ips), f_scv
    # [ne, nsh, dim] (SCV ips), f_scvf [ne, nip, dim] (SCVF ips); the source
    # arrays need `source` set (it switches the source term on)
    nu_ip: object = None
    f_scv: object = None
    f_scvf: object = None
    # viscosity at the boundary-face ips of boundary calls: [nb, 4] for a
    # list of sides (face j = the side's j-th corner), [4] for one side
    nu_bf: object = None
    # density at the integration points: rho_ip [ne, nip] (SCVF ips),
    # rho_scv [ne, nsh] (SCV ips, the corners)
    rho_ip: object = None
    rho_scv: object = None

    def c(self) -> _Params:
        p = _Params()
        p.dim = self.dim
        p.conv_upwind = -1 if (self.upwind is None or self.pac) else UPW[self.upwind]
        p.stab_type = {None: -1, "fields": STAB_FIELDS, "without": STAB_WITHOUT, "flow": STAB_FLOW}[self.stab]
        su = self.stab_upwind if self.stab_upwind is not None else self.upwind
        p.stab_upwind = -1 if su is None else UPW[su]
        p.pac = int(self.pac)
        p.difflen = DIFFLEN[self.difflen]
        p.stokes, p.laplace, p.peclet = int(self.stokes), int(self.laplace), int(self.peclet)
        p.exact, p.nu, p.rho = float(self.exact), float(self.nu), float(self.rho)
        p.has_source = int(self.source is not None)
        p.force_matrix = int(self.force_matrix)
        if self.source is not None:
            for i, v in enumerate(self.source):
                p.f[i] = float(v)
        keep = []
        for name in ("nu_ip", "f_scv", "f_scvf", "nu_bf", "rho_ip", "rho_scv"):
            a = getattr(self, name)
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.float64)
                keep.append(a)
                setattr(p, name, a.ctypes.data_as(C.POINTER(C.c_double)))
        p._keep = keep  # the arrays live as long as the structure
        return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip32 = C.POINTER(C.c_int32)
        ip64 = C.POINTER(C.c_int64)
        L.or_geometry.argtypes = [C.c_int, dp, C.POINTER(_Geo)]
        L.or_side_ray.argtypes = [C.c_int, dp, dp, dp, C.POINTER(C.c_int), dp, dp]
        L.or_upwind.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.or_elem_A.argtypes = [C.POINTER(_Params), C.c_int, dp, dp, dp, C.c_double, dp, dp]
        L.or_elem_M.argtypes = [C.POINTER(_Params), C.c_int, dp, dp, dp, dp]
        L.or_elem_rhs.argtypes = [C.POINTER(_Params), C.c_int, dp, dp]
        L.or_assemble.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int64, ip32, C.c_int64, C.c_int64,
                                  dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                  ip64, ip32, dp, dp]
        L.or_assemble_colored.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int64, ip32, ip64, ip64, C.c_int,
                                          dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                          ip64, ip32, dp, dp, C.c_int]
        L.or_assemble_abs.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int64, ip32, C.c_int64,
                                      dp, dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                      ip64, ip32, dp, dp]
        L.or_pattern.argtypes = [C.c_int, C.c_int64, ip32, C.c_int64, ip64, ip32, ip64]
        L.or_bnd_faces.argtypes = [C.c_int, dp, C.c_int, C.POINTER(_BFGeo)]
        L.or_outflow_elem.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int, dp, dp, dp]
        L.or_outflow.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int64, ip32, ip32, ip32, C.c_int64, dp,
                                 C.c_double, ip64, ip32, dp, dp]
        L.or_symmetry_elem.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int, dp, dp, dp]
        L.or_symmetry.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int64, ip32, ip32, ip32, C.c_int64, dp,
                                  C.c_double, ip64, ip32, dp, dp]
        L.or_sliding_elem.argtypes = [C.POINTER(_Params), C.c_int, dp, C.c_int, dp, C.c_double, C.c_double, dp, dp]
        L.or_sliding.argtypes = [C.POINTER(_Params), C.c_int, dp, ip32, ip32, ip32, C.c_int64, dp, C.c_double,
                                 C.c_double, C.c_double, ip64, ip32, dp, dp]
        L.or_export_elem.argtypes = [C.POINTER(_Params), C.c_int, dp, dp, dp, C.c_double, C.c_int, C.c_int, dp,
                                     C.c_int, dp]
        L.or_export.argtypes = [C.POINTER(_Params), C.c_int, dp, ip32, C.c_int64, dp, dp, C.c_double, C.c_int, C.c_int,
                                dp, C.c_int, dp]
        L.or_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"oracle error {code}: {msg}")
        self.code = code


def _chk(rc: int):
    if rc != 0:
        raise OracleError(rc, lib().or_last_error().decode())


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def geometry(et: int, corners) -> dict:
    co = _f64(corners)
    g = _Geo()
    _chk(lib().or_geometry(et, _dp(co), C.byref(g)))
    dim, nsh, nip = g.dim, g.nsh, g.nip
    a = lambda x: np.ctypeslib.as_array(x)
    return {
        "from": a(g.frm)[:nip].copy(), "to": a(g.to)[:nip].copy(),
        "ipl": a(g.ipl)[:nip, :dim].copy(), "ipg": a(g.ipg)[:nip, :dim].copy(),
        "n": a(g.n)[:nip, :dim].copy(), "phi": a(g.phi)[:nip, :nsh].copy(),
        "grad": a(g.g)[:nip, :nsh, :dim].copy(), "c0": a(g.c0)[:nip, :dim].copy(),
        "c2": a(g.c2)[:nip, :dim].copy(), "vol": a(g.vol)[:nsh].copy(),
    }


def side_ray(et: int, corners, frm, direction):
    co, f, d = _f64(corners), _f64(frm), _f64(direction)
    side = C.c_int(-1)
    gi = np.zeros(3)
    li = np.zeros(3)
    _chk(lib().or_side_ray(et, _dp(co), _dp(f), _dp(d), C.byref(side), _dp(gi), _dp(li)))
    return side.value, gi[: DIM[et]], li[: DIM[et]]


def upwind(kind: str, et: int, corners, ipvel, ip_shapes: bool = False):
    """Upwind corner shapes U [nip, nsh] and convection lengths [nip]
    (and, with ip_shapes, the ip shapes W [nip, nip])."""
    co, v = _f64(corners), _f64(ipvel)
    U = np.zeros((12, 8))
    W = np.zeros((12, 12))
    L = np.zeros(12)
    _chk(lib().or_upwind(UPW[kind], et, _dp(co), _dp(v), _dp(U), _dp(W), _dp(L)))
    if ip_shapes:
        return U[: NIP[et], : NSH[et]], W[: NIP[et], : NIP[et]], L[: NIP[et]]
    return U[: NIP[et], : NSH[et]], L[: NIP[et]]


def elem_A(p: Params, et: int, corners, uloc, uold=None, dt=0.0):
    nl = NSH[et] * (DIM[et] + 1)
    J = np.zeros((nl, nl))
    d = np.zeros(nl)
    pc = p.c()
    _chk(lib().or_elem_A(C.byref(pc), et, _dp(_f64(corners)), _dp(_f64(uloc)), _dp(_f64(uold)),
                         float(dt), _dp(J), _dp(d)))
    return J, d


def elem_M(p: Params, et: int, corners, uloc):
    nl = NSH[et] * (DIM[et] + 1)
    J = np.zeros((nl, nl))
    d = np.zeros(nl)
    pc = p.c()
    _chk(lib().or_elem_M(C.byref(pc), et, _dp(_f64(corners)), _dp(_f64(uloc)), _dp(J), _dp(d)))
    return J, d


def elem_rhs(p: Params, et: int, corners):
    nl = NSH[et] * (DIM[et] + 1)
    f = np.zeros(nl)
    pc = p.c()
    _chk(lib().or_elem_rhs(C.byref(pc), et, _dp(_f64(corners)), _dp(f)))
    return f


def pattern(et: int, nv: int, conn):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    ne = conn.shape[0]
    nc = DIM[et] + 1
    rowptr = np.zeros(nv * nc + 1, dtype=np.int64)
    nnz = C.c_int64(0)
    L = lib()
    ip32 = C.POINTER(C.c_int32)
    ip64 = C.POINTER(C.c_int64)
    _chk(L.or_pattern(et, nv, conn.ctypes.data_as(ip32), ne, rowptr.ctypes.data_as(ip64), None, C.byref(nnz)))
    colidx = np.zeros(nnz.value, dtype=np.int32)
    _chk(L.or_pattern(et, nv, conn.ctypes.data_as(ip32), ne, rowptr.ctypes.data_as(ip64),
                      colidx.ctypes.data_as(ip32), C.byref(nnz)))
    return rowptr, colidx


def assemble(p: Params, et: int, xyz, conn, u, uold=None, dt=0.0, s_a=1.0, s_m=0.0, s_f=0.0,
             parts=1, rowptr=None, colidx=None, e0=0, e1=None, want_J=True, want_d=True):
    xyz, u, uold = _f64(xyz), _f64(u), _f64(uold)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    nv = xyz.shape[0]
    if rowptr is None:
        rowptr, colidx = pattern(et, nv, conn)
    ndof = nv * (DIM[et] + 1)
    J = np.zeros(int(rowptr[-1])) if want_J else None
    d = np.zeros(ndof) if want_d else None
    e1 = conn.shape[0] if e1 is None else e1
    pc = p.c()
    _chk(lib().or_assemble(C.byref(pc), et, _dp(xyz), nv, conn.ctypes.data_as(C.POINTER(C.c_int32)), e0, e1,
                           _dp(u), _dp(uold), float(dt), float(s_a), float(s_m), float(s_f), int(parts),
                           rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                           colidx.ctypes.data_as(C.POINTER(C.c_int32)), _dp(J), _dp(d)))
    return J, d


def structured_colors(shape):
    """Element colouring of a lexicographic structured mesh (x fastest):
    colour = parity of each cell index, so elements of one colour share no
    node.  Returns (order, color_ptr)."""
    idx = np.indices(tuple(shape)[::-1]).reshape(len(shape), -1)[::-1]  # idx[0] = i (fastest)
    col = np.zeros(idx.shape[1], dtype=np.int64)
    for a in range(len(shape)):
        col += (idx[a] % 2) << a
    order = np.argsort(col, kind="stable").astype(np.int64)
    color_ptr = np.searchsorted(col[order], np.arange((1 << len(shape)) + 1)).astype(np.int64)
    return order, color_ptr


def assemble_colored(p: Params, et: int, xyz, conn, u, order, color_ptr, uold=None, dt=0.0, s_a=1.0, s_m=0.0,
                     s_f=0.0, parts=1, rowptr=None, colidx=None, threads=1):
    """assemble() with the elements of each colour spread over `threads`
    OpenMP threads (the all-cores CPU baseline)."""
    xyz, u, uold = _f64(xyz), _f64(u), _f64(uold)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    order = np.ascontiguousarray(order, dtype=np.int64)
    color_ptr = np.ascontiguousarray(color_ptr, dtype=np.int64)
    nv = xyz.shape[0]
    if rowptr is None:
        rowptr, colidx = pattern(et, nv, conn)
    J = np.zeros(int(rowptr[-1]))
    d = np.zeros(nv * (DIM[et] + 1))
    pc = p.c()
    ip64 = C.POINTER(C.c_int64)
    _chk(lib().or_assemble_colored(C.byref(pc), et, _dp(xyz), nv, conn.ctypes.data_as(C.POINTER(C.c_int32)),
                                   order.ctypes.data_as(ip64), color_ptr.ctypes.data_as(ip64), len(color_ptr) - 1,
                                   _dp(u), _dp(uold), float(dt), float(s_a), float(s_m), float(s_f), int(parts),
                                   rowptr.ctypes.data_as(ip64), colidx.ctypes.data_as(C.POINTER(C.c_int32)),
                                   _dp(J), _dp(d), int(threads)))
    return J, d


def assemble_abs(p: Params, et: int, xyz, conn, u, uold=None, dt=0.0, s_a=1.0, s_m=0.0, parts=1,
                 rowptr=None, colidx=None):
    """Sum of |element contributions| per entry: the scale for mixed tolerances."""
    xyz, u, uold = _f64(xyz), _f64(u), _f64(uold)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    nv = xyz.shape[0]
    if rowptr is None:
        rowptr, colidx = pattern(et, nv, conn)
    J = np.zeros(int(rowptr[-1]))
    d = np.zeros(nv * (DIM[et] + 1))
    pc = p.c()
    _chk(lib().or_assemble_abs(C.byref(pc), et, _dp(xyz), nv, conn.ctypes.data_as(C.POINTER(C.c_int32)),
                               conn.shape[0], _dp(u), _dp(uold), float(dt), float(s_a), float(s_m), int(parts),
                               rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                               colidx.ctypes.data_as(C.POINTER(C.c_int32)), _dp(J), _dp(d)))
    return J, d


def pattern_mixed(nv: int, groups):
    """CSR pattern of a mesh of several element types (groups of (type,
    connectivity) over nv nodes): node v's row holds every node that shares
    an element with it, of any group, sorted, in (dim+1)^2 blocks (an
    isolated node keeps its own block)."""
    dim = DIM[groups[0][0]]
    nc = dim + 1
    pairs = [np.stack([np.arange(nv), np.arange(nv)], 1)]
    for et, conn in groups:
        conn = np.asarray(conn, dtype=np.int64)
        n = conn.shape[1]
        a = np.repeat(conn, n, axis=1).ravel()
        b = np.tile(conn, (1, n)).ravel()
        pairs.append(np.stack([a, b], 1))
    pr = np.unique(np.concatenate(pairs), axis=0)  # sorted by row node, then column node
    deg = np.bincount(pr[:, 0], minlength=nv)
    rowptr = np.zeros(nv * nc + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(np.repeat(deg * nc, nc))
    colidx = np.empty(int(rowptr[-1]), dtype=np.int32)
    start = np.concatenate([[0], np.cumsum(deg)])
    for v in range(nv):
        nb = pr[start[v]:start[v + 1], 1]
        cols = (nb[:, None] * nc + np.arange(nc)[None, :]).ravel()
        for a in range(nc):
            colidx[rowptr[v * nc + a]:rowptr[v * nc + a + 1]] = cols
    return rowptr, colidx


def assemble_mixed(p: Params, groups, xyz, u, rowptr, colidx, **kw):
    """A mesh of several element types: the sum of assemble() over its
    groups in one pattern (the element loop of each type adds into the same
    matrix and defect)."""
    J = d = None
    for et, conn in groups:
        Jg, dg = assemble(p, et, xyz, conn, u, rowptr=rowptr, colidx=colidx, **kw)
        J = Jg if J is None else J + Jg
        d = dg if d is None else d + dg
    return J, d


def assemble_abs_mixed(p: Params, groups, xyz, u, rowptr, colidx, **kw):
    """assemble_abs() summed over the groups of a mixed mesh."""
    J = d = None
    for et, conn in groups:
        Jg, dg = assemble_abs(p, et, xyz, conn, u, rowptr=rowptr, colidx=colidx, **kw)
        J = Jg if J is None else J + Jg
        d = dg if d is None else d + dg
    return J, d


def bnd_faces(et: int, corners, side: int) -> dict:
    """FV1 boundary faces of one element side (restated ugcore FV1Geometry BF)."""
    co = _f64(corners)
    b = _BFGeo()
    _chk(lib().or_bnd_faces(et, _dp(co), int(side), C.byref(b)))
    dim, nsh, k = DIM[et], NSH[et], b.nbf
    a = lambda x: np.ctypeslib.as_array(x)
    return {"node": a(b.node)[:k].copy(), "n": a(b.n)[:k, :dim].copy(), "ipl": a(b.ipl)[:k, :dim].copy(),
            "ipg": a(b.ipg)[:k, :dim].copy(), "phi": a(b.phi)[:k, :nsh].copy(), "grad": a(b.g)[:k, :nsh, :dim].copy()}


def outflow_elem(p: Params, et: int, corners, side: int, uloc, kind: str = "outflow", k: float = 0.0,
                 lim: float = 0.0):
    """Element-local (J, d) of a boundary condition ("outflow": no normal
    stress; "symmetry"; "sliding": sliding wall with factor k and limit lim)
    on one side."""
    nl = NSH[et] * (DIM[et] + 1)
    J = np.zeros(nl * nl)
    d = np.zeros(nl)
    pc = p.c()
    co, ul = _dp(_f64(corners)), _dp(_f64(uloc))
    if kind == "sliding":
        _chk(lib().or_sliding_elem(C.byref(pc), et, co, int(side), ul, float(k), float(lim), _dp(J), _dp(d)))
    else:
        f = lib().or_outflow_elem if kind == "outflow" else lib().or_symmetry_elem
        _chk(f(C.byref(pc), et, co, int(side), ul, _dp(J), _dp(d)))
    return J.reshape(nl, nl), d


def outflow(p: Params, et: int, xyz, conn, belem, bside, u, s_a=1.0, rowptr=None, colidx=None, J=None, d=None,
            kind: str = "outflow", k: float = 0.0, lim: float = 0.0):
    """Adds s_a x the outflow contributions of the (element, side) list to J, d
    (new zero arrays when None).  Returns (J, d)."""
    xyz, u = _f64(xyz), _f64(u)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    belem = np.ascontiguousarray(belem, dtype=np.int32)
    bside = np.ascontiguousarray(bside, dtype=np.int32)
    nv = xyz.shape[0]
    if rowptr is None:
        rowptr, colidx = pattern(et, nv, conn)
    if J is None:
        J = np.zeros(int(rowptr[-1]))
    if d is None:
        d = np.zeros(nv * (DIM[et] + 1))
    pc = p.c()
    i32 = C.POINTER(C.c_int32)
    rp, ci = rowptr.ctypes.data_as(C.POINTER(C.c_int64)), colidx.ctypes.data_as(i32)
    if kind == "sliding":
        _chk(lib().or_sliding(C.byref(pc), et, _dp(xyz), conn.ctypes.data_as(i32), belem.ctypes.data_as(i32),
                              bside.ctypes.data_as(i32), len(belem), _dp(u), float(s_a), float(k), float(lim), rp, ci,
                              _dp(J), _dp(d)))
    else:
        f = lib().or_outflow if kind == "outflow" else lib().or_symmetry
        _chk(f(C.byref(pc), et, _dp(xyz), nv, conn.ctypes.data_as(i32), belem.ctypes.data_as(i32),
               bside.ctypes.data_as(i32), len(belem), _dp(u), float(s_a), rp, ci, _dp(J), _dp(d)))
    return J, d


EXPORTS = {"velocity": 0, "velocity_grad": 1, "velocity_ip": 2, "pressure": 3, "pressure_grad": 4}
AT = {"scvf": 0, "scv": 1, "local": 2}


def export_shape(et: int, what: str, at: str, nip: int = 0):
    """(points per element, components) of an export."""
    dim = DIM[et]
    npt = {"scvf": NIP[et], "scv": NSH[et], "local": nip}[at]
    ncomp = {"velocity": dim, "velocity_grad": dim * dim, "velocity_ip": dim, "pressure": 1,
             "pressure_grad": dim}[what]
    return npt, ncomp


def export(p: Params, et: int, xyz, conn, u, what: str, at: str = "scvf", local_ips=None, u_old=None, dt=0.0):
    """Data export of every element: [ne][points][components]."""
    xyz, u = _f64(xyz), _f64(u)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    ne = conn.shape[0]
    lips = _f64(local_ips if local_ips is not None else np.zeros((1, DIM[et])))
    nip = 0 if local_ips is None else lips.shape[0]
    npt, ncomp = export_shape(et, what, at, nip)
    out = np.zeros(ne * npt * ncomp)
    pc = p.c()
    uo = _f64(u_old) if u_old is not None else None
    _chk(lib().or_export(C.byref(pc), et, _dp(xyz), conn.ctypes.data_as(C.POINTER(C.c_int32)), ne, _dp(u),
                         _dp(uo) if uo is not None else None, float(dt), EXPORTS[what], AT[at], _dp(lips), nip,
                         _dp(out)))
    return out.reshape(ne, npt, ncomp)
