"""MI355X-native FV1 incompressible Navier-Stokes assembly.

``NavierStokesFV1`` mirrors the Lua/registry surface of the reference
``NavierStokesFV1`` element discretisation
(``incompressible/fv1/register_fv1.cpp:167-184``,
``register_navier_stokes.cpp:104-126``,
``incompressible/incompressible_navier_stokes_plugin.cpp:244-266``): the same
setter names, string keys, defaults and error cases.  Every call goes through
the C ABI of ``libnsfv1.so`` (declared in ``include/nsfv1.h``); device
buffers are torch tensors.  There is no CPU fallback: without the built HIP
library or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import meshes  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnsfv1.so")

TRIANGLE, QUADRILATERAL, TETRAHEDRON, HEXAHEDRON = 0, 1, 2, 3
PART_A, PART_M, PART_RHS = 1, 2, 4
F32, F64 = 32, 64

ERRORS = {
    -1: "ARGUMENT", -2: "FUNCTION_COUNT", -3: "TRIAL_SPACE", -4: "NO_STABILIZATION", -5: "NO_UPWIND",
    -6: "NO_VISCOSITY", -7: "HANGING_NODES", -8: "TIME_SERIES", -9: "CUT_SIDE", -10: "UNKNOWN_NAME",
    -11: "NOT_IMPLEMENTED", -12: "NO_MESH", -13: "DEVICE", -14: "DEGENERATE", -15: "SINGULAR_BLOCK",
}

# every symbol the C ABI exports (include/nsfv1.h)
EXPORTS = [
    "nsfv1_last_error", "nsfv1_create", "nsfv1_destroy", "nsfv1_prepare_setting", "nsfv1_disc_type",
    "nsfv1_set_kinematic_viscosity", "nsfv1_set_source", "nsfv1_set_exact_jacobian", "nsfv1_set_density",
    "nsfv1_set_kinematic_viscosity_ip", "nsfv1_set_source_ip", "nsfv1_set_density_ip", "nsfv1_ip_positions",
    "nsfv1_set_stokes", "nsfv1_set_laplace", "nsfv1_set_peclet_blend", "nsfv1_set_grad_div",
    "nsfv1_set_upwind", "nsfv1_set_stabilization", "nsfv1_set_stabilization_without",
    "nsfv1_set_stabilization_upwind", "nsfv1_set_pac_upwind", "nsfv1_set_mesh", "nsfv1_set_mesh_mixed", "nsfv1_bnd_ip_positions",
    "nsfv1_set_boundary_viscosity", "nsfv1_num_dofs", "nsfv1_nnz",
    "nsfv1_get_pattern", "nsfv1_pattern_device", "nsfv1_assemble", "nsfv1_assemble_noalloc", "nsfv1_sync_errors",
    "nsfv1_set_launch",
    "nsfv1_set_dirichlet", "nsfv1_apply_dirichlet", "nsfv1_spmv", "nsfv1_linear_solve",
    "nsfv1_set_preconditioner", "nsfv1_mg_set_hierarchy", "nsfv1_mg_setup", "nsfv1_set_upwind_decisions",
    "nsfv1_mlp_sizes", "nsfv1_mlp_pack", "nsfv1_mlp_forward", "nsfv1_mlp_mse", "nsfv1_mlp_backward", "nsfv1_mlp_adam",
    "nsfv1_predict_features", "nsfv1_predict_guess", "nsfv1_predict_target", "nsfv1_batch_set_instances",
    "nsfv1_set_assembly_path", "nsfv1_assembly_path_used", "nsfv1_set_outflow", "nsfv1_set_symmetry", "nsfv1_set_sliding_wall", "nsfv1_export_size",
    "nsfv1_export", "nsfv1_newton_step", "nsfv1_newton_graph_capture", "nsfv1_newton_graph_replay", "nsfv1_autotune",
]
# data exports (navier_stokes_fv1.cpp:1573-1577) and their points
EX = {"velocity": 0, "velocity_grad": 1, "velocity_ip": 2, "pressure": 3, "pressure_grad": 4}
AT = {"scvf": 0, "scv": 1, "local": 2}
UPWIND_FRESH, UPWIND_RECORD, UPWIND_REPLAY = 0, 1, 2
# device scratch of nsfv1_assemble_noalloc (NSFV1_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 128 << 20


class NSFV1Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"NSFV1_ERR_{ERRORS.get(code, code)} ({code}): {msg}")
        self.code = code
        self.name = ERRORS.get(code, str(code))


_lib = None


def load_library(path: str | None = None):
    """Load the HIP library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m plugin_navierstokes_amd.build`")
    L = C.CDLL(path)
    vp, dp = C.c_void_p, C.c_double
    H = C.c_void_p
    L.nsfv1_last_error.restype = C.c_char_p
    L.nsfv1_create.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.POINTER(H)]
    L.nsfv1_destroy.argtypes = [H]
    L.nsfv1_prepare_setting.argtypes = [H, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int]
    L.nsfv1_disc_type.argtypes = [H, C.c_char_p, C.c_size_t]
    for n in ["nsfv1_set_kinematic_viscosity", "nsfv1_set_exact_jacobian", "nsfv1_set_density", "nsfv1_set_grad_div"]:
        getattr(L, n).argtypes = [H, dp]
    for n in ["nsfv1_set_stokes", "nsfv1_set_laplace", "nsfv1_set_peclet_blend", "nsfv1_set_pac_upwind"]:
        getattr(L, n).argtypes = [H, C.c_int]
    L.nsfv1_set_source.argtypes = [H, C.POINTER(dp), C.c_int]
    L.nsfv1_set_kinematic_viscosity_ip.argtypes = [H, C.POINTER(dp), C.c_int64]
    L.nsfv1_set_source_ip.argtypes = [H, C.POINTER(dp), C.c_int64, C.POINTER(dp), C.c_int64]
    L.nsfv1_set_density_ip.argtypes = [H, C.POINTER(dp), C.c_int64, C.POINTER(dp), C.c_int64]
    L.nsfv1_ip_positions.argtypes = [H, C.c_int, C.POINTER(dp), C.c_int64]
    L.nsfv1_set_upwind.argtypes = [H, C.c_char_p]
    L.nsfv1_set_stabilization.argtypes = [H, C.c_char_p, C.c_char_p]
    L.nsfv1_set_stabilization_without.argtypes = [H]
    L.nsfv1_set_stabilization_upwind.argtypes = [H, C.c_char_p]
    L.nsfv1_set_mesh.argtypes = [H, C.c_int, vp, C.c_int64, vp, C.c_int64]
    L.nsfv1_bnd_ip_positions.argtypes = [H, C.c_int, C.POINTER(dp), C.c_int64]
    L.nsfv1_set_boundary_viscosity.argtypes = [H, C.c_int, C.POINTER(dp), C.c_int64]
    L.nsfv1_set_mesh_mixed.argtypes = [H, vp, C.c_int64, C.c_int, C.POINTER(C.c_int), C.POINTER(vp),
                                       C.POINTER(C.c_int64)]
    L.nsfv1_num_dofs.argtypes = [H, C.POINTER(C.c_int64)]
    L.nsfv1_nnz.argtypes = [H, C.POINTER(C.c_int64)]
    L.nsfv1_get_pattern.argtypes = [H, vp, vp]
    L.nsfv1_pattern_device.argtypes = [H, C.POINTER(vp), C.POINTER(vp)]
    L.nsfv1_assemble.argtypes = [H, vp, vp, dp, dp, dp, dp, C.c_int, vp, C.c_int, vp, vp]
    L.nsfv1_assemble_noalloc.argtypes = L.nsfv1_assemble.argtypes
    L.nsfv1_sync_errors.argtypes = [H, vp]
    L.nsfv1_set_launch.argtypes = [H, C.c_int, C.c_int]
    L.nsfv1_set_dirichlet.argtypes = [H, vp, C.c_int64]
    L.nsfv1_apply_dirichlet.argtypes = [H, vp, C.c_int, vp, vp, vp, vp]
    L.nsfv1_spmv.argtypes = [H, vp, C.c_int, vp, vp, vp]
    L.nsfv1_newton_step.argtypes = [H, vp, vp, dp, dp, vp, vp, vp, C.c_int, vp, C.c_int, vp, vp]
    L.nsfv1_newton_graph_capture.argtypes = L.nsfv1_newton_step.argtypes
    L.nsfv1_newton_graph_replay.argtypes = [H, vp]
    L.nsfv1_autotune.argtypes = [H, vp, vp, dp, dp, C.c_int, vp, C.c_int, vp, C.c_int, C.POINTER(dp),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), vp]
    L.nsfv1_linear_solve.argtypes = [H, vp, C.c_int, vp, vp, dp, C.c_int, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), vp]
    L.nsfv1_set_preconditioner.argtypes = [H, C.c_char_p]
    L.nsfv1_set_upwind_decisions.argtypes = [H, C.c_int]
    L.nsfv1_batch_set_instances.argtypes = [H, C.c_int64, vp, vp]
    L.nsfv1_set_assembly_path.argtypes = [H, C.c_int]
    L.nsfv1_assembly_path_used.argtypes = [H, C.POINTER(C.c_int)]
    L.nsfv1_set_outflow.argtypes = [H, vp, vp, C.c_int64]
    L.nsfv1_set_symmetry.argtypes = [H, vp, vp, C.c_int64]
    L.nsfv1_set_sliding_wall.argtypes = [H, vp, vp, C.c_int64, C.c_double, C.c_double]
    L.nsfv1_export_size.argtypes = [H, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)]
    L.nsfv1_export.argtypes = [H, C.c_int, C.c_int, vp, C.c_int, vp, vp, C.c_double, vp, vp]
    L.nsfv1_mlp_sizes.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.nsfv1_mlp_pack.argtypes = [vp, vp, vp]
    L.nsfv1_mlp_forward.argtypes = [vp, vp, vp, C.c_int64, vp, vp, vp, vp]
    L.nsfv1_mlp_mse.argtypes = [vp, vp, C.c_int64, vp, vp, C.POINTER(C.c_double), vp]
    L.nsfv1_mlp_backward.argtypes = [vp, vp, vp, vp, vp, C.c_int64, vp, vp, vp]
    L.nsfv1_mlp_adam.argtypes = [vp, vp, vp, vp, C.c_int, dp, dp, dp, dp, vp, vp]
    L.nsfv1_predict_features.argtypes = [H, vp, vp, vp, vp, vp, vp, vp, vp]
    L.nsfv1_predict_guess.argtypes = [H, vp, vp, vp, vp, vp, vp, vp, vp]
    L.nsfv1_predict_target.argtypes = [H, vp, vp, vp, vp, vp, vp, vp, vp]
    L.nsfv1_mg_set_hierarchy.argtypes = [H, C.c_int, C.POINTER(H), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                         C.POINTER(vp), C.c_int, C.c_int, dp]
    L.nsfv1_mg_setup.argtypes = [H, C.c_int, vp, vp, dp, dp, dp, C.c_int, vp]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise NSFV1Error(rc, load_library().nsfv1_last_error().decode())


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        return None
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


class NavierStokesFV1:
    """Host object of the GPU assembly; mirrors ``NavierStokesFV1{2d,3d}``."""

    def __init__(self, functions="u,v,p", subsets="Inner", dim: int | None = None):
        if isinstance(functions, (list, tuple)):
            functions = ",".join(functions)
        if isinstance(subsets, (list, tuple)):
            subsets = ",".join(subsets)
        nf = len([f for f in functions.split(",") if f.strip()])
        self.dim = dim if dim is not None else nf - 1
        L = load_library()
        h = C.c_void_p()
        _check(L.nsfv1_create(self.dim, functions.encode(), subsets.encode(), C.byref(h)))
        self._h = h
        self.mesh = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.nsfv1_destroy(h)
            self._h = None

    # --- NavierStokesBase ---
    def set_kinematic_viscosity(self, nu):
        """A number (constant user data), or a function of position: a
        callable x [npts, dim] -> [npts], evaluated at the SCVF integration
        points (navier_stokes_fv1.cpp:192, 243), or the values there
        ([ne, nip]).  A function is evaluated again after a new mesh."""
        self._nu_fn = None
        if callable(nu) or isinstance(nu, np.ndarray):
            self._nu_fn = nu
            self._apply_nu_fn()
        else:
            _check(_lib.nsfv1_set_kinematic_viscosity(self._h, float(nu)))

    def set_source(self, f):
        """A vector (constant user data), None (no source), a function of
        position x [npts, dim] -> [npts, dim] evaluated at the SCV and SCVF
        integration points (navier_stokes_fv1.cpp:195-196, 246-247), or the
        pair of values there ([ne, nsh, dim], [ne, nip, dim])."""
        self._f_fn = None
        if callable(f) or (isinstance(f, tuple) and len(f) == 2 and isinstance(f[0], np.ndarray)):
            self._f_fn = f
            self._apply_f_fn()
            return
        arr = (C.c_double * len(f))(*[float(x) for x in f]) if f is not None else None
        _check(_lib.nsfv1_set_source(self._h, arr, 0 if f is None else len(f)))

    def ip_positions(self, kind: str = "scvf") -> np.ndarray:
        """Global integration points: "scvf" [ne, nip, dim] or "scv" [ne, nsh, dim] (the corners);
        on a mixed mesh [points, dim], group after group."""
        k = {"scvf": 0, "scv": 1}[kind]
        per = {0: (3, 3), 1: (4, 4), 2: (6, 4), 3: (12, 8), 4: (9, 6)}
        groups = getattr(self, "_groups", [(self._et, self.ne)])
        if len(groups) == 1:
            out = np.zeros((self.ne, per[self._et][k], self.dim))
        else:
            out = np.zeros((sum(ne * per[et][k] for et, ne in groups), self.dim))
        _check(_lib.nsfv1_ip_positions(self._h, k, out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def _apply_nu_fn(self):
        fn = getattr(self, "_nu_fn", None)
        if fn is None:
            return
        if isinstance(fn, np.ndarray):
            vals = np.ascontiguousarray(fn, dtype=np.float64)
        else:
            x = self.ip_positions("scvf")
            vals = np.ascontiguousarray(np.asarray(fn(x.reshape(-1, self.dim)), dtype=np.float64).reshape(x.shape[:-1]))
        _check(_lib.nsfv1_set_kinematic_viscosity_ip(self._h, vals.ctypes.data_as(C.POINTER(C.c_double)), vals.size))
        if callable(fn):
            for c in getattr(self, "_mg_levels", []):
                c.set_kinematic_viscosity(fn)
            for kind in (0, 1):
                self._apply_bnd_nu(kind)

    def _apply_f_fn(self):
        fn = getattr(self, "_f_fn", None)
        if fn is None:
            return
        if isinstance(fn, tuple):
            a, b = (np.ascontiguousarray(v, dtype=np.float64) for v in fn)
        else:
            xs, xf = self.ip_positions("scv"), self.ip_positions("scvf")
            a = np.ascontiguousarray(np.asarray(fn(xs.reshape(-1, self.dim)), dtype=np.float64).reshape(xs.shape))
            b = np.ascontiguousarray(np.asarray(fn(xf.reshape(-1, self.dim)), dtype=np.float64).reshape(xf.shape))
        dp = C.POINTER(C.c_double)
        _check(_lib.nsfv1_set_source_ip(self._h, a.ctypes.data_as(dp), a.size, b.ctypes.data_as(dp), b.size))
        if callable(fn):
            for c in getattr(self, "_mg_levels", []):
                c.set_source(fn)

    def set_exact_jacobian(self, factor):
        _check(_lib.nsfv1_set_exact_jacobian(self._h, (1.0 if factor else 0.0) if isinstance(factor, bool) else float(factor)))

    def disc_type(self) -> str:
        buf = C.create_string_buffer(16)
        _check(_lib.nsfv1_disc_type(self._h, buf, 16))
        return buf.value.decode()

    # --- IncompressibleNavierStokesBase ---
    def set_density(self, rho):
        """A number (constant user data), a function of position x [npts, dim]
        -> [npts] evaluated at the SCVF and SCV integration points
        (navier_stokes_fv1.cpp:118-121, 193-194, 244-245), or the pair of
        values there ([ne, nip], [ne, nsh])."""
        self._rho_fn = None
        if callable(rho) or isinstance(rho, tuple):
            self._rho_fn = rho
            self._apply_rho_fn()
            return
        _check(_lib.nsfv1_set_density(self._h, float(rho)))

    def _apply_rho_fn(self):
        fn = getattr(self, "_rho_fn", None)
        if fn is None:
            return
        if isinstance(fn, tuple):
            a, b = (np.ascontiguousarray(v, dtype=np.float64) for v in fn)
        else:
            xf, xs = self.ip_positions("scvf"), self.ip_positions("scv")
            a = np.ascontiguousarray(np.asarray(fn(xf.reshape(-1, self.dim)), dtype=np.float64).reshape(xf.shape[:-1]))
            b = np.ascontiguousarray(np.asarray(fn(xs.reshape(-1, self.dim)), dtype=np.float64).reshape(xs.shape[:-1]))
        dp = C.POINTER(C.c_double)
        _check(_lib.nsfv1_set_density_ip(self._h, a.ctypes.data_as(dp), a.size, b.ctypes.data_as(dp), b.size))
        if callable(fn):
            for c in getattr(self, "_mg_levels", []):
                c.set_density(fn)

    def set_stokes(self, on: bool):
        _check(_lib.nsfv1_set_stokes(self._h, int(bool(on))))

    def set_laplace(self, on: bool):
        _check(_lib.nsfv1_set_laplace(self._h, int(bool(on))))

    def set_peclet_blend(self, on: bool):
        _check(_lib.nsfv1_set_peclet_blend(self._h, int(bool(on))))

    def set_grad_div(self, factor: float):
        _check(_lib.nsfv1_set_grad_div(self._h, float(factor)))

    # --- NavierStokesFV1 ---
    def set_upwind(self, name: str):
        _check(_lib.nsfv1_set_upwind(self._h, name.encode()))

    def set_stabilization(self, name: str, diff_length: str | None = None):
        _check(_lib.nsfv1_set_stabilization(self._h, name.encode(), None if diff_length is None else diff_length.encode()))

    def set_stabilization_without(self):
        _check(_lib.nsfv1_set_stabilization_without(self._h))

    def set_stabilization_upwind(self, name: str):
        _check(_lib.nsfv1_set_stabilization_upwind(self._h, name.encode()))

    def set_pac_upwind(self, on: bool):
        _check(_lib.nsfv1_set_pac_upwind(self._h, int(bool(on))))

    def prepare_setting(self, lfe_types, orders, non_regular_grid=False):
        """lfe_types: LFEID types per function, as codes (0 = Lagrange) or
        names ("lagrange", "crouzeix-raviart", "piecewise-constant")."""
        names = {"lagrange": 0, "crouzeix-raviart": 1, "piecewise-constant": 2}
        codes = [names.get(str(x).strip().lower(), -1) if isinstance(x, str) else int(x) for x in lfe_types]
        n = len(codes)
        t = (C.c_int * n)(*codes)
        o = (C.c_int * n)(*orders)
        _check(_lib.nsfv1_prepare_setting(self._h, t, o, n, int(bool(non_regular_grid))))

    # --- mesh / pattern ---
    def set_mesh(self, mesh=None, et=None, xyz=None, conn=None):
        if mesh is not None and getattr(mesh, "groups", None) is not None:
            return self._set_mesh_mixed(mesh)
        if mesh is not None:
            et, xyz, conn = mesh.et, mesh.xyz, mesh.conn
            self.mesh = mesh
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        _check(_lib.nsfv1_set_mesh(self._h, int(et), xyz.ctypes.data, xyz.shape[0], conn.ctypes.data, conn.shape[0]))
        self.nv = xyz.shape[0]
        self.ne = conn.shape[0]
        self.dim = 2 if int(et) in (0, 1) else 3
        self._et = int(et)
        self._groups = [(self._et, self.ne)]
        self._mg_levels = []  # a new mesh drops the multigrid hierarchy
        self._bnd_n = [0, 0, 0]  # and the boundary lists
        # functions of position are imports: evaluated again on the new mesh
        if callable(getattr(self, "_nu_fn", None)):
            self._apply_nu_fn()
        if callable(getattr(self, "_f_fn", None)):
            self._apply_f_fn()
        if callable(getattr(self, "_rho_fn", None)):
            self._apply_rho_fn()

    def _set_mesh_mixed(self, mesh):
        """A mesh of several element types (meshes.MixedMesh): groups of
        (element type, connectivity) over the same nodes."""
        xyz = np.ascontiguousarray(mesh.xyz, dtype=np.float64)
        conns = [np.ascontiguousarray(c, dtype=np.int32) for _, c in mesh.groups]
        ng = len(conns)
        ets = (C.c_int * ng)(*[int(et) for et, _ in mesh.groups])
        ptrs = (C.c_void_p * ng)(*[c.ctypes.data for c in conns])
        nes = (C.c_int64 * ng)(*[c.shape[0] for c in conns])
        _check(_lib.nsfv1_set_mesh_mixed(self._h, xyz.ctypes.data, xyz.shape[0], ng, ets, ptrs, nes))
        self.mesh = mesh
        self.nv = xyz.shape[0]
        self.ne = sum(c.shape[0] for c in conns)
        self.dim = xyz.shape[1]
        self._et = int(mesh.groups[0][0])
        self._groups = [(int(et), c.shape[0]) for et, c in mesh.groups]
        self._mg_levels = []  # a new mesh drops the multigrid hierarchy
        self._bnd_n = [0, 0, 0]  # and the boundary lists
        # functions of position are imports: evaluated again on the new mesh
        if callable(getattr(self, "_nu_fn", None)):
            self._apply_nu_fn()
        if callable(getattr(self, "_f_fn", None)):
            self._apply_f_fn()
        if callable(getattr(self, "_rho_fn", None)):
            self._apply_rho_fn()

    def num_dofs(self) -> int:
        n = C.c_int64()
        _check(_lib.nsfv1_num_dofs(self._h, C.byref(n)))
        return n.value

    def nnz(self) -> int:
        n = C.c_int64()
        _check(_lib.nsfv1_nnz(self._h, C.byref(n)))
        return n.value

    def pattern(self):
        rowptr = np.zeros(self.num_dofs() + 1, dtype=np.int64)
        colidx = np.zeros(self.nnz(), dtype=np.int32)
        _check(_lib.nsfv1_get_pattern(self._h, rowptr.ctypes.data, colidx.ctypes.data))
        return rowptr, colidx

    def rowptr(self):
        """Row pointers only (the column indices stay on the device)."""
        rowptr = np.zeros(self.num_dofs() + 1, dtype=np.int64)
        _check(_lib.nsfv1_get_pattern(self._h, rowptr.ctypes.data, None))
        return rowptr

    def pattern_device(self):
        r, c = C.c_void_p(), C.c_void_p()
        _check(_lib.nsfv1_pattern_device(self._h, C.byref(r), C.byref(c)))
        return r.value, c.value

    def set_launch(self, threads=0, lds_bytes=0):
        _check(_lib.nsfv1_set_launch(self._h, int(threads), int(lds_bytes)))

    # --- assembly ---
    def assemble(self, u, u_old=None, dt=0.0, s_a=1.0, s_m=0.0, s_f=0.0, parts=PART_A, jprec=F64,
                 J=None, d=None, want_J=True, want_d=True, stream=None, check=True, noalloc=False):
        """Assemble J = s_a J_A + s_m J_M and d = s_a d_A + s_m d_M + s_f f.

        ``u``/``u_old``: CUDA float64 tensors [nv, dim+1]; returns (J, d).
        ``noalloc`` calls the entry that allocates no device memory (J and d
        must then be given when wanted)."""
        import torch

        dev = u.device
        if noalloc and ((want_J and J is None) or (want_d and d is None)):
            raise ValueError("the no-allocation assembly writes into given J and d only")
        if want_J and J is None:
            J = torch.empty(self.nnz(), dtype=torch.float32 if jprec == F32 else torch.float64, device=dev)
        if want_d and d is None:
            d = torch.empty(self.num_dofs(), dtype=torch.float64, device=dev)
        fn = _lib.nsfv1_assemble_noalloc if noalloc else _lib.nsfv1_assemble
        _check(fn(self._h, _ptr(u), _ptr(u_old), float(dt), float(s_a), float(s_m), float(s_f), int(parts),
                  _ptr(J) if want_J else None, int(jprec), _ptr(d) if want_d else None, _stream_ptr(stream)))
        if check:
            self.sync_errors(stream)
        return (J if want_J else None), (d if want_d else None)

    # --- Newton-loop services ---
    def set_dirichlet(self, dofs):
        dofs = np.ascontiguousarray(dofs, dtype=np.int64)
        self._ndir = len(dofs)
        _check(_lib.nsfv1_set_dirichlet(self._h, dofs.ctypes.data, len(dofs)))

    def apply_dirichlet(self, J=None, jprec=F64, d=None, u=None, g=None, stream=None):
        _check(_lib.nsfv1_apply_dirichlet(self._h, _ptr(J), int(jprec), _ptr(d), _ptr(u), _ptr(g),
                                          _stream_ptr(stream)))

    def spmv(self, J, x, y, jprec=F64, stream=None):
        _check(_lib.nsfv1_spmv(self._h, _ptr(J), int(jprec), _ptr(x), _ptr(y), _stream_ptr(stream)))

    # --- Newton step on the device, its HIP graph, the native autotuner ---
    def _newton_args(self, u, u_old, dt, s_m, dconst, g, J, d, krylov, res2, jprec):
        import torch

        dev = u.device
        if J is None:
            J = torch.empty(self.nnz(), dtype=torch.float32 if jprec == F32 else torch.float64, device=dev)
        if d is None:
            d = torch.empty(self.num_dofs(), dtype=torch.float64, device=dev)
        return (self._h, _ptr(u), _ptr(u_old), float(dt), float(s_m), _ptr(dconst), _ptr(g), _ptr(J), int(jprec),
                _ptr(d), int(krylov), _ptr(res2)), J, d

    def newton_step(self, u, u_old=None, dt=0.0, s_m=0.0, dconst=None, g=None, J=None, d=None, krylov=20,
                    res2=None, jprec=F32, stream=None):
        """One device Newton step (nsfv1_newton_step): u is updated in place;
        returns (J, d).  g: Dirichlet values (device, sorted like the rows
        given to set_dirichlet); res2: optional device float64 scalar."""
        args, J, d = self._newton_args(u, u_old, dt, s_m, dconst, g, J, d, krylov, res2, jprec)
        _check(_lib.nsfv1_newton_step(*args, _stream_ptr(stream)))
        return J, d

    def newton_graph_capture(self, u, u_old=None, dt=0.0, s_m=0.0, dconst=None, g=None, J=None, d=None,
                             krylov=20, res2=None, jprec=F32, stream=None):
        """Capture the Newton step into the handle's HIP graph (stream must
        not be the null stream).  The tensors are kept referenced for the
        replays."""
        args, J, d = self._newton_args(u, u_old, dt, s_m, dconst, g, J, d, krylov, res2, jprec)
        _check(_lib.nsfv1_newton_graph_capture(*args, _stream_ptr(stream)))
        self._graph_keep = (u, u_old, dconst, g, J, d, res2)
        return J, d

    def newton_graph_replay(self, stream=None):
        _check(_lib.nsfv1_newton_graph_replay(self._h, _stream_ptr(stream)))

    def autotune_native(self, u, u_old=None, dt=0.0, s_m=0.0, parts=PART_A, jprec=F32, J=None, d=None, reps=3,
                        stream=None):
        """The C++ launch-shape autotuner (nsfv1_autotune): returns ((path,
        threads, lds), [(path, threads, lds, ms), ...]); the best shape is
        kept on the handle."""
        import torch

        if J is None:
            J = torch.empty(self.nnz(), dtype=torch.float32 if jprec == F32 else torch.float64, device=u.device)
        if d is None:
            d = torch.empty(self.num_dofs(), dtype=torch.float64, device=u.device)
        cap = 64
        table = (C.c_double * (4 * cap))()
        nt = C.c_int(cap)
        best = (C.c_int * 3)()
        _check(_lib.nsfv1_autotune(self._h, _ptr(u), _ptr(u_old), float(dt), float(s_m), int(parts), _ptr(J),
                                   int(jprec), _ptr(d), int(reps), table, C.byref(nt), best, _stream_ptr(stream)))
        rows = [(int(table[4 * i]), int(table[4 * i + 1]), int(table[4 * i + 2]), float(table[4 * i + 3]))
                for i in range(min(nt.value, cap))]
        return (best[0], best[1], best[2]), rows

    def linear_solve(self, J, b, x, jprec=F64, rtol=1e-6, maxit=500, restart=30, stream=None):
        it = C.c_int(0)
        rr = C.c_double(0.0)
        _check(_lib.nsfv1_linear_solve(self._h, _ptr(J), int(jprec), _ptr(b), _ptr(x), float(rtol), int(maxit),
                                       int(restart), C.byref(it), C.byref(rr), _stream_ptr(stream)))
        return it.value, rr.value

    def set_outflow(self, elem, side):
        """No-normal-stress outflow on the (element, local side) list."""
        e = np.ascontiguousarray(elem, dtype=np.int32)
        s = np.ascontiguousarray(side, dtype=np.int32)
        _check(_lib.nsfv1_set_outflow(self._h, e.ctypes.data, s.ctypes.data, len(e)))
        self._bnd_count()[0] = len(e)
        self._apply_bnd_nu(0)

    def set_symmetry(self, elem, side):
        """Symmetry condition (NavierStokesSymBCFV1) on the (element, local side) list."""
        e = np.ascontiguousarray(elem, dtype=np.int32)
        s = np.ascontiguousarray(side, dtype=np.int32)
        _check(_lib.nsfv1_set_symmetry(self._h, e.ctypes.data, s.ctypes.data, len(e)))
        self._bnd_count()[1] = len(e)
        self._apply_bnd_nu(1)

    def _bnd_count(self):
        if not hasattr(self, "_bnd_n"):
            self._bnd_n = [0, 0, 0]
        return self._bnd_n

    def bnd_ip_positions(self, kind: int) -> np.ndarray:
        """Boundary-face ips of list `kind` (0 outflow, 1 symmetry, 2 sliding
        wall) in input order: [sides, 4, dim] (unused slots 0)."""
        out = np.zeros((self._bnd_count()[kind], 4, self.dim))
        _check(_lib.nsfv1_bnd_ip_positions(self._h, int(kind), out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def set_boundary_viscosity(self, kind: int, values):
        """The viscosity at the boundary-face ips of list `kind` ([sides, 4]);
        None clears."""
        if values is None:
            _check(_lib.nsfv1_set_boundary_viscosity(self._h, int(kind), None, 0))
            return
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(_lib.nsfv1_set_boundary_viscosity(self._h, int(kind), v.ctypes.data_as(C.POINTER(C.c_double)), v.size))

    def _apply_bnd_nu(self, kind: int):
        """A viscosity function evaluated at the outflow / symmetry faces."""
        fn = getattr(self, "_nu_fn", None)
        if not callable(fn) or kind == 2 or self._bnd_count()[kind] == 0:
            return
        x = self.bnd_ip_positions(kind)
        self.set_boundary_viscosity(kind, np.asarray(fn(x.reshape(-1, self.dim)), dtype=np.float64).reshape(x.shape[:2]))

    def set_sliding_wall(self, elem, side, sliding_factor: float, sliding_limit: float = 0.0):
        """Sliding wall (NavierStokesWSBCFV1) on the (element, local side) list."""
        e = np.ascontiguousarray(elem, dtype=np.int32)
        s = np.ascontiguousarray(side, dtype=np.int32)
        _check(_lib.nsfv1_set_sliding_wall(self._h, e.ctypes.data, s.ctypes.data, len(e), float(sliding_factor),
                                           float(sliding_limit)))
        self._bnd_count()[2] = len(e)

    def export(self, u, what: str, at: str = "scvf", local_ips=None, u_old=None, dt=0.0, out=None, stream=None,
               check=True):
        """Data export ("velocity", "velocity_grad", "velocity_ip", "pressure",
        "pressure_grad") at the SCVF ips, the corners ("scv") or the given
        local points of every element: a CUDA float64 tensor
        [ne, points, components] ([points, components] on a mixed mesh,
        group after group)."""
        import torch

        lips = None if local_ips is None else np.ascontiguousarray(local_ips, dtype=np.float64)
        nip = 0 if lips is None else lips.shape[0]
        cnt = C.c_int64(0)
        _check(_lib.nsfv1_export_size(self._h, EX[what], AT[at], nip, C.byref(cnt)))
        if out is None:
            out = torch.empty(cnt.value, dtype=torch.float64, device=u.device)
        _check(_lib.nsfv1_export(self._h, EX[what], AT[at], None if lips is None else lips.ctypes.data, nip, _ptr(u),
                                 _ptr(u_old), float(dt), _ptr(out), _stream_ptr(stream)))
        if check:
            self.sync_errors(stream)
        ncomp = self.dim * self.dim if what == "velocity_grad" else (1 if what == "pressure" else self.dim)
        if len(getattr(self, "_groups", [])) > 1:
            return out.view(-1, ncomp)  # mixed mesh: points group after group
        return out.view(self.ne, -1, ncomp)

    def set_assembly_path(self, path: int):
        _check(_lib.nsfv1_set_assembly_path(self._h, int(path)))

    def assembly_path_used(self) -> int:
        """Kernel family of the last assembly: 3 plane march, 0 split, 2 split
        with row lanes, 1 fused, 4 coupled, -1 none yet."""
        p = C.c_int(-1)
        _check(_lib.nsfv1_assembly_path_used(self._h, C.byref(p)))
        return p.value

    def batch_set_instances(self, elem_offset, nu):
        off = np.ascontiguousarray(elem_offset, dtype=np.int64)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        _check(_lib.nsfv1_batch_set_instances(self._h, len(nu), off.ctypes.data, nu.ctypes.data))

    def set_upwind_decisions(self, mode: int):
        _check(_lib.nsfv1_set_upwind_decisions(self._h, int(mode)))

    def set_preconditioner(self, name: str):
        _check(_lib.nsfv1_set_preconditioner(self._h, name.encode()))

    def mg_set_hierarchy(self, coarse, prolongations, injections, pre=2, post=2, omega=0.7):
        """coarse: list of NavierStokesFV1 (meshes and Dirichlet rows set),
        prolongations: list of (rowptr int64, colidx int32, weights float64)
        node-level CSR per level pair, injections: int64 arrays."""
        nl = len(coarse) + 1
        Hs = (C.c_void_p * (nl - 1))(*[c._h.value for c in coarse])
        keep = []
        def arr(xs, dt):
            out = []
            for x in xs:
                a = np.ascontiguousarray(x, dtype=dt)
                keep.append(a)
                out.append(a.ctypes.data)
            return (C.c_void_p * (nl - 1))(*out)
        rp = arr([p[0] for p in prolongations], np.int64)
        ci = arr([p[1] for p in prolongations], np.int32)
        w = arr([p[2] for p in prolongations], np.float64)
        ij = arr(injections, np.int64)
        _check(_lib.nsfv1_mg_set_hierarchy(self._h, nl, Hs, rp, ci, w, ij, int(pre), int(post), float(omega)))
        self._mg_levels = list(coarse)  # the library borrows the coarse handles
        # functions of position are imported on every level
        for c in self._mg_levels:
            c._inherit_functions(self)

    def _inherit_functions(self, fine):
        """Evaluate the fine level's viscosity / source functions on this
        (coarse) level's integration points."""
        nu_fn, f_fn = getattr(fine, "_nu_fn", None), getattr(fine, "_f_fn", None)
        if callable(nu_fn):
            self.set_kinematic_viscosity(nu_fn)
        if callable(f_fn):
            self.set_source(f_fn)
        rho_fn = getattr(fine, "_rho_fn", None)
        if callable(rho_fn):
            self.set_density(rho_fn)

    def mg_setup(self, jprec, u, u_old=None, dt=0.0, s_a=1.0, s_m=0.0, parts=PART_A, stream=None):
        _check(_lib.nsfv1_mg_setup(self._h, int(jprec), _ptr(u), _ptr(u_old), float(dt), float(s_a), float(s_m),
                                   int(parts), _stream_ptr(stream)))

    def sync_errors(self, stream=None):
        _check(_lib.nsfv1_sync_errors(self._h, _stream_ptr(stream)))


def NavierStokes(fcts, subsets, disc_type: str = "fv1"):
    """``NavierStokes(fcts, subsets, discType)`` of lua/lua-include.lua:36-47,
    restricted to the FV1 discretisation built here."""
    if disc_type is None:
        disc_type = "fv1"
    if disc_type != "fv1":
        raise NSFV1Error(-11, f"NavierStokes: no disc type '{disc_type}' available in this build")
    return NavierStokesFV1(fcts, subsets)
