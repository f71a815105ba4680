"""Host-side mirror of pdekit's explicit continuous RD method.

``PGRDMethodExplicit`` keeps the member names of the reference class
(src/solver/rdm/PGRDMethodExplicit.hpp:19-193 over the RDMethod API of
src/solver/rdm/RDMethod.hpp:33-139) and runs every step on the GPU through
the C ABI of include/pdekit_rd.h. Nodal fields are numpy arrays in the
reference's AoS layout [N][m].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .fe import element_tables
from .mesh import Mesh

PHYSICS = {"Advection2D": L.ADVECTION2D, "Burgers2D": L.BURGERS2D, "Euler2DCons": L.EULER2D,
           "Euler3DCons": L.EULER3D}
SCHEMES = {"PGLDA": L.LDA, "PGN": L.N, "PGB": L.B, "PSI": L.PSI, "BPSI": L.BPSI}
NEQ = {L.ADVECTION2D: 1, L.BURGERS2D: 1, L.EULER2D: 4, L.EULER3D: 5, L.ROTATION2D: 1}
FIELDS = {"solution": L.SOLUTION, "residuals": L.RESIDUAL, "time_step": L.TIME_STEP,
          "wave_speed": L.WAVE_SPEED, "dual_volume": L.DUAL_VOLUME, "blending_coeff": L.BLEND_COEFF,
          "artificial_viscosity": L.ART_VISC}


class PGRDMethodExplicit:
    """One explicit continuous RD method on the current GPU.

    physics: 'Advection2D' (constant velocity params=(ax, ay)), 'Burgers2D',
    'Euler2DCons', 'Euler3DCons' (params=(gamma, R)); scheme: 'PGLDA', 'PGN',
    'PGB', 'PSI' or 'BPSI' (the PSI / LDA blend by PGB's cell theta).
    """

    def __init__(self, physics: str, scheme: str, params=None):
        self.physics = PHYSICS[physics] if isinstance(physics, str) else int(physics)
        self.scheme = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        self.params = np.zeros(4)
        if params is not None:
            self.params[:len(params)] = params
            self.nb_params = len(params)
        else:
            self.nb_params = 0
        self.neq = NEQ[self.physics]
        self._raw = None     # C handle; pdrd_initialize runs at the first use that needs the device
        self._ready = False
        self.quad_order = 4
        self.mesh: Mesh | None = None
        self.dirichlet_nodes = np.zeros(0, dtype=np.uint32)

    # -- RDMethod::configure_mesh_data / initialize_work_data ---------------------------
    def configure_mesh_data(self, mesh: Mesh):
        self.mesh = mesh
        return self

    def initialize_work_data(self, quad_order=4, point_set=L.WARPBLEND, scatter=None, patch_cells=None):
        """Builds the FE tables of every element type (one quadrature order, or a
        dict shape -> order) and uploads mesh and tables. scatter: None (the
        library's default, patch assembly), 0 FP64 atomics, 1 colours, 2 patches.
        patch_cells: None (default), 64 / 128 / 256 cells per patch workgroup."""
        if self.mesh is None:
            raise RuntimeError("configure_mesh_data first")
        m = self.mesh
        lib = L.lib()
        h = C.c_void_p()
        coords = np.ascontiguousarray(m.coords, dtype=np.float64)
        L.check(lib.pdrd_create(m.dim, coords.shape[0], L.ptr(coords), m.nb_nodes, self.physics, self.scheme,
                                L.ptr(self.params), self.nb_params, C.byref(h)), "pdrd_create")
        self._raw = h
        self._ready = False
        self.quad_order = max(quad_order.values()) if isinstance(quad_order, dict) else quad_order
        self.tables = {}
        for blk in m.blocks:
            q = quad_order[blk.shape] if isinstance(quad_order, dict) else quad_order
            t = element_tables(blk.shape, m.order, q, 1, point_set)
            self.tables[blk.shape] = t
            L.check(lib.pdrd_add_cells(h, C.byref(t.desc), L.ptr(t.packed), blk.geo_conn.shape[0],
                                       L.ptr(blk.geo_conn, C.c_uint32), L.ptr(blk.sol_conn, C.c_uint32)),
                    "pdrd_add_cells")
        if scatter is not None:
            L.check(lib.pdrd_set_option(h, L.OPT_SCATTER, int(scatter)), "pdrd_set_option")
        if patch_cells is not None:
            L.check(lib.pdrd_set_option(h, L.OPT_PATCH_CELLS, int(patch_cells)), "pdrd_set_option")
        return self

    @property
    def _h(self):
        """The C handle, initialised (device upload, colouring, patch plan) on
        first use, so that weak boundary conditions can still be added after
        initialize_work_data as in the reference's call order."""
        if self._raw is None:
            raise RuntimeError("initialize_work_data first")
        if not self._ready:
            L.check(L.lib().pdrd_initialize(self._raw), "pdrd_initialize")
            self._ready = True
        return self._raw

    def set_option(self, option: int, value: int):
        L.check(L.lib().pdrd_set_option(self._raw, option, int(value)), "pdrd_set_option")

    def get_option(self, option: int) -> int:
        v = C.c_int(0)
        L.check(L.lib().pdrd_get_option(self._h, option, C.byref(v)), "pdrd_get_option")
        return v.value

    def __del__(self):
        h = getattr(self, "_raw", None)
        if h is not None and h.value:
            L.lib().pdrd_destroy(h)
            self._raw = None

    def destroy(self):
        self.__del__()

    # -- vec functions ------------------------------------------------------------------
    def _shape(self, which: int):
        n = self.mesh.nb_nodes
        return (n, self.neq) if which in (L.SOLUTION, L.RESIDUAL) else (n,)

    def set_vec_function(self, which, values):
        """'solution', 'residuals', ... (a device field), or 'sources': the
        nodal source field [N][m] of the reaction residuals (None removes it)."""
        if which == "sources":
            self.set_sources(values)
            return
        w = FIELDS[which] if isinstance(which, str) else which
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), self._shape(w)))
        L.check(L.lib().pdrd_set_field(self._h, w, L.ptr(a)), "pdrd_set_field")

    def set_sources(self, src):
        """set_vec_function(sources, f): nodal source field [N][m]; None
        switches it off."""
        if src is None:
            L.check(L.lib().pdrd_set_sources(self._h, None), "pdrd_set_sources")
            return
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(src, dtype=np.float64), (self.mesh.nb_nodes, self.neq)))
        L.check(L.lib().pdrd_set_sources(self._h, L.ptr(a)), "pdrd_set_sources")

    def vec_function(self, which) -> np.ndarray:
        w = FIELDS[which] if isinstance(which, str) else which
        a = np.zeros(self._shape(w))
        L.check(L.lib().pdrd_get_field(self._h, w, L.ptr(a)), "pdrd_get_field")
        return a

    # -- boundary conditions --------------------------------------------------------------
    WEAK = {"WeakWall": L.WEAK_WALL, "WeakSubInlet": L.WEAK_SUBINLET, "WeakSubOutlet": L.WEAK_SUBOUTLET,
            "WeakFarfield": L.WEAK_FARFIELD, "WeakSymmetry": L.WEAK_SYMMETRY, "WeakSuperInlet": L.WEAK_SUPERINLET,
            "WeakDirichlet": L.WEAK_DIRICHLET}

    def add_boundary_condition(self, bc_type: str, name: str, nodes, values=None, params=None):
        """'StrongDirichlet': values [n][m] or a callable f(node_coords[n][dim]) -> [n][m].
        Weak conditions, added before the first use of the method's device data:
        'WeakWall', 'WeakSymmetry', 'WeakSuperInlet' (no parameters),
        'WeakSubInlet' (params: total_temperature, total_pressure),
        'WeakSubOutlet' (params: total_pressure), 'WeakFarfield' (values: the
        reference state, m values) for the Euler equations; 'WeakDirichlet'
        (values: the boundary state, m values) for every physics."""
        if bc_type in self.WEAK:
            nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
            if bc_type in ("WeakFarfield", "WeakDirichlet"):
                if values is None:
                    raise ValueError(f"{bc_type} needs values: the boundary state")
                par = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
                if par.size != self.neq:
                    raise ValueError(f"{bc_type} takes one state of {self.neq} values")
                npar = par.size
            else:
                params = params or {}
                keys = {"WeakWall": [], "WeakSymmetry": [], "WeakSuperInlet": [],
                        "WeakSubInlet": ["total_temperature", "total_pressure"],
                        "WeakSubOutlet": ["total_pressure"]}[bc_type]
                par = np.array([float(params[k]) for k in keys] + [0.0, 0.0])
                npar = len(keys)
            if self._ready:
                raise RuntimeError("weak boundary conditions are added before the method's first use")
            L.check(L.lib().pdrd_add_weak_bc(self._raw, self.WEAK[bc_type], nodes.size, L.ptr(nodes, C.c_uint32),
                                             L.ptr(par), npar, int(self.quad_order)), "pdrd_add_weak_bc")
            return
        if bc_type != "StrongDirichlet":
            raise NotImplementedError(f"boundary condition {bc_type} is not built on the GPU path")
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        if callable(values):
            values = values(self.mesh.node_coords[nodes])
        vals = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64),
                                                    (nodes.size, self.neq)))
        L.check(L.lib().pdrd_add_strong_dirichlet(self._h, nodes.size, L.ptr(nodes, C.c_uint32), L.ptr(vals)),
                "pdrd_add_strong_dirichlet")
        self.dirichlet_nodes = np.union1d(self.dirichlet_nodes, nodes).astype(np.uint32)

    # -- the iteration -----------------------------------------------------------------------
    def assemble_rhs(self):
        L.check(L.lib().pdrd_assemble_rhs(self._h), "pdrd_assemble_rhs")

    def assemble_lhs_and_rhs(self, cfl: float):
        L.check(L.lib().pdrd_assemble_lhs_and_rhs(self._h, float(cfl)), "pdrd_assemble_lhs_and_rhs")

    def solve(self, solver_options=None):
        L.check(L.lib().pdrd_solve(self._h), "pdrd_solve")

    def compute_residual_norm(self) -> np.ndarray:
        out = np.zeros(self.neq)
        L.check(L.lib().pdrd_compute_residual_norm(self._h, L.ptr(out)), "pdrd_compute_residual_norm")
        return out

    def residual_norm_L2(self, exclude_dirichlet: bool = False) -> np.ndarray:
        out = np.zeros(self.neq)
        L.check(L.lib().pdrd_residual_norm_L2(self._h, int(exclude_dirichlet), L.ptr(out)), "pdrd_residual_norm_L2")
        return out

    def iterate(self, cfl: float, n_iter: int, exclude_dirichlet: bool = False) -> np.ndarray:
        out = np.zeros(self.neq)
        L.check(L.lib().pdrd_iterate(self._h, float(cfl), int(n_iter), int(exclude_dirichlet), L.ptr(out)),
                "pdrd_iterate")
        return out

    def check_status(self):
        L.check(L.lib().pdrd_check_status(self._h), "pdrd_check_status")

    def synchronize(self):
        L.check(L.lib().pdrd_synchronize(self._h), "pdrd_synchronize")

    def info(self) -> dict:
        a = np.zeros(6, dtype=np.int64)
        L.check(L.lib().pdrd_method_info(self._h, L.ptr(a, C.c_int64)), "pdrd_method_info")
        return dict(nb_nodes=int(a[0]), neq=int(a[1]), nb_blocks=int(a[2]), nb_cells=int(a[3]),
                    nb_colours=int(a[4]), device_bytes=int(a[5]))

    # -- nodal inputs of the cell loop ------------------------------------------------------
    def set_blending_coeff(self, theta):
        """set_blending_coeff: nodal theta [N] that the B scheme uses instead
        of its cell-wise theta (PGB.hpp:141-148, 467-477); None switches it off."""
        if theta is None:
            L.check(L.lib().pdrd_set_blending_coeff(self._h, None), "pdrd_set_blending_coeff")
            return
        t = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(self.mesh.nb_nodes))
        L.check(L.lib().pdrd_set_blending_coeff(self._h, L.ptr(t)), "pdrd_set_blending_coeff")

    def set_artificial_viscosity(self, nu):
        """set_artificial_viscosity: nodal nu [N] of the LDA scheme's
        shock-capturing term (PGLDA.hpp:245-256); None switches it off."""
        if nu is None:
            L.check(L.lib().pdrd_set_artificial_viscosity(self._h, None), "pdrd_set_artificial_viscosity")
            return
        v = np.ascontiguousarray(np.asarray(nu, dtype=np.float64).reshape(self.mesh.nb_nodes))
        L.check(L.lib().pdrd_set_artificial_viscosity(self._h, L.ptr(v)), "pdrd_set_artificial_viscosity")

    def compute_blending_coeff(self, kind: str = "LF_Blend") -> np.ndarray:
        """RDBlendingCoeff::calculate(kind) on the device from the current
        solution, kept as the method's blending coefficient; returns it."""
        if kind != "LF_Blend":
            raise NotImplementedError(f"blending coefficient {kind} is not built on the GPU path")
        L.check(L.lib().pdrd_compute_blending_coeff(self._h, L.BLEND_LF), "pdrd_compute_blending_coeff")
        return self.vec_function("blending_coeff")

    def compute_artificial_viscosity(self, kappa: float = 5.0, factor: float = 1.0) -> np.ndarray:
        """ArtificialViscosity::compute_artificial_viscosity_nodal on the
        device from the current solution, kept as the method's artificial
        viscosity; returns it."""
        L.check(L.lib().pdrd_compute_artificial_viscosity(self._h, float(kappa), float(factor)),
                "pdrd_compute_artificial_viscosity")
        return self.vec_function("artificial_viscosity")

    def element_jacobians(self, block: int = 0):
        """compute_adv_res_and_jacobian of every cell of one block (its cells
        in mesh order): phi[ne][nd][m], jac[ne][nd*m][nd*m], wave[ne]."""
        blk = self.mesh.blocks[block]
        ne, nd = blk.sol_conn.shape
        phi = np.zeros((ne, nd, self.neq))
        jac = np.zeros((ne, nd * self.neq, nd * self.neq))
        wave = np.zeros(ne)
        L.check(L.lib().pdrd_element_jacobians(self._h, int(block), L.ptr(phi), L.ptr(jac), L.ptr(wave)),
                "pdrd_element_jacobians")
        return phi, jac, wave

    def weak_bc_info(self, k: int) -> dict:
        a = np.zeros(3, dtype=np.int64)
        L.check(L.lib().pdrd_weak_bc_info(self._h, int(k), L.ptr(a, C.c_int64)), "pdrd_weak_bc_info")
        return dict(facets=int(a[0]), nodes_per_facet=int(a[1]), quadrature_points=int(a[2]))

    def weak_bc_points(self, k: int) -> np.ndarray:
        """Coordinates [nf][nq][dim] of the facet points of weak BC k."""
        info = self.weak_bc_info(k)
        xq = np.zeros((info["facets"], info["quadrature_points"], self.mesh.dim))
        L.check(L.lib().pdrd_weak_bc_points(self._h, int(k), L.ptr(xq)), "pdrd_weak_bc_points")
        return xq

    def set_weak_bc_state(self, k: int, state):
        """WeakDirichlet k's boundary state at every facet point, [nf][nq][m]
        (e.g. an expression evaluated at weak_bc_points(k)); None restores the
        constant state."""
        if state is None:
            L.check(L.lib().pdrd_set_weak_bc_state(self._h, int(k), None, 0), "pdrd_set_weak_bc_state")
            return
        info = self.weak_bc_info(k)
        shape = (info["facets"], info["quadrature_points"], self.neq)
        a = np.asarray(state, dtype=np.float64)
        if a.shape != shape and not (self.neq == 1 and a.shape == shape[:2]):
            raise ValueError(f"weak BC {k} takes a state of shape {shape} ([facets][points][m]), got {a.shape}")
        a = np.ascontiguousarray(a.reshape(shape))
        L.check(L.lib().pdrd_set_weak_bc_state(self._h, int(k), L.ptr(a), a.size), "pdrd_set_weak_bc_state")

    def last_timing(self):
        c = C.c_double(0)
        n = C.c_double(0)
        L.check(L.lib().pdrd_last_timing(self._h, C.byref(c), C.byref(n)), "pdrd_last_timing")
        return c.value, n.value


class Batch:
    """Iterations of several independent methods in shared launches
    (pdrd_batch_iterate): one launch per element type / physics / scheme and
    iteration for the patches of all methods, one for their shared nodes.
    Every method advances exactly as its own iterate() would. graphs: replay
    the iterations of a call from a captured HIP graph instead of launching
    them directly."""

    def __init__(self, graphs: bool = False):
        h = C.c_void_p()
        L.check(L.lib().pdrd_batch_create(C.byref(h)), "pdrd_batch_create")
        self._h = h
        L.check(L.lib().pdrd_batch_set_graphs(h, int(graphs)), "pdrd_batch_set_graphs")

    def iterate(self, methods, cfl, n_iter: int, exclude_dirichlet: bool = False):
        """methods: PGRDMethodExplicit list; cfl: one value or one per method.
        Returns the norm_L2 of the last residual of every method."""
        n = len(methods)
        hs = (C.c_void_p * n)(*[m._h.value for m in methods])
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(cfl, dtype=np.float64), (n,)))
        norms = np.zeros(sum(m.neq for m in methods))
        L.check(L.lib().pdrd_batch_iterate(self._h, n, hs, L.ptr(c), int(n_iter), int(exclude_dirichlet),
                                           L.ptr(norms)), "pdrd_batch_iterate")
        out, off = [], 0
        for m in methods:
            out.append(norms[off:off + m.neq].copy())
            off += m.neq
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().pdrd_batch_destroy(h)
            self._h = None
