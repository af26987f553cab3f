"""Host side of the projection step and the interface transport, over the
simcogpu C ABI.

``Projection`` mirrors the reference's ``PoissonEquationSolver``
(ProjectionP_Mod.f90:26-99), the correction part of ``UpdatePUV``
(ComputePUV_Mod.f90:17-225) and ``Coupled_LS_VOF`` (Clsvof_Mod.f90:730-959):
same inputs (predictor velocities, PU%Dp / PV%Dp, the P-cell vof / phi,
cut-cell apertures, TVar%u / v), same outputs (Proj%Pp with its ghost cells;
corrected TVar%u / v / p and the mass residual; advected PCell%vof / phi /
nx / ny).  Arrays are
numpy arrays in the reference's Fortran layout: a cell array has shape
(nj, ni) in C order, a ghosted array (nj+2, ni+2).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import F, Params, SolveInfo, check

_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_DP)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_IP)


def default_params(**kw):
    p = Params()
    check(_lib.load().sg_params_default(ctypes.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def params_for(fluids, tol=1e-10, precision=_lib.SG_PREC_FP64, bc=(1, 1, 1, 0), **kw):
    """Parameters for a cases.Fluids description."""
    return default_params(rho_l=fluids.rho_l, rho_g=fluids.rho_g, rho_ref=fluids.rho_ref,
                          nu_l=fluids.mu_l / fluids.rho_l, nu_g=fluids.mu_g / fluids.rho_g,
                          nu_ref=fluids.mu_l / fluids.rho_l, bc_w=bc[0], bc_e=bc[1], bc_s=bc[2], bc_n=bc[3],
                          tol=tol, precision=precision, **kw)


class Projection:
    """A device context for one grid (all device memory allocated here)."""

    def __init__(self, grid, params):
        self.lib = _lib.load()
        self.grid = grid
        self.ni, self.nj = grid.ni, grid.nj
        self.params = params
        ctx = ctypes.c_void_p()
        keep = [_d(grid.dx), _d(grid.dy), _d(grid.udx), _d(grid.vdy)]
        check(self.lib.sg_ctx_create(self.ni, self.nj, keep[0][1], keep[1][1], keep[2][1], keep[3][1],
                                     ctypes.byref(params), ctypes.byref(ctx)))
        self.ctx = ctx
        if getattr(grid, "xc", None) is not None:
            xc, yc = _d(grid.xc), _d(grid.yc)
            check(self.lib.sg_set_centres(self.ctx, xc[1], yc[1]))

    # -- lifetime --------------------------------------------------------
    def close(self):
        if self.ctx:
            self.lib.sg_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, params):
        self.params = params
        check(self.lib.sg_set_params(self.ctx, ctypes.byref(params)))

    # -- fields ----------------------------------------------------------
    def _shape(self, name):
        if name in _lib.GHOST_FIELDS:
            return (self.nj + 2, self.ni + 2)
        return (self.nj, self.ni)

    def upload(self, name, arr):
        shape = self._shape(name)
        if name in _lib.INT_FIELDS:
            a = np.ascontiguousarray(arr, dtype=np.int32)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        check(self.lib.sg_upload(self.ctx, F[name], a.ctypes.data_as(ctypes.c_void_p)))

    def download(self, name):
        dt = np.int32 if name in _lib.INT_FIELDS else np.float64
        a = np.empty(self._shape(name), dtype=dt)
        check(self.lib.sg_download(self.ctx, F[name], a.ctypes.data_as(ctypes.c_void_p)))
        return a

    def field_ptr(self, name):
        return self.lib.sg_field_ptr(self.ctx, F[name])

    def set_geometry(self, geo):
        args = [_d(geo.vofS), _d(geo.nyS), _d(geo.WlLh), _d(geo.aE), _d(geo.aW), _d(geo.aN), _d(geo.aS)]
        ints = [_i(geo.uflag), _i(geo.vflag)]
        tail = [_d(geo.vofSU), _d(geo.vofSV)]
        check(self.lib.sg_set_geometry(self.ctx, *[a[1] for a in args], *[a[1] for a in ints],
                                       *[a[1] for a in tail]))
        # the solid line of the three-phase cells (interface transport)
        if getattr(geo, "nxS", None) is not None:
            self.upload("NXS", geo.nxS)
            self.upload("PHIS", geo.phiS)

    def set_interface(self, vof, phi):
        self.upload("VOF", vof)
        self.upload("PHI", phi)

    def set_predictor(self, upred, vpred, DpU, DpV):
        self.upload("UPRED", upred)
        self.upload("VPRED", vpred)
        self.upload("DPU", DpU)
        self.upload("DPV", DpV)

    # -- projection --------------------------------------------------------
    def build(self, vb=0.0):
        check(self.lib.sg_poisson_build_resident(self.ctx, float(vb)))

    def solve(self, allow_not_converged=False):
        info = SolveInfo()
        rc = self.lib.sg_poisson_solve_resident(self.ctx, ctypes.byref(info))
        check(rc, allow=(_lib.SG_NOT_CONVERGED,) if allow_not_converged else ())
        return info

    def poisson_solve(self, u_star, v_star, DpU, DpV, vof, phi, vb=0.0):
        """Drop-in PoissonEquationSolver: host arrays in, Proj%Pp out."""
        ins = [_d(a) for a in (u_star, v_star, DpU, DpV, vof, phi)]
        Pp = np.zeros((self.nj + 2, self.ni + 2))
        info = SolveInfo()
        rc = self.lib.sg_poisson_solve(self.ctx, *[a[1] for a in ins], float(vb), Pp.ctypes.data_as(_DP),
                                       ctypes.byref(info))
        check(rc)
        return Pp, info

    def matrix(self):
        """Assembled rows in the reference's matr layout, shape (5, nj, ni)."""
        m = np.zeros((5, self.nj, self.ni))
        check(self.lib.sg_poisson_matrix(self.ctx, m.ctypes.data_as(_DP)))
        return m

    def rhs(self):
        b = np.zeros((self.nj, self.ni))
        check(self.lib.sg_poisson_rhs(self.ctx, b.ctypes.data_as(_DP)))
        return b

    def apply(self, x):
        x, xp = _d(x)
        y = np.zeros((self.nj, self.ni))
        check(self.lib.sg_poisson_apply(self.ctx, xp, y.ctypes.data_as(_DP)))
        return y

    def velocity_correct(self, vb=0.0):
        check(self.lib.sg_velocity_correct_resident(self.ctx, float(vb)))

    def synchronize(self):
        check(self.lib.sg_synchronize(self.ctx))

    # -- interface transport -------------------------------------------------
    def clsvof_params(self, nsub=6, wave=None):
        p = _lib.ClsvofParams()
        check(self.lib.sg_clsvof_params_default(ctypes.byref(p)))
        p.nsub = nsub
        if wave is not None:
            p.depth, p.amp, p.kw, p.omega = wave
        return p

    def clsvof_advect_resident(self, dt, time=0.0, nsub=6, wave=None):
        """Coupled_LS_VOF on the device fields VOF, PHI (in/out), NX, NY (out)."""
        p = self.clsvof_params(nsub, wave)
        check(self.lib.sg_clsvof_advect_resident(self.ctx, float(dt), float(time), ctypes.byref(p)))

    def clsvof_advect(self, vof, phi, u, v, dt, time=0.0, nsub=6, wave=None):
        """Drop-in Coupled_LS_VOF: host arrays in, (vof, phi, nx, ny) out."""
        vof = np.array(vof, dtype=np.float64, order="C")
        phi = np.array(phi, dtype=np.float64, order="C")
        nx = np.zeros_like(vof)
        ny = np.zeros_like(vof)
        u, up = _d(u)
        v, vp = _d(v)
        if vof.shape != (self.nj, self.ni) or u.shape != (self.nj + 2, self.ni + 2) or v.shape != u.shape:
            raise ValueError("clsvof_advect: vof/phi must be (nj, ni), u/v (nj+2, ni+2)")
        p = self.clsvof_params(nsub, wave)
        check(self.lib.sg_clsvof_advect(self.ctx, vof.ctypes.data_as(_DP), phi.ctypes.data_as(_DP),
                                        nx.ctypes.data_as(_DP), ny.ctypes.data_as(_DP), up, vp, float(dt),
                                        float(time), ctypes.byref(p)))
        return vof, phi, nx, ny

    # -- momentum predictor ----------------------------------------------------
    def set_staggered_geometry(self, which, st):
        """U (which=0) or V (which=1) cell geometry (a cases.StagGeometry)."""
        keep = []
        cell = []
        for k in _lib.STAG_CELL:
            a, ptr = _d(getattr(st, k))
            keep.append(a)
            cell.append(ptr)
        ints = []
        for k in _lib.STAG_INT:
            a, ptr = _i(getattr(st, k))
            keep.append(a)
            ints.append(ptr)
        gh = []
        for k in _lib.STAG_GHOST:
            a, ptr = _d(getattr(st, k))
            keep.append(a)
            gh.append(ptr)
        g = _lib.StagGeometry(*cell, *ints, *gh)
        check(self.lib.sg_set_staggered_geometry(self.ctx, int(which), ctypes.byref(g)))

    def predictor_resident(self, dt, vb=0.0, first_step=False):
        info = _lib.PredictorInfo()
        check(self.lib.sg_predictor_resident(self.ctx, float(dt), float(vb), int(bool(first_step)),
                                             ctypes.byref(info)))
        return info

    def predictor_rows(self):
        mu = np.zeros((5, self.nj, self.ni)); ru = np.zeros((self.nj, self.ni))
        mv = np.zeros((5, self.nj, self.ni)); rv = np.zeros((self.nj, self.ni))
        check(self.lib.sg_predictor_rows(self.ctx, mu.ctypes.data_as(_DP), ru.ctypes.data_as(_DP),
                                         mv.ctypes.data_as(_DP), rv.ctypes.data_as(_DP)))
        return mu, ru, mv, rv

    # -- Lagrangian particles (TrackingParticles, Particles.f90:71-205) ----
    def particles_set(self, x, y, u, v, d):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, u, v, d)]
        n = len(arrs[0])
        if any(len(a) != n for a in arrs):
            raise ValueError("particle arrays must have one length")
        self._npart = n
        check(self.lib.sg_particles_set(self.ctx, n, *[_d(a)[1] for a in arrs]))

    def particles_get(self):
        n = getattr(self, "_npart", 0)
        out = [np.zeros(n) for _ in range(4)]
        check(self.lib.sg_particles_get(self.ctx, n, *[_d(a)[1] for a in out]))
        return tuple(out)

    def particles_track(self, dt, params):
        """params: dict of sg_particle_params fields; returns the number of
        particles that met the reference's failure test."""
        p = _lib.ParticleParams(**{k: float(v) for k, v in params.items()})
        flagged = ctypes.c_int(0)
        check(self.lib.sg_particles_track(self.ctx, float(dt), ctypes.byref(p), ctypes.byref(flagged)))
        return flagged.value

    def geometry_rebuild(self, kind, cx=0.0, cy=0.0, a=0.0, b=0.0, c=0.0):
        """Cut-cell geometry of a box (half sizes a, b), circle (radius a),
        ellipse (semi-axes a, b) or boom (cylinder radius a, bar half width b,
        bar length c) centred at (cx, cy), or of no solid, rebuilt on the
        device (P, U and V cells)."""
        s = _lib.SolidShape(_lib.SOLID_KINDS[kind], float(cx), float(cy), float(a), float(b), float(c))
        check(self.lib.sg_geometry_rebuild(self.ctx, ctypes.byref(s)))

    def object_force(self, vs=0.0):
        """ComputeForceObject on the device fields: (force, clp, clf)."""
        f, a, b = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self.lib.sg_object_force(self.ctx, float(vs), ctypes.byref(f), ctypes.byref(a), ctypes.byref(b)))
        return f.value, a.value, b.value

    def set_moving_solid(self, kind, cx=0.0, cy=0.0, a=0.0, b=0.0, c=0.0, us=0.0, vs=0.0, asx=0.0, asy=0.0):
        """Arms the solid motion inside the transport's sub-steps ("none"
        disarms)."""
        s = _lib.SolidShape(_lib.SOLID_KINDS[kind], float(cx), float(cy), float(a), float(b), float(c))
        check(self.lib.sg_set_moving_solid(self.ctx, ctypes.byref(s), float(us), float(vs), float(asx), float(asy)))

    def get_moving_solid(self):
        """(kind, cx, cy, a, b, c), us, vs of the moved solid."""
        s = _lib.SolidShape()
        us, vs = ctypes.c_double(), ctypes.c_double()
        check(self.lib.sg_get_moving_solid(self.ctx, ctypes.byref(s), ctypes.byref(us), ctypes.byref(vs)))
        return (s.kind, s.cx, s.cy, s.a, s.b, s.c), us.value, vs.value

    def geometry_snapshot(self):
        """Keeps the geometry of the step's start for inter_new_var."""
        check(self.lib.sg_geometry_snapshot(self.ctx))

    def inter_new_var(self, vb=0.0):
        """InterNewVar on the device velocities after a geometry rebuild."""
        check(self.lib.sg_inter_new_var(self.ctx, float(vb)))

    def stag_download(self, which):
        """The device's staggered geometry of U (0) or V (1) cells as a dict
        of the cases.StagGeometry array names."""
        out = {}
        names = _lib.STAG_CELL[:8] + _lib.STAG_GHOST
        for k, name in enumerate(names):
            shape = (self.nj, self.ni) if k < 8 else (self.nj + 2, self.ni + 2)
            a = np.empty(shape)
            check(self.lib.sg_stag_download(self.ctx, which, k, a.ctypes.data_as(ctypes.c_void_p)))
            out[name] = a
        for k, name in ((len(names), "moex"), (len(names) + 1, "active")):
            a = np.empty((self.nj, self.ni), dtype=np.int32)
            check(self.lib.sg_stag_download(self.ctx, which, k, a.ctypes.data_as(ctypes.c_void_p)))
            out[name] = a
        return out

    def diagnostics_record(self):
        """Enqueues one row of per-step diagnostics (no host wait)."""
        check(self.lib.sg_diagnostics_record(self.ctx))

    def diagnostics_read(self, max_rows=_lib.DIAG_RING):
        """Rows recorded and not yet read, oldest first: (rows (n, 8), lost)."""
        out = np.zeros((max(max_rows, 1), _lib.DIAG_COLS))
        n = ctypes.c_int(0)
        lost = ctypes.c_longlong(0)
        check(self.lib.sg_diagnostics_read(self.ctx, max_rows, _d(out)[1], ctypes.byref(n), ctypes.byref(lost)))
        return out[:n.value].copy(), lost.value

    def state_save(self):
        """The context's time-stepping state (sg_state_save) as bytes."""
        n = ctypes.c_size_t(0)
        check(self.lib.sg_state_bytes(self.ctx, ctypes.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        check(self.lib.sg_state_save(self.ctx, buf.ctypes.data_as(ctypes.c_void_p), n.value))
        return buf.tobytes()

    def state_load(self, blob):
        """Restores a state saved by state_save on a context of the same grid."""
        buf = np.frombuffer(blob, dtype=np.uint8)
        check(self.lib.sg_state_load(self.ctx, buf.ctypes.data_as(ctypes.c_void_p), buf.size))
        # the particle count is the header's last field (bytes 168-175)
        self._npart = int(np.frombuffer(blob[168:176], dtype=np.int64)[0])

    def interface_to_uv(self):
        """Initial_ClsVofUV on the device fields; returns the U and V
        (vof, phi, nx, ny)."""
        check(self.lib.sg_interface_to_uv_resident(self.ctx))
        return (tuple(self.download(k) for k in ("UVOF", "UPHI", "UNX", "UNY")),
                tuple(self.download(k) for k in ("VVOF", "VPHI", "VNX", "VNY")))
