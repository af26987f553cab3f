"""ctypes front end of the CPU FP64 oracle (test infrastructure only).

Only tests/, the smoke check and the CPU-baseline leg of bench.py import this
module.  Arrays follow the reference's Fortran layout: a cell array is a
numpy array of shape (nj, ni) in C order (so that x is the fastest index, as
in ``PCell%vof(1:Isize,1:Jsize)``), a ghosted array has shape (nj+2, ni+2).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsg_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_is_wet.argtypes = [ctypes.c_double] * 3
        L.orc_is_wet.restype = ctypes.c_int
        L.orc_poisson_coef.argtypes = [ctypes.c_int, ctypes.c_int] + [_D] * 9 + [ctypes.c_double] * 3 + [_D]
        L.orc_poisson_matrix.argtypes = [ctypes.c_int, ctypes.c_int] + [_D] * 6 + [ctypes.c_int] * 4 + [_D]
        L.orc_poisson_rhs.argtypes = [ctypes.c_int, ctypes.c_int] + [_D] * 11 + [ctypes.c_double, _D]
        L.orc_matr_apply.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D]
        L.orc_max_row_residual.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]
        L.orc_max_row_residual.restype = ctypes.c_double
        L.orc_pressure_scatter_bc.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D] + [ctypes.c_int] * 4 + [_D]
        L.orc_velocity_correct.argtypes = ([ctypes.c_int, ctypes.c_int] + [_D] * 13 + [_I, _I] + [_D] * 7 +
                                           [ctypes.c_double] * 4 + [_D] * 4)
        L.orc_pcg_mg_solve.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
        L.orc_pcg_mg_solve.restype = ctypes.c_int
        L.orc_pcg_jacobi_solve.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, ctypes.c_double,
                                           ctypes.c_int, _D]
        L.orc_pcg_jacobi_solve.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def poisson_coef(g, f, DpU, DpV, fl):
    """ComputePossionMatrixCoefficient -> PoCoef of shape (4, nj, ni) (S,W,E,N)."""
    ni, nj = g.ni, g.nj
    out = np.zeros((4, nj, ni))
    lib().orc_poisson_coef(ni, nj, _d(g.dx), _d(g.dy), _d(g.udx), _d(g.vdy),
                           _d(f.vof), _d(f.phi), _d(f.vofS), _d(DpU), _d(DpV),
                           fl.rho_l, fl.rho_g, fl.rho_ref, _d(out))
    return out


def poisson_matrix(g, geo, PoCoef, bc=(1, 1, 1, 0)):
    """SetPoissonMatrix -> matr of shape (5, nj, ni) (S,W,diag,E,N)."""
    ni, nj = g.ni, g.nj
    out = np.zeros((5, nj, ni))
    lib().orc_poisson_matrix(ni, nj, _d(PoCoef), _d(geo.aE), _d(geo.aW), _d(geo.aN), _d(geo.aS),
                             _d(geo.vofS), bc[0], bc[1], bc[2], bc[3], _d(out))
    return out


def poisson_rhs(g, geo, u, v, vb=0.0):
    ni, nj = g.ni, g.nj
    out = np.zeros((nj, ni))
    lib().orc_poisson_rhs(ni, nj, _d(g.dx), _d(g.dy), _d(u), _d(v), _d(geo.aE), _d(geo.aW),
                          _d(geo.aN), _d(geo.aS), _d(geo.nyS), _d(geo.WlLh), _d(geo.vofS), vb, _d(out))
    return out


def matr_apply(g, matr, x):
    y = np.zeros((g.nj, g.ni))
    lib().orc_matr_apply(g.ni, g.nj, _d(matr), _d(np.ascontiguousarray(x)), _d(y))
    return y


def max_row_residual(g, geo, matr, x, b):
    return lib().orc_max_row_residual(g.ni, g.nj, _d(matr), _d(np.ascontiguousarray(x)),
                                      _d(np.ascontiguousarray(b)), _d(geo.vofS))


def pressure_scatter_bc(g, geo, x, bc=(1, 1, 1, 0)):
    Pp = np.zeros((g.nj + 2, g.ni + 2))
    lib().orc_pressure_scatter_bc(g.ni, g.nj, _d(np.ascontiguousarray(x)), _d(geo.vofS),
                                  bc[0], bc[1], bc[2], bc[3], _d(Pp))
    return Pp


def velocity_correct(g, geo, f, fl, DpU, DpV, upred, vpred, Pp, u, v, p, vb=0.0):
    """UpdatePUV's correction; returns (u, v, p, mres) (copies of the inputs updated)."""
    u = u.copy(); v = v.copy(); p = p.copy()
    mres = np.zeros((g.nj, g.ni))
    lib().orc_velocity_correct(g.ni, g.nj, _d(g.dx), _d(g.dy), _d(g.udx), _d(g.vdy),
                               _d(f.vof), _d(f.phi), _d(geo.vofS), _d(geo.nyS), _d(geo.WlLh),
                               _d(geo.aE), _d(geo.aW), _d(geo.aN), _d(geo.aS),
                               _i(geo.uflag), _i(geo.vflag), _d(geo.vofSU), _d(geo.vofSV),
                               _d(DpU), _d(DpV), _d(upred), _d(vpred), _d(Pp), vb,
                               fl.rho_l, fl.rho_g, fl.rho_ref, _d(u), _d(v), _d(p), _d(mres))
    return u, v, p, mres


def pcg_mg_solve(g, matr, b, tol=1e-10, max_iter=500, pre=1, post=1, x0=None):
    x = np.zeros((g.nj, g.ni)) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
    rr = ctypes.c_double(0.0)
    it = lib().orc_pcg_mg_solve(g.ni, g.nj, _d(matr), _d(np.ascontiguousarray(b)), _d(x), tol,
                                max_iter, pre, post, ctypes.byref(rr))
    return x, it, rr.value


def pcg_jacobi_solve(g, matr, b, tol=1e-10, max_iter=100000):
    x = np.zeros((g.nj, g.ni))
    rr = ctypes.c_double(0.0)
    it = lib().orc_pcg_jacobi_solve(g.ni, g.nj, _d(matr), _d(np.ascontiguousarray(b)), _d(x), tol,
                                    max_iter, ctypes.byref(rr))
    return x, it, rr.value


def matr_to_scipy(g, geo, matr):
    """Assemble the oracle's row-wise matrix over fluid rows as scipy CSR
    (index = i + j*ni, x fastest); solid rows become identity rows."""
    import scipy.sparse as sp
    ni, nj = g.ni, g.nj
    N = ni * nj
    idx = np.arange(N).reshape(nj, ni)
    fluid = (geo.vofS < 1.0 - 1e-3)
    rows, cols, vals = [], [], []
    offs = {0: (-1, 0), 1: (0, -1), 3: (0, 1), 4: (1, 0)}  # S, W, E, N as (dj, di)
    for k, (dj, di) in offs.items():
        m = matr[k]
        jj, ii = np.nonzero(m)
        rows.append(idx[jj, ii]); cols.append(idx[jj + dj, ii + di]); vals.append(m[jj, ii])
    jj, ii = np.nonzero(fluid)
    rows.append(idx[jj, ii]); cols.append(idx[jj, ii]); vals.append(matr[2][jj, ii])
    jj, ii = np.nonzero(~fluid)
    rows.append(idx[jj, ii]); cols.append(idx[jj, ii]); vals.append(np.ones(len(jj)))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    return A


# ---------------------------------------------------------------------------
# Coupled level-set / VOF transport (Clsvof_Mod.f90)
class _ClsGeom(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int),
                ("dx", _D), ("dy", _D), ("xc", _D), ("yc", _D),
                ("vofS", _D), ("nxS", _D), ("nyS", _D), ("phiS", _D),
                ("aE", _D), ("aW", _D), ("aN", _D), ("aS", _D),
                ("uflag", _I), ("vflag", _I)]


class _Wave(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_double), ("amp", ctypes.c_double),
                ("kw", ctypes.c_double), ("omega", ctypes.c_double)]


_cls_declared = False


def _cls_lib():
    global _cls_declared
    L = lib()
    if not _cls_declared:
        dbl = ctypes.c_double
        G = ctypes.POINTER(_ClsGeom)
        L.orc_frac.argtypes = [dbl] * 5
        L.orc_frac.restype = dbl
        L.orc_plic_distance.argtypes = [dbl] * 5
        L.orc_plic_distance.restype = dbl
        L.orc_cellgeocal.argtypes = [dbl] * 8 + [_D, _D]
        L.orc_distance_fluid.argtypes = [dbl] * 8 + [_D]
        L.orc_distance_fluid.restype = ctypes.c_int
        L.orc_normal_irre.argtypes = [G, _D, _D, ctypes.c_int, ctypes.c_int, _D, _D]
        L.orc_reconstruct.argtypes = [G, _D, _D, _D, _D, _D]
        L.orc_masked_velocities.argtypes = [G, _D, _D, _D, _D]
        L.orc_x_sweep.argtypes = [G, _D, _D, _D, _D, _D, _D, dbl, dbl, _D, _D]
        L.orc_y_sweep.argtypes = [G, _D, _D, _D, _D, _D, _D, dbl, _D, _D]
        L.orc_boundary_copy.argtypes = [ctypes.c_int, ctypes.c_int, _D]
        L.orc_redistance.argtypes = [G, _D, _D, _D, _D, _D]
        L.orc_inflow_level.argtypes = [G, ctypes.POINTER(_Wave), dbl, ctypes.c_int, dbl]
        L.orc_inflow_level.restype = dbl
        L.orc_clsvof_advect.argtypes = [G, _D, _D, _D, _D, _D, _D, dbl, dbl, ctypes.c_int,
                                        ctypes.POINTER(_Wave)]
        L.orc_interface_to_uv.argtypes = [G, ctypes.c_int] + [_D] * 10
        _cls_declared = True
    return L


class ClsGeom:
    """Grid + static solid geometry for the transport oracle (keeps the
    numpy arrays alive for the C struct)."""

    def __init__(self, g, geo):
        c = np.ascontiguousarray
        self.ni, self.nj = g.ni, g.nj
        self._keep = dict(dx=c(g.dx, dtype=np.float64), dy=c(g.dy, dtype=np.float64),
                          xc=c(g.xc, dtype=np.float64), yc=c(g.yc, dtype=np.float64),
                          vofS=c(geo.vofS), nxS=c(geo.nxS), nyS=c(geo.nyS), phiS=c(geo.phiS),
                          aE=c(geo.aE), aW=c(geo.aW), aN=c(geo.aN), aS=c(geo.aS),
                          uflag=c(geo.uflag, dtype=np.int32), vflag=c(geo.vflag, dtype=np.int32))
        k = self._keep
        self.s = _ClsGeom(g.ni, g.nj, _d(k["dx"]), _d(k["dy"]), _d(k["xc"]), _d(k["yc"]),
                          _d(k["vofS"]), _d(k["nxS"]), _d(k["nyS"]), _d(k["phiS"]),
                          _d(k["aE"]), _d(k["aW"]), _d(k["aN"]), _d(k["aS"]),
                          _i(k["uflag"]), _i(k["vflag"]))

    @property
    def ref(self):
        return ctypes.byref(self.s)


def frac1(nx, ny, d, dx, dy):
    return _cls_lib().orc_frac(nx, ny, d, dx, dy)


def plic_distance(vf, nx, ny, dx, dy):
    return _cls_lib().orc_plic_distance(vf, nx, ny, dx, dy)


def cellgeocal(nxs, nys, phis, nxl, nyl, phil, dx, dy):
    a, b = ctypes.c_double(), ctypes.c_double()
    _cls_lib().orc_cellgeocal(nxs, nys, phis, nxl, nyl, phil, dx, dy, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def distance_fluid(nxs, nys, phis, nxf, nyf, volf, dx, dy, init=0.0):
    out = ctypes.c_double(init)
    ok = _cls_lib().orc_distance_fluid(nxs, nys, phis, nxf, nyf, volf, dx, dy, ctypes.byref(out))
    return out.value, bool(ok)


def normal_irre(cg, phi, i, j):
    a, b = ctypes.c_double(), ctypes.c_double()
    _cls_lib().orc_normal_irre(cg.ref, _d(phi), _d(cg._keep["vofS"]), i, j, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def reconstruct(cg, vof, phi):
    nx = np.zeros((cg.nj, cg.ni)); ny = np.zeros_like(nx); dis = np.zeros_like(nx)
    _cls_lib().orc_reconstruct(cg.ref, _d(np.ascontiguousarray(vof)), _d(np.ascontiguousarray(phi)),
                               _d(nx), _d(ny), _d(dis))
    return nx, ny, dis


def masked_velocities(cg, u, v):
    ue = np.zeros((cg.nj, cg.ni + 1)); ve = np.zeros((cg.nj + 1, cg.ni))
    _cls_lib().orc_masked_velocities(cg.ref, _d(u), _d(v), _d(ue), _d(ve))
    return ue, ve


def x_sweep(cg, vof, phi, nx, ny, dis, ue, dtv, level, temvf, temls):
    temvf = temvf.copy(); temls = temls.copy()
    _cls_lib().orc_x_sweep(cg.ref, _d(vof), _d(phi), _d(nx), _d(ny), _d(dis), _d(ue), dtv, level,
                           _d(temvf), _d(temls))
    return temvf, temls


def y_sweep(cg, vof, phi, nx, ny, dis, ve, dtv, temvf, temls):
    temvf = temvf.copy(); temls = temls.copy()
    _cls_lib().orc_y_sweep(cg.ref, _d(vof), _d(phi), _d(nx), _d(ny), _d(dis), _d(ve), dtv,
                           _d(temvf), _d(temls))
    return temvf, temls


def boundary_copy(a):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    _cls_lib().orc_boundary_copy(a.shape[1], a.shape[0], _d(a))
    return a


def redistance(cg, vof, nx, ny, dis, phi):
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    _cls_lib().orc_redistance(cg.ref, _d(np.ascontiguousarray(vof)), _d(nx), _d(ny), _d(dis), _d(phi))
    return phi


def inflow_level(cg, wave, time, k, dtv):
    """Depthw + etau of X_Sweep for sub-step k (wave = depth, amp, kw, omega)."""
    w = _Wave(*wave)
    return _cls_lib().orc_inflow_level(cg.ref, ctypes.byref(w), time, k, dtv)


def clsvof_advect(cg, vof, phi, u, v, dt, time=0.0, nsub=6, wave=(1.0e30, 0.0, 0.0, 0.0)):
    """Coupled_LS_VOF: returns (vof, phi, nx, ny) after nsub sub-steps."""
    vof = np.ascontiguousarray(vof, dtype=np.float64).copy()
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    nx = np.zeros_like(vof); ny = np.zeros_like(vof)
    w = _Wave(*wave)
    _cls_lib().orc_clsvof_advect(cg.ref, _d(vof), _d(phi), _d(nx), _d(ny), _d(np.ascontiguousarray(u)),
                                 _d(np.ascontiguousarray(v)), dt, time, nsub, ctypes.byref(w))
    return vof, phi, nx, ny


def interface_to_uv(cg, which, vof, phi, nx, ny, vofST, phiST):
    """Initial_ClsVofUV (Clsvof_Mod.f90:402-710) for U (which=0) or V
    (which=1) cells; returns (vof, phi, nx, ny) of the staggered cells."""
    c = np.ascontiguousarray
    ins = [c(a, dtype=np.float64) for a in (vof, phi, nx, ny, vofST, phiST)]
    outs = [np.zeros((cg.nj, cg.ni)) for _ in range(4)]
    _cls_lib().orc_interface_to_uv(cg.ref, which, *[_d(a) for a in ins], *[_d(a) for a in outs])
    return tuple(outs)


# ---------------------------------------------------------------------------
# Momentum predictor (PredictorUV_mod.f90)
_STAG_D1 = ["vofS", "nxS", "nyS", "phiS", "WlLh", "delh", "ccx", "ccy"]
_STAG_I = ["moex", "active"]
_STAG_D2 = ["vof_old", "vof_new", "aE", "aW", "aN", "aS", "EtaE", "EtaN", "AlE", "AlN", "DAlE", "DAlN",
            "SxE", "SyN", "fcex", "fcey", "fcnx", "fcny"]


class _Stag(ctypes.Structure):
    _fields_ = ([(k, _D) for k in _STAG_D1] + [("moex", _I), ("flag", _I)] +
                [(k, _D) for k in _STAG_D2])


class _PredIn(ctypes.Structure):
    _fields_ = ([("ni", ctypes.c_int), ("nj", ctypes.c_int)] +
                [(k, _D) for k in ("dx", "dy", "udx", "vdy")] +
                [("U", _Stag), ("V", _Stag)] +
                [(k, _D) for k in ("pvof", "pphi", "pnx", "pny", "pvofS", "pnxS", "pnyS", "pphiS", "u", "v")] +
                [(k, ctypes.c_double) for k in ("rho_l", "rho_g", "rho_ref", "nu_l", "nu_g", "nu_ref", "Re", "Fr",
                                                "dt", "vb")] +
                [("first_step", ctypes.c_int)])


_pred_declared = False


def _pred_lib():
    global _pred_declared
    L = lib()
    if not _pred_declared:
        L.orc_predictor_assemble.argtypes = [ctypes.POINTER(_PredIn)] + [_D] * 7
        L.orc_predictor_finish.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D]
        _pred_declared = True
    return L


def _stag_struct(st, vof_old, vof_new, keep):
    c = np.ascontiguousarray
    args = {}
    for k in _STAG_D1:
        a = c(getattr(st, k), dtype=np.float64); keep.append(a); args[k] = _d(a)
    mo = c(st.moex, dtype=np.int32); fl = c(st.active, dtype=np.int32)
    keep += [mo, fl]
    vo = c(vof_old, dtype=np.float64); vn = c(vof_new, dtype=np.float64)
    keep += [vo, vn]
    rest = {}
    for k in _STAG_D2[2:]:
        a = c(getattr(st, k), dtype=np.float64); keep.append(a); rest[k] = _d(a)
    return _Stag(*[args[k] for k in _STAG_D1], _i(mo), _i(fl), _d(vo), _d(vn), *[rest[k] for k in _STAG_D2[2:]])


def predictor_assemble(g, Ust, Vst, uvof_old, uvof_new, vvof_old, vvof_new, pcell, u, v, phys, dt, vb=0.0,
                       first_step=False, fluxn1=None, DpU=None, DpV=None):
    """Predictor_UV up to the linear systems.  pcell: dict with the P-cell
    arrays at time n (vof, phi, nx, ny, vofS, nxS, nyS, phiS); phys: dict
    rho_l, rho_g, rho_ref, nu_l, nu_g, nu_ref, Re, Fr.  Returns dict with
    fluxn1 (2, nj+2, ni+2), matr_u / matr_v (5, nj, ni), rhs_u / rhs_v,
    DpU / DpV (ghosted; active cells updated)."""
    c = np.ascontiguousarray
    ni, nj = g.ni, g.nj
    keep = []
    gx = [c(a, dtype=np.float64) for a in (g.dx, g.dy, g.udx, g.vdy)]
    keep += gx
    Us = _stag_struct(Ust, uvof_old, uvof_new, keep)
    Vs = _stag_struct(Vst, vvof_old, vvof_new, keep)
    pk = [c(pcell[k], dtype=np.float64) for k in ("vof", "phi", "nx", "ny", "vofS", "nxS", "nyS", "phiS")]
    uv = [c(u, dtype=np.float64), c(v, dtype=np.float64)]
    keep += pk + uv
    s = _PredIn(ni, nj, *[_d(a) for a in gx], Us, Vs, *[_d(a) for a in pk], *[_d(a) for a in uv],
                *[float(phys[k]) for k in ("rho_l", "rho_g", "rho_ref", "nu_l", "nu_g", "nu_ref", "Re", "Fr")],
                float(dt), float(vb), int(bool(first_step)))
    f1 = np.zeros((2, nj + 2, ni + 2)) if fluxn1 is None else np.array(fluxn1, dtype=np.float64, order="C")
    DU = np.zeros((nj + 2, ni + 2)) if DpU is None else np.array(DpU, dtype=np.float64, order="C")
    DV = np.zeros((nj + 2, ni + 2)) if DpV is None else np.array(DpV, dtype=np.float64, order="C")
    mu = np.zeros((5, nj, ni)); ru = np.zeros((nj, ni))
    mv = np.zeros((5, nj, ni)); rv = np.zeros((nj, ni))
    _pred_lib().orc_predictor_assemble(ctypes.byref(s), _d(f1), _d(mu), _d(ru), _d(mv), _d(rv), _d(DU), _d(DV))
    return dict(fluxn1=f1, matr_u=mu, rhs_u=ru, matr_v=mv, rhs_v=rv, DpU=DU, DpV=DV)


def _pred_struct(g, Ust, Vst, uvof_old, uvof_new, vvof_old, vvof_new, pcell, u, v, phys, dt, vb, first_step, keep):
    c = np.ascontiguousarray
    ni, nj = g.ni, g.nj
    gx = [c(a, dtype=np.float64) for a in (g.dx, g.dy, g.udx, g.vdy)]
    keep += gx
    Us = _stag_struct(Ust, uvof_old, uvof_new, keep)
    Vs = _stag_struct(Vst, vvof_old, vvof_new, keep)
    pk = [c(pcell[k], dtype=np.float64) for k in ("vof", "phi", "nx", "ny", "vofS", "nxS", "nyS", "phiS")]
    uv = [c(u, dtype=np.float64), c(v, dtype=np.float64)]
    keep += pk + uv
    return _PredIn(ni, nj, *[_d(a) for a in gx], Us, Vs, *[_d(a) for a in pk], *[_d(a) for a in uv],
                   *[float(phys[k]) for k in ("rho_l", "rho_g", "rho_ref", "nu_l", "nu_g", "nu_ref", "Re", "Fr")],
                   float(dt), float(vb), int(bool(first_step)))


def muscl_limiter(r, tolim):
    L = _pred_lib()
    L.orc_muscl_limiter.argtypes = [ctypes.c_double, ctypes.c_double]
    L.orc_muscl_limiter.restype = ctypes.c_double
    return L.orc_muscl_limiter(r, tolim)


def conv_fluxes(g, Ust, Vst, uvof, vvof, pcell, u, v, phys, dt):
    """The high-order convective face fluxes: (CFluxEW, CFluxNS), each of
    shape (2, nj+2, ni+2) (components U, V)."""
    keep = []
    s = _pred_struct(g, Ust, Vst, uvof, uvof, vvof, vvof, pcell, u, v, phys, dt, 0.0, False, keep)
    L = _pred_lib()
    L.orc_conv_fluxes.argtypes = [ctypes.POINTER(_PredIn), _D, _D]
    F1 = np.zeros((2, g.nj + 2, g.ni + 2))
    F2 = np.zeros((2, g.nj + 2, g.ni + 2))
    L.orc_conv_fluxes(ctypes.byref(s), _d(F1), _d(F2))
    return F1, F2


def mixing(g, Ust, Vst, comp, flux, pcell, phys):
    """Mixing_Procedure of the U (comp 0) or V cells on the cell array flux."""
    keep = []
    z = np.zeros(g.shape)
    zg = np.zeros((g.nj + 2, g.ni + 2))
    s = _pred_struct(g, Ust, Vst, z, z, z, z, pcell, zg, zg, phys, 1.0, 0.0, False, keep)
    L = _pred_lib()
    L.orc_mixing.argtypes = [ctypes.POINTER(_PredIn), ctypes.c_int, _D]
    f = np.array(flux, dtype=np.float64, order="C")
    L.orc_mixing(ctypes.byref(s), comp, _d(f))
    return f


def predictor_finish(g, DpU, DpV, vpred):
    DU = np.array(DpU, dtype=np.float64, order="C")
    DV = np.array(DpV, dtype=np.float64, order="C")
    vp = np.array(vpred, dtype=np.float64, order="C")
    _pred_lib().orc_predictor_finish(g.ni, g.nj, _d(DU), _d(DV), _d(vp))
    return DU, DV, vp


def momentum_solve(matr, rhs, active):
    """Direct sparse solve of one momentum system in the matr layout
    (S,W,diag,E,N); inactive cells are 0 (DeltaGetValues,
    PredictorUV_mod.f90:847-885)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    nj, ni = rhs.shape
    idx = -np.ones((nj, ni), dtype=np.int64)
    jj, ii = np.nonzero(active)
    idx[jj, ii] = np.arange(len(jj))
    rows, cols, vals = [], [], []
    offs = {0: (-1, 0), 1: (0, -1), 3: (0, 1), 4: (1, 0)}
    for k, (dj, di) in offs.items():
        m = matr[k]
        sel = (m != 0) & (active != 0)
        j2, i2 = np.nonzero(sel)
        rows.append(idx[j2, i2]); cols.append(idx[j2 + dj, i2 + di]); vals.append(m[j2, i2])
    rows.append(idx[jj, ii]); cols.append(idx[jj, ii]); vals.append(matr[2][jj, ii])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(len(jj), len(jj)))
    x = spl.spsolve(A.tocsc(), rhs[jj, ii])
    out = np.zeros((nj, ni))
    out[jj, ii] = x
    return out


# ---------------------------------------------------------------------------
# Time-step control (Solver_Mod.f90:321-344) and the closed-box ghost layer
_step_declared = False


def _step_lib():
    global _step_declared
    L = lib()
    if not _step_declared:
        dbl = ctypes.c_double
        L.orc_time_step.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _D] + [dbl] * 6
        L.orc_time_step.restype = dbl
        L.orc_closed_box_bc.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D]
        _step_declared = True
    return L


def time_step(g, u, v, cfl, Fr, vint=0.0, uref=1.0, vs=0.0, asy=0.0):
    c = np.ascontiguousarray
    return _step_lib().orc_time_step(g.ni, g.nj, _d(c(g.udx, dtype=np.float64)), _d(c(g.vdy, dtype=np.float64)),
                                     _d(c(u)), _d(c(v)), cfl, Fr, vint, uref, vs, asy)


def closed_box_bc(u, v, p):
    """In place on ghosted (nj+2, ni+2) arrays; returns them."""
    nj, ni = u.shape[0] - 2, u.shape[1] - 2
    _step_lib().orc_closed_box_bc(ni, nj, _d(u), _d(v), _d(p))
    return u, v, p


# ---------------------------------------------------------------------------
# Lagrangian particles (Particles.f90:71-249)
class ParticleParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("rho_l", "rho_g", "nu_l", "nu_g", "rho_p", "gx", "gy", "lref", "uref",
                                               "us", "vs")]


_part_declared = False


def _part_lib():
    global _part_declared
    L = lib()
    if not _part_declared:
        dbl, ci = ctypes.c_double, ctypes.c_int
        L.orc_particle_drag.argtypes = [dbl]
        L.orc_particle_drag.restype = dbl
        L.orc_particle_position.argtypes = [ci, ci, _D, _D, _D, _D, dbl, dbl, dbl, ctypes.POINTER(ci),
                                            ctypes.POINTER(ci)]
        L.orc_particle_track.argtypes = [ci, ci, _D, _D, _D, _D, _D, _D, _D, _D, ci, _D, _D, _D, _D, _D,
                                         ctypes.POINTER(ParticleParams), dbl]
        L.orc_particle_track.restype = ci
        _part_declared = True
    return L


def particle_drag(re):
    return _part_lib().orc_particle_drag(float(re))


def particle_position(g, x, y, lref=1.0):
    """ParticlePosition: 1-based (ii, jj), -1 where the point is outside."""
    c = np.ascontiguousarray
    ii, jj = ctypes.c_int(), ctypes.c_int()
    _part_lib().orc_particle_position(g.ni, g.nj, _d(c(g.xc, dtype=np.float64)), _d(c(g.yc, dtype=np.float64)),
                                      _d(c(g.dx, dtype=np.float64)), _d(c(g.dy, dtype=np.float64)), float(lref),
                                      float(x), float(y), ctypes.byref(ii), ctypes.byref(jj))
    return ii.value, jj.value


def particle_track(g, vof, vofS, u, v, px, py, pu, pv, pd, params, dt):
    """TrackingParticles over one step; returns (px, py, pu, pv, flagged)."""
    c = np.ascontiguousarray
    keep = [c(a, dtype=np.float64) for a in (g.xc, g.yc, g.dx, g.dy, vof, vofS, u, v)]
    out = [np.array(a, dtype=np.float64, order="C") for a in (px, py, pu, pv)]
    d = c(pd, dtype=np.float64)
    P = ParticleParams(**{k: float(val) for k, val in params.items()})
    bad = _part_lib().orc_particle_track(g.ni, g.nj, *[_d(a) for a in keep], len(d), *[_d(a) for a in out], _d(d),
                                         ctypes.byref(P), float(dt))
    return (*out, bad)


# ---------------------------------------------------------------------------
# Force on the solid (Clsvof_Mod.f90:2400-2549)
_force_declared = False


def object_force(g, geo, V, p, v, vs, Re):
    """ComputeForceObject on host arrays: geo a cases.Geometry, V the V-cell
    cases.StagGeometry, p and v ghosted; returns (force, clp, clf)."""
    global _force_declared
    L = lib()
    if not _force_declared:
        dbl, ci = ctypes.c_double, ctypes.c_int
        L.orc_object_force.argtypes = [ci, ci] + [_D] * 8 + [ctypes.POINTER(ci), _D, _D, dbl, dbl, _D, _D]
        L.orc_object_force.restype = dbl
        _force_declared = True
    c = np.ascontiguousarray
    arrs = [c(a, dtype=np.float64) for a in (g.dx, geo.vofS, geo.nyS, geo.WlLh, geo.aE, geo.aW, p, V.vofS)]
    moex = c(V.moex, dtype=np.int32)
    rest = [c(a, dtype=np.float64) for a in (V.delh, v)]
    clp, clf = ctypes.c_double(), ctypes.c_double()
    f = L.orc_object_force(g.ni, g.nj, *[_d(a) for a in arrs], moex.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           *[_d(a) for a in rest], float(vs), float(Re), ctypes.byref(clp), ctypes.byref(clf))
    return f, clp.value, clf.value


def inter_new_var(g, geo_old, geo_new, u, v, vb):
    """InterNewVar (Solver_Mod.f90:360-495) on host arrays (cases.Geometry at
    the step's start and now); returns the new (u, v)."""
    L = lib()
    c = np.ascontiguousarray
    L.orc_inter_new_var.argtypes = [ctypes.c_int, ctypes.c_int] + [_D] * 8 + [ctypes.c_double, _D, _D]
    arrs = [c(a, dtype=np.float64) for a in (geo_old.vofSU, geo_new.vofSU, geo_old.vofSV, geo_new.vofSV, geo_old.aE,
                                             geo_new.aE, geo_old.aN, geo_new.aN)]
    uo, vo = np.array(u, dtype=np.float64, order="C"), np.array(v, dtype=np.float64, order="C")
    L.orc_inter_new_var(g.ni, g.nj, *[_d(a) for a in arrs], float(vb), _d(uo), _d(vo))
    return uo, vo


# ---------------------------------------------------------------------------
# Cut-cell geometry (Cutcell_Mod.f90: Grid_Preprocess + NewCellFace)
class _CGrid(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("x", _D), ("y", _D), ("dx", _D), ("dy", _D)]


_CCELL_D1 = ["vofS", "phiS", "nxS", "nyS", "WlLh", "delh", "ccx", "ccy"]
_CCELL_I = ["moex", "posnu"]
_CCELL_D2 = ["aE", "aW", "aN", "aS", "fcex", "fcey", "fcnx", "fcny",
             "EtaE", "EtaN", "AlE", "AlN", "DAlE", "DAlN", "SxE", "SyN"]


class _CCell(ctypes.Structure):
    _fields_ = ([(k, _D) for k in _CCELL_D1] + [(k, _I) for k in _CCELL_I] + [(k, _D) for k in _CCELL_D2])


_geo_declared = False


def _geo_lib():
    global _geo_declared
    L = lib()
    if not _geo_declared:
        dbl = ctypes.c_double
        G = ctypes.POINTER(_CGrid)
        C = ctypes.POINTER(_CCell)
        L.orc_edge_geo_cal.argtypes = [dbl] * 7 + [_D, _D, _D]
        L.orc_solid_fraction.argtypes = [ctypes.c_int, ctypes.c_int] + [_D] * 6
        L.orc_cell_geo_cal.argtypes = [G, C, ctypes.c_int]
        L.orc_cut_cell_geometry.argtypes = [G, G, G, C, C, C, ctypes.c_int]
        _geo_declared = True
    return L


def edge_geo_cal(p1, p2, nx, ny, d):
    """Edge_Geo_Cal: (solid fraction, face centre x, face centre y)."""
    a, fx, fy = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _geo_lib().orc_edge_geo_cal(p1[0], p1[1], p2[0], p2[1], nx, ny, d, ctypes.byref(a), ctypes.byref(fx),
                                ctypes.byref(fy))
    return a.value, fx.value, fy.value


def grid_spacings(g):
    """(x, y, dx, dy) per i / per j of the P, U and V grids (Mesh_Mod.f90:123-175)."""
    P = (g.xc, g.yc, g.dx, g.dy)
    U = (g.xc + 0.5 * g.dx, g.yc, g.udx, g.dy)
    V = (g.xc, g.yc + 0.5 * g.dy, g.dx, g.vdy)
    return {"P": P, "U": U, "V": V}


def cut_cell_geometry(g, planes, prev_AlN=None, icorprob=True):
    """Grid_Preprocess + NewCellFace for solid lines `planes` = {"P"|"U"|"V":
    (phiS, nxS, nyS)} (cell arrays, phiS > 0 in the fluid).  The solid
    fractions are formed from the lines (frac + clipping).  prev_AlN = {"U":
    AlN, "V": AlN} of the previous call (ghosted); None runs the set-up twice
    so that AlE(0,j) = AlN(1,j) reads the same geometry's north factor.
    Returns {"P": dict, "U": dict, "V": dict} of numpy arrays."""
    L = _geo_lib()
    ni, nj = g.ni, g.nj
    sp = grid_spacings(g)
    keep = []

    def arr(v):
        a = np.ascontiguousarray(v, dtype=np.float64)
        keep.append(a)
        return a

    grids = {}
    for k in "PUV":
        x, y, dx, dy = (arr(a) for a in sp[k])
        grids[k] = _CGrid(ni, nj, _d(x), _d(y), _d(dx), _d(dy))

    def run(alN):
        out = {}
        cells = {}
        for k in "PUV":
            phiS, nxS, nyS = (np.array(a, dtype=np.float64, order="C") for a in planes[k])
            d = {"phiS": phiS, "nxS": nxS, "nyS": nyS, "vofS": np.zeros((nj, ni)), "WlLh": np.zeros((nj, ni)),
                 "posnu": np.zeros((nj, ni), dtype=np.int32)}
            x, y, dx, dy = sp[k]
            L.orc_solid_fraction(ni, nj, _d(arr(dx)), _d(arr(dy)), _d(phiS), _d(nxS), _d(nyS), _d(d["vofS"]))
            for n in ("aE", "aW", "aN", "aS", "fcex", "fcey", "fcnx", "fcny"):
                d[n] = np.zeros((nj + 2, ni + 2))
            if k != "P":
                for n in ("delh", "ccx", "ccy"):
                    d[n] = np.zeros((nj, ni))
                d["moex"] = np.zeros((nj, ni), dtype=np.int32)
                for n in ("EtaE", "EtaN", "AlE", "DAlE", "DAlN", "SxE", "SyN"):
                    d[n] = np.zeros((nj + 2, ni + 2))
                d["AlN"] = np.ascontiguousarray(alN[k], dtype=np.float64).copy()
            c = _CCell()
            for n in _CCELL_D1 + _CCELL_D2:
                setattr(c, n, _d(d[n]) if n in d else _D())
            for n in _CCELL_I:
                setattr(c, n, _i(d[n]) if n in d else _I())
            cells[k] = c
            out[k] = d
        L.orc_cut_cell_geometry(ctypes.byref(grids["P"]), ctypes.byref(grids["U"]), ctypes.byref(grids["V"]),
                                ctypes.byref(cells["P"]), ctypes.byref(cells["U"]), ctypes.byref(cells["V"]),
                                1 if icorprob else 0)
        # a row of the momentum system: numbered by NumberExternalCell
        # (the last column of U cells and last row of V cells never are)
        for k, (iu, iv) in (("U", (1, 0)), ("V", (0, 1))):
            act = (out[k]["posnu"] != -1).astype(np.int32)
            if iu:
                act[:, ni - 1] = 0
            if iv:
                act[nj - 1, :] = 0
            out[k]["active"] = act
        return out

    if prev_AlN is None:
        first = run({"U": np.ones((nj + 2, ni + 2)), "V": np.ones((nj + 2, ni + 2))})
        prev_AlN = {"U": first["U"]["AlN"], "V": first["V"]["AlN"]}
    return run(prev_AlN)
