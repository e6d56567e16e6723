"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker (or the timed CPU
baseline), never the product path.

Arrays follow the reference layout: a field is a numpy array of shape
(mx, my, mz) in Fortran order, so F(i,j,k) of the Fortran code is
F[i-1, j-1, k-1] here; particles are 1-D arrays q,x,y,z,u,v,w (float64)
and flv (int32) of length `cap`, the first `used` entries live.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle.so")

c_dp = ctypes.POINTER(ctypes.c_double)
c_ip = ctypes.POINTER(ctypes.c_int32)


class OrcParams(ctypes.Structure):
    _fields_ = (
        [(n, ctypes.c_int32) for n in ("ndim", "nx", "ny", "nz", "mx", "my", "mz", "epc")]
        + [(n, ctypes.c_double) for n in (
            "c", "c_ompe", "g0", "mi_me", "me", "mi", "ompe", "qi", "qe", "qmi", "qme", "masse", "massi",
            "fldc", "fld_halfc", "cinv", "sqc", "inv_dx", "inv_dy", "inv_dz",
            "xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "xlen", "ylen", "zlen",
            "Bx_ext0", "By_ext0", "Bz_ext0")]
        + [("nflvr", ctypes.c_int32), ("reorder_period", ctypes.c_int32), ("flvrqm", ctypes.c_double * 8)]
        + [("fdtd", ctypes.c_int32), ("curr_filter", ctypes.c_int32), ("lehe", ctypes.c_double * 12)]
        + [("periodic", ctypes.c_int32 * 3), ("bc_fld", ctypes.c_int32 * 6), ("bc_prtl", ctypes.c_int32 * 6)]
        + [(n, ctypes.c_double * 6) for n in ("bc_pos_fld", "bc_pos_prtl", "bc_speed", "pml_L")]
        + [("cond_top_as_written", ctypes.c_int32)]
    )


# face types (pictor_oracle.h ORC_BC_*) by the reference's names
BC_TYPES = {"prdc": 0, "cond": 1, "open": 2, "pml": 3, "refl": 4, "remv": 5}


class OrcCyl(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_double) for n in ("dtheta", "inv_dtheta", "ax_perm_area", "rshift", "rmin", "r_proc")]
                + [("second_half_new_cell", ctypes.c_int32), ("push", ctypes.c_int32),
                   ("substep_limit", ctypes.c_int64), ("substeps", ctypes.c_int64), ("lehe", ctypes.c_double * 12)])


class OrcFields(ctypes.Structure):
    _fields_ = [(n, c_dp) for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz")] + [("pml", c_dp * 12)]


class OrcParticles(ctypes.Structure):
    _fields_ = ([("cap", ctypes.c_int64), ("used", ctypes.c_int64)]
                + [(n, c_dp) for n in ("q", "x", "y", "z", "u", "v", "w")]
                + [("flv", c_ip)])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(_LIBPATH)
        P = ctypes.POINTER(OrcParams)
        F = ctypes.POINTER(OrcFields)
        S = ctypes.POINTER(OrcParticles)
        W = ctypes.c_void_p
        L.orc_params_init.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_set_lehe.argtypes = [P]
        L.orc_ws_create.argtypes = [P, ctypes.c_int]
        L.orc_ws_create.restype = W
        L.orc_ws_destroy.argtypes = [W]
        for name in ("orc_update_b_half", "orc_update_e", "orc_add_current", "orc_reset_current"):
            getattr(L, name).argtypes = [P, F]
        L.orc_field_halo.argtypes = [P, c_dp, c_dp, c_dp]
        L.orc_current_halo_add.argtypes = [P, c_dp, c_dp, c_dp]
        L.orc_smoothen_current.argtypes = [P, F]
        L.orc_gather_vec_e.argtypes = [P, W, F]
        L.orc_gather_vec_b.argtypes = [P, W, F]
        L.orc_move_deposit.argtypes = [P, W, F, S]
        L.orc_ws_vec.argtypes = [W, ctypes.c_int]
        L.orc_ws_vec.restype = c_dp
        L.orc_ws_vec_len.argtypes = [W]
        L.orc_exchange_particles.argtypes = [P, S]
        L.orc_reorder.argtypes = [P, S]
        L.orc_step.argtypes = [P, W, F, S, ctypes.c_int]
        L.orc_run.argtypes = [P, W, F, S, ctypes.c_int, ctypes.c_int]
        L.orc_energy.argtypes = [P, F, S, c_dp]
        L.orc_div_e.argtypes = [P, F, c_dp]
        L.orc_rho.argtypes = [P, S, c_dp]
        L.orc_gauss.argtypes = [P, F, S, c_dp]
        L.orc_momentum.argtypes = [P, S, c_dp]
        L.orc_bc_post_movdep.argtypes = [P, F, S]
        L.orc_bc_final.argtypes = [P, F]
        L.orc_pml_b.argtypes = [P, F]
        L.orc_pml_e.argtypes = [P, F]
        L.orc_pml_coeffs.argtypes = [P, ctypes.c_double, ctypes.c_int, c_dp]
        C = ctypes.POINTER(OrcCyl)
        L.orc_cyl_init.argtypes = [C, P]
        for name in ("orc_cyl_update_e", "orc_cyl_update_b_half", "orc_cyl_add_current"):
            getattr(L, name).argtypes = [P, C, F]
        L.orc_cyl_axis_current_bc.argtypes = [P, F]
        L.orc_cyl_field_halo.argtypes = [P, c_dp, c_dp, c_dp]
        L.orc_cyl_current_halo_add.argtypes = [P, c_dp, c_dp, c_dp]
        L.orc_cyl_move_deposit.argtypes = [P, C, W, F, S]
        L.orc_cyl_energy.argtypes = [P, C, F, S, c_dp]
        L.orc_cyl_exchange_particles.argtypes = [P, S]
        L.orc_cyl_step.argtypes = [P, C, W, F, S]
        _lib = L
    return _lib


def _ptr(a, t=c_dp):
    return a.ctypes.data_as(t)


class Oracle:
    """FP64 reference CPU path on one periodic rank (restated PICTOR)."""

    FIELD_NAMES = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz")

    def __init__(self, ndim=2, nx=16, ny=16, nz=1, c=0.45, c_ompe=10.0, g0=1.0, mi_me=1.0, epc=16,
                 nthreads=1, cap=1024, reorder_period=20, fdtd="yee", curr_filter=0):
        L = lib()
        self.p = OrcParams()
        rc = L.orc_params_init(ctypes.byref(self.p), ndim, nx, ny, nz, c, c_ompe, g0, mi_me, epc)
        if rc:
            raise ValueError("invalid oracle configuration")
        if fdtd == "lehe":
            L.orc_set_lehe(ctypes.byref(self.p))
        elif fdtd != "yee":
            raise ValueError("fdtd must be 'yee' or 'lehe'")
        self.p.reorder_period = reorder_period
        self.p.curr_filter = curr_filter
        self.shape = (self.p.mx, self.p.my, self.p.mz)
        self.F = {n: np.zeros(self.shape, order="F") for n in self.FIELD_NAMES}
        self.f = OrcFields(*[_ptr(self.F[n]) for n in self.FIELD_NAMES])
        self._pml = [None] * 12
        self._alloc_particles(cap)
        self.ws = L.orc_ws_create(ctypes.byref(self.p), nthreads)
        if not self.ws:
            raise MemoryError("oracle workspace allocation failed")

    def __del__(self):
        try:
            if getattr(self, "ws", None):
                lib().orc_ws_destroy(self.ws)
                self.ws = None
        except Exception:
            pass

    # ---- particles ----
    def _alloc_particles(self, cap):
        self.P = {n: np.zeros(cap) for n in ("q", "x", "y", "z", "u", "v", "w")}
        self.P["flv"] = np.zeros(cap, dtype=np.int32)
        self.s = OrcParticles(cap, 0, *[_ptr(self.P[n]) for n in ("q", "x", "y", "z", "u", "v", "w")],
                              _ptr(self.P["flv"], c_ip))

    def load_particles(self, q, x, y, z, u, v, w, flv, cap=None):
        n = len(x)
        cap = max(cap or 0, int(n * 1.25) + 1024)
        self._alloc_particles(cap)
        for name, arr in zip(("q", "x", "y", "z", "u", "v", "w"), (q, x, y, z, u, v, w)):
            self.P[name][:n] = arr
        self.P["flv"][:n] = flv
        self.s.used = n

    def particles(self, live_only=True):
        n = self.s.used
        out = {k: v[:n].copy() for k, v in self.P.items()}
        if live_only:
            keep = out["flv"] != 0
            out = {k: v[keep] for k, v in out.items()}
        return out

    @property
    def used(self):
        return self.s.used

    # ---- fields ----
    def set_fields(self, **arrays):
        for k, v in arrays.items():
            self.F[k][...] = v

    # ---- phases of MAIN.F90 ----
    def update_b_half(self):
        lib().orc_update_b_half(ctypes.byref(self.p), ctypes.byref(self.f))

    def update_e(self):
        lib().orc_update_e(ctypes.byref(self.p), ctypes.byref(self.f))

    def add_current(self):
        lib().orc_add_current(ctypes.byref(self.p), ctypes.byref(self.f))

    def reset_current(self):
        lib().orc_reset_current(ctypes.byref(self.p), ctypes.byref(self.f))

    def field_halo(self, which):
        names = {"E": ("Ex", "Ey", "Ez"), 0: ("Ex", "Ey", "Ez"), "J": ("Jx", "Jy", "Jz")}.get(which, ("Bx", "By", "Bz"))
        lib().orc_field_halo(ctypes.byref(self.p), *[_ptr(self.F[n]) for n in names])

    def current_halo_add(self):
        lib().orc_current_halo_add(ctypes.byref(self.p), *[_ptr(self.F[n]) for n in ("Jx", "Jy", "Jz")])

    def smoothen_current(self):
        lib().orc_smoothen_current(ctypes.byref(self.p), ctypes.byref(self.f))

    def move_deposit(self):
        lib().orc_move_deposit(ctypes.byref(self.p), self.ws, ctypes.byref(self.f), ctypes.byref(self.s))

    # ---- boundary faces (bc.F90) ----
    @staticmethod
    def _face(axis, side):
        a = {"x": 0, "y": 1, "z": 2}[axis] if isinstance(axis, str) else int(axis)
        s = {"lower": 0, "upper": 1}[side] if isinstance(side, str) else int(side)
        return a, 2 * a + s

    def set_bc_fld(self, axis, position, kind, side, thickness=4.0):
        """SetBC_Fld (bc.F90:23-50): a field face at a global position."""
        a, f = self._face(axis, side)
        self.p.bc_fld[f] = BC_TYPES[kind]
        self.p.bc_pos_fld[f] = position
        self.p.pml_L[f] = thickness
        self.p.periodic[a] = 0
        if kind == "pml" and self._pml[0] is None:
            for k in range(12):
                self._pml[k] = np.zeros(self.shape, order="F")
                self.f.pml[k] = _ptr(self._pml[k])

    def set_bc_prtl(self, axis, position, kind, side):
        """SetBC_Prtl (bc.F90:52-63): a particle face at a global position."""
        a, f = self._face(axis, side)
        self.p.bc_prtl[f] = BC_TYPES[kind]
        self.p.bc_pos_prtl[f] = position
        self.p.periodic[a] = 0

    def set_boundary_speed(self, axis, side, speed):
        """SetBoundarySpeed (bc.F90:210-217)."""
        self.p.bc_speed[self._face(axis, side)[1]] = speed

    @property
    def pml_psi(self):
        return self._pml

    def bc_post_movdep(self):
        lib().orc_bc_post_movdep(ctypes.byref(self.p), ctypes.byref(self.f), ctypes.byref(self.s))

    def bc_final(self):
        lib().orc_bc_final(ctypes.byref(self.p), ctypes.byref(self.f))

    def pml_b(self):
        lib().orc_pml_b(ctypes.byref(self.p), ctypes.byref(self.f))

    def pml_e(self):
        lib().orc_pml_e(ctypes.byref(self.p), ctypes.byref(self.f))

    def pml_coeffs(self, x, axis):
        out = np.zeros(3)
        lib().orc_pml_coeffs(ctypes.byref(self.p), float(x), int(axis), _ptr(out))
        return out

    def gather_vec(self):
        L = lib()
        L.orc_gather_vec_e(ctypes.byref(self.p), self.ws, ctypes.byref(self.f))
        L.orc_gather_vec_b(ctypes.byref(self.p), self.ws, ctypes.byref(self.f))
        nv = L.orc_ws_vec_len(self.ws)
        ncell = self.p.mx * self.p.my * self.p.mz
        out = []
        for c in range(6):
            ptr = L.orc_ws_vec(self.ws, c)
            out.append(np.ctypeslib.as_array(ptr, shape=(ncell * nv,)).reshape(ncell, nv).copy())
        return out

    def exchange(self):
        rc = lib().orc_exchange_particles(ctypes.byref(self.p), ctypes.byref(self.s))
        if rc:
            raise RuntimeError(f"oracle exchange failed ({rc})")

    def reorder(self):
        rc = lib().orc_reorder(ctypes.byref(self.p), ctypes.byref(self.s))
        if rc:
            raise RuntimeError(f"oracle reorder failed ({rc})")

    def step(self, t):
        rc = lib().orc_step(ctypes.byref(self.p), self.ws, ctypes.byref(self.f), ctypes.byref(self.s), t)
        if rc:
            raise RuntimeError(f"oracle step failed ({rc})")

    def run(self, t0, nsteps):
        rc = lib().orc_run(ctypes.byref(self.p), self.ws, ctypes.byref(self.f), ctypes.byref(self.s), t0, nsteps)
        if rc:
            raise RuntimeError(f"oracle run failed ({rc})")

    # ---- diagnostics ----
    def energy(self):
        out = np.zeros(4)
        lib().orc_energy(ctypes.byref(self.p), ctypes.byref(self.f), ctypes.byref(self.s), _ptr(out))
        return out

    def gauss(self):
        out = np.zeros(2)
        lib().orc_gauss(ctypes.byref(self.p), ctypes.byref(self.f), ctypes.byref(self.s), _ptr(out))
        return out

    def momentum(self):
        out = np.zeros(3)
        lib().orc_momentum(ctypes.byref(self.p), ctypes.byref(self.s), _ptr(out))
        return out

    def rho(self):
        out = np.zeros(self.shape, order="F")
        lib().orc_rho(ctypes.byref(self.p), ctypes.byref(self.s), _ptr(out))
        return out

    def div_e(self):
        out = np.zeros(self.shape, order="F")
        lib().orc_div_e(ctypes.byref(self.p), ctypes.byref(self.f), _ptr(out))
        return out


class CylOracle(Oracle):
    """FP64 reference CPU path of the cylindrical build (-Dcyl, inc_axis) on the
    rank holding the axis: x = r (grid index i at r = i - 3.5), y = theta
    (ny cells over 2 pi), z = z.  The outer x ghosts are a conducting wall."""

    def __init__(self, ndim=2, nr=16, ntheta=16, nz=1, c=0.45, c_ompe=10.0, g0=1.0, mi_me=1.0, epc=16,
                 cap=1024, fdtd="yee", second_half_new_cell=True, push="boris"):
        super().__init__(ndim, nr, ntheta, nz, c=c, c_ompe=c_ompe, g0=g0, mi_me=mi_me, epc=epc, nthreads=1,
                         cap=cap, reorder_period=0)
        self.cy = OrcCyl()
        if lib().orc_cyl_init(ctypes.byref(self.cy), ctypes.byref(self.p)):
            raise ValueError("the cylindrical grid needs an even number of theta cells")
        self.p.fdtd = {"yee": 0, "lehe": 1}[fdtd]
        self.cy.second_half_new_cell = 1 if second_half_new_cell else 0
        self.cy.push = {"boris": 0, "strang": 1, "strang_as_written": 2}[push]

    def energy(self):
        """CalcEnergy with -Dcyl: B energy, E energy (radius-weighted), flavour 1, flavour 2"""
        out = np.zeros(4)
        lib().orc_cyl_energy(ctypes.byref(self.p), ctypes.byref(self.cy), ctypes.byref(self.f), ctypes.byref(self.s),
                             _ptr(out))
        return out

    def update_b_half(self):
        lib().orc_cyl_update_b_half(ctypes.byref(self.p), ctypes.byref(self.cy), ctypes.byref(self.f))

    def update_e(self):
        lib().orc_cyl_update_e(ctypes.byref(self.p), ctypes.byref(self.cy), ctypes.byref(self.f))

    def add_current(self):
        lib().orc_cyl_add_current(ctypes.byref(self.p), ctypes.byref(self.cy), ctypes.byref(self.f))

    def field_halo(self, which):
        names = {"E": ("Ex", "Ey", "Ez"), 0: ("Ex", "Ey", "Ez")}.get(which, ("Bx", "By", "Bz"))
        lib().orc_cyl_field_halo(ctypes.byref(self.p), *[_ptr(self.F[n]) for n in names])

    def current_halo_add(self):
        lib().orc_cyl_current_halo_add(ctypes.byref(self.p), *[_ptr(self.F[n]) for n in ("Jx", "Jy", "Jz")])

    def move_deposit(self):
        rc = lib().orc_cyl_move_deposit(ctypes.byref(self.p), ctypes.byref(self.cy), self.ws, ctypes.byref(self.f),
                                        ctypes.byref(self.s))
        if rc:
            raise RuntimeError("cylindrical push failed: a particle reached the outer wall or the axis sub-steps diverged")
        return self.cy.substeps

    def exchange(self):
        lib().orc_cyl_exchange_particles(ctypes.byref(self.p), ctypes.byref(self.s))

    def step(self, t=0):
        rc = lib().orc_cyl_step(ctypes.byref(self.p), ctypes.byref(self.cy), self.ws, ctypes.byref(self.f),
                                ctypes.byref(self.s))
        if rc:
            raise RuntimeError(f"cylindrical oracle step failed ({rc})")

    def run(self, t0, nsteps):
        for t in range(t0, t0 + nsteps):
            self.step(t)

    def radius(self, i):
        """r of grid index i (1-based, the reference's i)"""
        return (i - 3.0) + self.cy.rmin


class OrcCylm(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int32) for n in ("nr", "nz", "mx", "mz")]
                + [(n, ctypes.c_double) for n in ("dr", "inv_dr", "inv_dz", "c", "fldc", "fld_halfc", "cinv", "sqc",
                                                  "ompe", "qi")]
                + [("flvrqm", ctypes.c_double * 2)]
                + [(n, ctypes.c_double) for n in ("zmin", "zmax", "zlen")]
                + [("fdtd", ctypes.c_int32), ("zopen", ctypes.c_int32)]
                + [(n, ctypes.c_double) for n in ("l_rr", "l_rz", "l_tt", "l_tz", "l_zz", "l_zr", "l_zd", "l_zm")])


class CylModalOracle:
    """FP64 statement of the azimuthal-mode (m = 0, 1) cylindrical step (no
    reference counterpart; built from the reference's r-z staggering, metric
    factors and axis rules).  Fields: numpy arrays (mx, 3, mz) in Fortran order,
    slot 0 = m 0, 1 = m 1 cos, 2 = m 1 sin; particles x = radial grid coordinate,
    y = theta (radians), z = axial grid coordinate."""

    FIELD_NAMES = Oracle.FIELD_NAMES

    def __init__(self, nr=16, nz=16, dr=1.0, c=0.45, c_ompe=10.0, g0=1.0, mi_me=1e30, epc=16, cap=1024, fdtd=0):
        L = lib()
        P = ctypes.POINTER(OrcCylm)
        F = ctypes.POINTER(OrcFields)
        S = ctypes.POINTER(OrcParticles)
        if not getattr(L, "_cylm_declared", False):
            L.orc_cylm_init.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
            for n in ("orc_cylm_b_half", "orc_cylm_current_halo", "orc_cylm_reset_current"):
                getattr(L, n).argtypes = [P, F]
            L.orc_cylm_e_phase.argtypes = [P, F, ctypes.c_int, ctypes.c_int]
            L.orc_cylm_set_fdtd.argtypes = [P, ctypes.c_int]
            L.orc_cylm_set_zopen.argtypes = [P, ctypes.c_int]
            L.orc_cylm_set_zopen.restype = None
            L.orc_cylm_move_deposit.argtypes = [P, F, S]
            L.orc_cylm_step.argtypes = [P, F, S]
            L.orc_cylm_energy.argtypes = [P, F, c_dp]
            L.orc_cylm_gauss.argtypes = [P, F, S, c_dp]
            L.orc_cylm_load.argtypes = [P, S, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_uint64, ctypes.c_int64]
            L.orc_cylm_load.restype = ctypes.c_int64
            L.orc_cylm_shift.argtypes = [P, F, S, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                         ctypes.c_int64]
            L._cylm_declared = True
        self.m = OrcCylm()
        if L.orc_cylm_init(ctypes.byref(self.m), nr, nz, dr, c, c_ompe, g0, mi_me, epc) or \
                L.orc_cylm_set_fdtd(ctypes.byref(self.m), fdtd):
            raise ValueError("invalid modal configuration")
        self.epc = epc
        self.shape = (self.m.mx, 3, self.m.mz)
        self.F = {n: np.zeros(self.shape, order="F") for n in self.FIELD_NAMES}
        self.f = OrcFields(*[_ptr(self.F[n]) for n in self.FIELD_NAMES])
        self._pml = [None] * 12
        self._alloc(cap)

    def _alloc(self, cap):
        self.P = {n: np.zeros(cap) for n in ("q", "x", "y", "z", "u", "v", "w")}
        self.P["flv"] = np.zeros(cap, dtype=np.int32)
        self.s = OrcParticles(cap, 0, *[_ptr(self.P[n]) for n in ("q", "x", "y", "z", "u", "v", "w")],
                              _ptr(self.P["flv"], c_ip))

    def load_particles(self, q, x, y, z, u, v, w, flv, cap=None):
        n = len(x)
        self._alloc(max(cap or 0, n + 1024))
        for name, arr in zip(("q", "x", "y", "z", "u", "v", "w"), (q, x, y, z, u, v, w)):
            self.P[name][:n] = arr
        self.P["flv"][:n] = flv
        self.s.used = n

    def load_plasma(self, nr_p, k0, nzc, ppc, seed=1, shift=0, cap=None):
        """uniform cold electrons (orc_cylm_load); returns the particle arrays"""
        need = nr_p * nzc * ppc
        self._alloc(max(cap or 0, need + 1024))
        lib().orc_cylm_load(ctypes.byref(self.m), ctypes.byref(self.s), nr_p, k0, nzc, ppc, self.epc, seed, shift)
        return self.particles(live_only=False)

    def particles(self, live_only=True):
        n = self.s.used
        out = {k: v[:n].copy() for k, v in self.P.items()}
        if live_only:
            keep = out["flv"] != 0
            out = {k: v[keep] for k, v in out.items()}
        return out

    def set_fields(self, **arrays):
        for k, v in arrays.items():
            self.F[k][...] = v

    def b_half(self):
        lib().orc_cylm_b_half(ctypes.byref(self.m), ctypes.byref(self.f))

    def e_phase(self, curl=True, add=False):
        lib().orc_cylm_e_phase(ctypes.byref(self.m), ctypes.byref(self.f), int(curl), int(add))

    def current_halo(self):
        lib().orc_cylm_current_halo(ctypes.byref(self.m), ctypes.byref(self.f))

    def reset_current(self):
        lib().orc_cylm_reset_current(ctypes.byref(self.m), ctypes.byref(self.f))

    def move_deposit(self):
        if lib().orc_cylm_move_deposit(ctypes.byref(self.m), ctypes.byref(self.f), ctypes.byref(self.s)):
            raise RuntimeError("modal push: a particle reached the outer wall")

    def step(self):
        if lib().orc_cylm_step(ctypes.byref(self.m), ctypes.byref(self.f), ctypes.byref(self.s)):
            raise RuntimeError("modal oracle step failed")

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()

    def set_zopen(self, on=True):
        """open z ends (the moving window): ghosts hold no field, no z fold of J,
        particles leaving through either end are dropped"""
        lib().orc_cylm_set_zopen(ctypes.byref(self.m), 1 if on else 0)

    def shift(self, nr_p, ppc, seed, shift):
        """the moving window's shift by one axial cell (shift = the shift count after it)"""
        lib().orc_cylm_shift(ctypes.byref(self.m), ctypes.byref(self.f), ctypes.byref(self.s), nr_p, ppc, self.epc,
                             seed, shift)

    def energy(self):
        out = np.zeros(2)
        lib().orc_cylm_energy(ctypes.byref(self.m), ctypes.byref(self.f), _ptr(out))
        return out

    def gauss(self):
        g = np.zeros(self.shape, order="F")
        lib().orc_cylm_gauss(ctypes.byref(self.m), ctypes.byref(self.f), ctypes.byref(self.s), _ptr(g))
        return g

    def radius(self, i):
        """r of grid index i (1-based)"""
        return (i - 3.5) * self.m.dr
