"""Python host of the gfx950 particle step, mirroring PICTOR's MAIN loop.

The reference drives one time step by calling zero-argument subroutines in
a fixed order (src/MAIN.F90:42-284).  `Simulation` exposes the same phases
under the same names, each one a call through the C ABI of
libpictor_hip.so; `Simulation.step()` is the fused form of one whole step.
Arrays are numpy arrays in the reference layout: fields (mx, my, mz) in
Fortran order, particles 1-D SoA.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import (BC_TYPES, FDTD_LEHE, FDTD_YEE, FORM_FP32_SPLIT, FORM_FP64, GEOM_CARTESIAN, GEOM_CYL, GEOM_CYL_MODAL,
                   PictorConfig, check, lib)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


DIAG_KEYS = ("B_energy", "E_energy", "prtl_energy_1", "prtl_energy_2", "px", "py", "pz", "gauss_residual", "rho_max")


class Simulation:
    """One periodic domain on one GPU (parameters.F90 as runtime settings).

    form: "fp64" (reference SoA in FP64) or "fp32_split" (int32 cell +
    FP32 offset, FP32 push/fields, fixed-point current).
    """

    def __init__(self, nx, ny, nz=1, ndim=2, epc=16, c=0.45, c_ompe=10.0, g0=1.0, mi_me=1.0,
                 form="fp64", tile=0, capacity_factor=0.0, fdtd="yee", geometry="cartesian", curr_filter=0, dr=0.0):
        if fdtd not in ("yee", "lehe"):
            raise ValueError("fdtd must be 'yee' or 'lehe'")
        if geometry not in ("cartesian", "cyl", "cyl_modal"):
            raise ValueError("geometry must be 'cartesian', 'cyl' or 'cyl_modal'")
        self.geometry = geometry
        self.cfg = PictorConfig(ndim, nx, ny, nz, epc, FORM_FP64 if form == "fp64" else FORM_FP32_SPLIT,
                                tile, FDTD_LEHE if fdtd == "lehe" else FDTD_YEE,
                                {"cyl": GEOM_CYL, "cyl_modal": GEOM_CYL_MODAL}.get(geometry, GEOM_CARTESIAN),
                                curr_filter, c, c_ompe, g0, mi_me, capacity_factor, dr)
        if form not in ("fp64", "fp32_split"):
            raise ValueError("form must be 'fp64' or 'fp32_split'")
        self.form = form
        self._ctx = ctypes.c_void_p()
        check(lib().pictor_init(ctypes.byref(self.cfg), ctypes.byref(self._ctx)))
        sc = np.zeros(16)
        check(lib().pictor_get_scalars(self._ctx, _p(sc)))
        self.qi = sc[0]
        self.shape = (int(sc[9]), int(sc[10]), int(sc[11]))
        self.tile = int(sc[12])
        self.ntiles = int(sc[13])
        self.t = 0

    def save_state(self, path, t=None):
        """SaveRestartData (savedata.F90:73-141): the context's state to one file; t defaults to self.t."""
        t = self.t if t is None else int(t)
        check(lib().pictor_save_state(self._ctx, os.fsencode(path), t))

    def load_state(self, path):
        """The reload of reload.F90:17-338 into this context (same configuration); returns and sets t."""
        t = ctypes.c_int64(0)
        check(lib().pictor_load_state(self._ctx, os.fsencode(path), ctypes.byref(t)))
        self.t = int(t.value)
        return self.t

    def save_total_energy(self, path):
        """save_total_energy (savedata.F90:856-890): Benergy, Eenergy, PrtlEnergy as a .npz file."""
        check(lib().pictor_save_energy(self._ctx, os.fsencode(path)))

    def save_field_collective(self, path, res=1, save_current=False):
        """save_field_collective (savedata.F90:356-497): CollectFld node samples as a .npz file."""
        check(lib().pictor_save_fields(self._ctx, os.fsencode(path), int(res), 1 if save_current else 0))

    def relayout_stats(self):
        """(re-layouts after a tile-bin overflow, particles taken from the spill buffer) since init."""
        sc = np.zeros(16)
        check(lib().pictor_get_scalars(self._ctx, _p(sc)))
        return int(sc[14]), int(sc[15])

    def close(self):
        if self._ctx:
            lib().pictor_finalize(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state transfer (reference layout) ----
    def upload_fields(self, Ex, Ey, Ez, Bx, By, Bz):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (Ex, Ey, Ez, Bx, By, Bz)]
        for a in arrs:
            if a.shape != self.shape:
                raise ValueError(f"field shape {a.shape} != {self.shape}")
        check(lib().pictor_upload_fields(self._ctx, *[_p(a) for a in arrs]))

    def download_fields(self):
        out = {n: np.zeros(self.shape, order="F") for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz")}
        check(lib().pictor_download_fields(self._ctx, *[_p(out[n]) for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz",
                                                                              "Jx", "Jy", "Jz")]))
        return out

    def upload_particles(self, q, x, y, z, u, v, w, flv, tag=None):
        """Particles in the reference SoA; tag (optional, int32 per particle): the
        reference's tagp, carried with each particle (download_particles(with_tag=True))."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (q, x, y, z, u, v, w)]
        f = np.ascontiguousarray(flv, dtype=np.int32)
        n = len(f)
        if tag is None:
            check(lib().pictor_upload_particles(self._ctx, n, *[_p(a) for a in arrs], _p(f, _ip)))
        else:
            t = np.ascontiguousarray(tag, dtype=np.int32)
            if len(t) != n:
                raise ValueError("one tag per particle")
            check(lib().pictor_upload_particles_tagged(self._ctx, n, *[_p(a) for a in arrs], _p(f, _ip), _p(t, _ip)))

    def particle_count(self):
        n = ctypes.c_int64()
        check(lib().pictor_particle_count(self._ctx, ctypes.byref(n)))
        return n.value

    def filtered_efield(self, passes):
        """SetFilteredEfldGPU (fields_gpu.F90:70-90): E after `passes` rounds of the 1-2-1
        filter and the halo copy (the field the reference's test particles feel)."""
        out = {n: np.zeros(self.shape, order="F") for n in ("Ex", "Ey", "Ez")}
        check(lib().pictor_filtered_efield(self._ctx, passes, *[_p(out[n]) for n in ("Ex", "Ey", "Ez")]))
        return out

    def download_particles(self, with_tile=False, with_tag=False):
        n = self.particle_count()
        out = {k: np.zeros(n) for k in ("q", "x", "y", "z", "u", "v", "w")}
        out["flv"] = np.zeros(n, dtype=np.int32)
        tile = np.zeros(n, dtype=np.int32) if with_tile else None
        nout = ctypes.c_int64()
        cols = [_p(out[k]) for k in ("q", "x", "y", "z", "u", "v", "w")]
        if with_tag:
            tag = np.zeros(n, dtype=np.int32)
            check(lib().pictor_download_particles_tagged(self._ctx, n, *cols, _p(out["flv"], _ip), _p(tile, _ip),
                                                         _p(tag, _ip), ctypes.byref(nout)))
            out["tag"] = tag
        else:
            check(lib().pictor_download_particles(self._ctx, n, *cols, _p(out["flv"], _ip), _p(tile, _ip),
                                                  ctypes.byref(nout)))
        if with_tile:
            out["tile"] = tile
        return out

    # ---- the phases of MAIN.F90, same names as the reference ----
    def UpdateBfieldHalf(self):          # em_update.F90:41-77 + SendRecvFlds(B)
        check(lib().pictor_update_b_half(self._ctx))

    def MoveDepositPrtl(self):           # movdep.F90:12-258
        check(lib().pictor_move_deposit(self._ctx))

    def ExchangePrtl(self):              # comm_prtl.F90:36-65 + AppendParticles
        check(lib().pictor_exchange_particles(self._ctx))

    def UpdateEfield(self):              # em_update.F90:9-39 + SendRecvFlds(E)
        check(lib().pictor_update_e(self._ctx))

    def UpdateCurrentAllFaces(self):     # comm_fld.F90:15-33
        check(lib().pictor_current_halo_add(self._ctx))

    # ---- the azimuthal-mode cylinder (geometry "cyl_modal") ----
    def cyl_push_scheme(self, scheme):
        """push of the cylindrical grid: "boris" (cyl_movdep.F90, the default) or "strang"
        (cyl_movdep_strang.F90, the reference's CPU -Dcyl build)"""
        check(lib().pictor_cyl_push_scheme(self._ctx, {"boris": 0, "strang": 1}[scheme]))

    def cyl_sort_period(self, period):
        """modal cylinder: sort the particles by cell (and drop holes) every `period` steps"""
        check(lib().pictor_cyl_sort_period(self._ctx, period))

    def cyl_plasma(self, nr_plasma, k0, nzc, ppc, seed=1):
        """cold uniform electrons in radial cells 4..4+nr_plasma-1, axial cells k0..k0+nzc-1"""
        check(lib().pictor_cyl_plasma(self._ctx, nr_plasma, k0, nzc, ppc, seed))

    def cyl_laser(self, a0, omega0, tau, w0, zc):
        """Gaussian x-polarised pulse along +z (omega0 in omega_p, tau in 1/omega_p, w0 in c/omega_p, zc in cells)"""
        check(lib().pictor_cyl_laser(self._ctx, a0, omega0, tau, w0, zc))

    def moving_window(self, on=True, nr_plasma=0, ppc=0, seed=1):
        check(lib().pictor_moving_window(self._ctx, 1 if on else 0, nr_plasma, ppc, seed))

    def smoothen_current(self):          # current.F90:19-30 (curr_filter passes)
        check(lib().pictor_smoothen_current(self._ctx))

    def AddCurrent(self):                # em_update.F90:79-102 + SendRecvFlds(E)
        check(lib().pictor_add_current(self._ctx))

    def SendRecvFlds(self, which):       # comm_fld.F90:75-91; which "E" or "B"
        check(lib().pictor_field_halo(self._ctx, 0 if which == "E" else 1))

    def ResetCurrent(self):              # current.F90:13-17
        check(lib().pictor_reset_current(self._ctx))

    # ---- boundary faces (src/bc), the reference's setup calls ----
    @staticmethod
    def _face(axis, side):
        return {"x": 0, "y": 1, "z": 2}[axis], {"lower": 0, "upper": 1}[side]

    def SetBC_Fld(self, Axis, Position, Type, Side, AttnThickness=4.0):   # bc.F90:23-50
        a, s = self._face(Axis, Side)
        check(lib().pictor_set_bc_fld(self._ctx, a, s, BC_TYPES[Type], float(Position), float(AttnThickness)))

    def SetBC_Prtl(self, Axis, Position, Type, Side):                     # bc.F90:52-63
        a, s = self._face(Axis, Side)
        check(lib().pictor_set_bc_prtl(self._ctx, a, s, BC_TYPES[Type], float(Position)))

    def SetBoundarySpeed(self, Axis, Side, Speed):                         # bc.F90:210-217
        a, s = self._face(Axis, Side)
        check(lib().pictor_set_boundary_speed(self._ctx, a, s, float(Speed)))

    def UpdateMagFld_PML(self):          # bc_pml.F90:185-290 + SendRecvFlds(B)
        check(lib().pictor_pml(self._ctx, 1))

    def UpdateElcFld_PML(self):          # bc_pml.F90:85-183 + SendRecvFlds(E)
        check(lib().pictor_pml(self._ctx, 0))

    def EnforceBC_Final(self):           # bc.F90:289-340
        check(lib().pictor_enforce_bc_final(self._ctx))

    def ReorderPrtl(self):               # mem_prtl.F90:162-301
        check(lib().pictor_sort_tiles(self._ctx))

    def step_phases(self):
        """One time step as the sequence of separate phase calls (MAIN.F90 order)."""
        self.ResetCurrent()
        self.UpdateBfieldHalf()
        self.MoveDepositPrtl()
        self.ExchangePrtl()
        self.UpdateBfieldHalf()
        self.UpdateMagFld_PML()
        self.UpdateEfield()
        self.UpdateElcFld_PML()
        self.UpdateCurrentAllFaces()
        self.smoothen_current()
        self.AddCurrent()
        self.EnforceBC_Final()
        self.t += 1

    def step(self, n=1):
        """n fused time steps."""
        check(lib().pictor_step(self._ctx, n))
        self.t += n

    def synchronize(self):
        check(lib().pictor_synchronize(self._ctx))

    def diagnostics(self):
        out = np.zeros(9)
        check(lib().pictor_diagnostics(self._ctx, _p(out)))
        return dict(zip(DIAG_KEYS, out.tolist()))

    def timer_reset(self):
        check(lib().pictor_timer_reset(self._ctx))

    def timer(self, which):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(lib().pictor_timer_get(self._ctx, {"push": 0, "fields": 1, "step": 2}[which], ctypes.byref(ms),
                                     ctypes.byref(n)))
        return ms.value, n.value

    def memory_bytes(self):
        b = ctypes.c_int64()
        check(lib().pictor_memory_bytes(self._ctx, ctypes.byref(b)))
        return b.value

    # ---- adjoint of the 2-D step (pictor_adjoint_*) ----
    def adjoint_forward(self, q, x, y, z, u, v, w, flv, nsteps, segment):
        """Objective 1/2 sum B^2 after nsteps steps from these particles and the
        current fields, keeping a checkpoint every `segment` steps."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (q, x, y, z, u, v, w)]
        f = np.ascontiguousarray(flv, dtype=np.int32)
        L = ctypes.c_double()
        check(lib().pictor_adjoint_forward(self._ctx, len(arrs[1]), *[_p(a) for a in arrs], _p(f, _ip), nsteps,
                                           segment, ctypes.byref(L)))
        self._adj_n = len(arrs[1])
        return L.value

    def adjoint_backward(self, scale=1.0):
        """scale * dL/d(u, v, w) of the initial momenta (numpy, caller's order)."""
        n = self._adj_n
        du, dv, dw = np.zeros(n), np.zeros(n), np.zeros(n)
        check(lib().pictor_adjoint_backward(self._ctx, scale, _p(du), _p(dv), _p(dw)))
        return du, dv, dw

    def adjoint_forward_device(self, n, q, x, y, z, u, v, w, flv, nsteps, segment):
        """adjoint_forward from device addresses (contiguous FP64 arrays q .. w and
        an int32 flv of n particles on this context's GPU); the caller's writes
        to them must be complete (e.g. torch.cuda.synchronize())."""
        L = ctypes.c_double()
        ptrs = [ctypes.c_void_p(int(a)) for a in (q, x, y, z, u, v, w, flv)]
        check(lib().pictor_adjoint_forward_device(self._ctx, int(n), *ptrs, nsteps, segment, ctypes.byref(L)))
        self._adj_n = int(n)
        return L.value

    def adjoint_backward_device(self, du, dv, dw, scale=1.0):
        """scale * dL/d(u, v, w) written to the device addresses du, dv, dw
        (contiguous FP64 arrays of n elements); complete when the call returns."""
        ptrs = [ctypes.c_void_p(int(a)) for a in (du, dv, dw)]
        check(lib().pictor_adjoint_backward_device(self._ctx, scale, *ptrs))

    def adjoint_stats(self):
        out = np.zeros(10)
        check(lib().pictor_adjoint_stats(self._ctx, _p(out)))
        keys = ("forward_ms", "recompute_ms", "reverse_ms", "device_bytes", "nsteps", "segment", "checkpoints",
                "objective", "engine", "relayouts")
        return dict(zip(keys, out.tolist()))

    def adjoint_release(self):
        check(lib().pictor_adjoint_release(self._ctx))

    def adjoint_engine(self, engine):
        """'sorted' (tile-sorted FP32 kernels, the default) or 'unsorted'."""
        check(lib().pictor_adjoint_engine(self._ctx, {"sorted": 0, "unsorted": 1}[engine]))


def setup_weibel(nx, ny, nz=1, ndim=2, epc=16, drift=0.5, uth=0.01, seed=12345):
    """Seeded counter-streaming pair plasma in the reference SoA (host arrays)."""
    ncell = nx * ny * (nz if ndim == 3 else 1)
    n = 2 * epc * ncell
    out = {k: np.zeros(n) for k in ("q", "x", "y", "z", "u", "v", "w")}
    out["flv"] = np.zeros(n, dtype=np.int32)
    nout = ctypes.c_int64()
    check(lib().pictor_setup_weibel(ndim, nx, ny, nz, epc, drift, uth, seed, n,
                                    *[_p(out[k]) for k in ("q", "x", "y", "z", "u", "v", "w")],
                                    _p(out["flv"], _ip), ctypes.byref(nout)))
    return out


__all__ = ["Simulation", "setup_weibel", "DIAG_KEYS", "FORM_FP64", "FORM_FP32_SPLIT", "_lib"]
