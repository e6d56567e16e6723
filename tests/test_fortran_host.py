"""The Fortran host (pictor_amd/host): the reference's MAIN loop over the C ABI.

CPU: the host compiles with amdflang against the interface module.
GPU: a run of the Fortran binary gives the same energies as the Python host
on the same seeded input (split form: deterministic, so field energies agree
bit for bit; kinetic energies are FP64 sums over particles in storage order).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "pictor_amd", "host")
FC = "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
def test_fortran_host_compiles(tmp_path):
    srcs = ["parameters.F90", "pictor_cabi.F90", "vars.F90", "setup_weibel.F90", "MAIN.F90"]
    for s in srcs:
        shutil.copy(os.path.join(HOST, s), tmp_path)
    r = subprocess.run([FC, "-O0", "-c"] + srcs, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]


def _energies(out):
    rows = {}
    for line in out.splitlines():
        if line.startswith("energy"):
            f = line.split()
            rows[int(f[1])] = [float(x) for x in f[2:]]
    return rows


@pytest.mark.gpu
@pytest.mark.parametrize("fused,curr_filter", [(0, 0), (1, 0), (0, 2)])
def test_fortran_host_matches_python_host(fused, curr_filter):
    exe = os.path.join(HOST, "pictor_main")
    assert os.path.exists(exe), "pictor_main is not built (__graft_entry__.build)"
    print(f"[case] fortran host fused={fused} curr_filter={curr_filter}", flush=True)
    r = subprocess.run([exe, "nx=32", "ny=32", "epc=8", "tfinish=20", "form=1", f"fused={fused}",
                        f"curr_filter={curr_filter}", "energy_save_period=10", "seed=7"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    F = _energies(r.stdout)
    from pictor_amd import Simulation, setup_weibel
    P = setup_weibel(32, 32, epc=8, drift=0.5, uth=0.01, seed=7)
    with Simulation(32, 32, epc=8, form="fp32_split", curr_filter=curr_filter) as sim:
        sim.upload_particles(**P)
        for t in (10, 20):
            sim.step(10)
            d = sim.diagnostics()
            want = [d["B_energy"], d["E_energy"], d["prtl_energy_1"], d["prtl_energy_2"], d["gauss_residual"]]
            got = F[t]
            assert got[0] == want[0] and got[1] == want[1], (t, got, want)
            assert np.allclose(got[2:4], want[2:4], rtol=1e-12, atol=0)
