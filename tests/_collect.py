"""numpy restatement of CollectFld (src/io/savedata_routines.F90:207-261) for one
rank over the whole box (GetSizeofCollectFld, :114-173): the node values the
field files hold.  Test infrastructure; pinned by tests/test_output_collect.py."""
import numpy as np


def collect(F, ndim, res):
    """F: dict of (mx, my, mz) Fortran-order FP64 arrays (Ex..Bz, optionally Jx..Jz);
    returns the FP32 arrays of the field file."""
    mx, my, mz = F["Ex"].shape
    nx, ny, nz = mx - 5, my - 5, mz - 5
    i = np.arange(3, res * (nx // res) + 3 + 1, res) - 1          # 0-based node indices
    j = np.arange(3, res * (ny // res) + 3 + 1, res) - 1
    k = np.arange(3, res * (nz // res) + 3 + 1, res) - 1 if ndim == 3 else np.array([0])
    I, J, K = np.meshgrid(i, j, k, indexing="ij")

    def at(A, di=0, dj=0, dk=0):
        return A[I + di, J + dj, K + dk]

    out = {}
    out["Ex"] = (at(F["Ex"], -1) + at(F["Ex"])) * 0.5
    out["Ey"] = (at(F["Ey"], 0, -1) + at(F["Ey"])) * 0.5
    if ndim == 3:
        out["Ez"] = (at(F["Ez"], 0, 0, -1) + at(F["Ez"])) * 0.5
        out["Bx"] = (at(F["Bx"]) + at(F["Bx"], 0, -1) + at(F["Bx"], 0, 0, -1) + at(F["Bx"], 0, -1, -1)) * 0.25
        out["By"] = (at(F["By"]) + at(F["By"], -1) + at(F["By"], 0, 0, -1) + at(F["By"], -1, 0, -1)) * 0.25
    else:
        out["Ez"] = at(F["Ez"])
        out["Bx"] = (at(F["Bx"]) + at(F["Bx"], 0, -1)) * 0.5
        out["By"] = (at(F["By"]) + at(F["By"], -1)) * 0.5
    out["Bz"] = (at(F["Bz"]) + at(F["Bz"], -1) + at(F["Bz"], 0, -1) + at(F["Bz"], -1, -1)) * 0.25
    for n in ("Jx", "Jy", "Jz"):
        if n in F:
            out[n] = at(F[n])
    return {n: np.asfortranarray(v.astype(np.float32)) for n, v in out.items()}
