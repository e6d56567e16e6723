// ctx.h -- the context behind the opaque pictor_ctx handle (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "../../include/pictor_hip.h"
#include "pictor_common.h"
#include "cyl.h"
#include "cylm.h"

struct pictor_ctx;

namespace pictor {

// message of the last failed call on this thread (pictor_last_error)
extern thread_local std::string g_err;

inline int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPTRY(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return ::pictor::fail(PICTOR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline void dfree(void *p) { (void)hipFree(p); }

template <typename T>
hipError_t dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, n * sizeof(T));
}

// boundary faces (bc.hip), called by the step: the current folds of the
// reflecting faces after a push, the PML after the B (b = true) or E update,
// the field faces and face motion at the end of a step, and the release
int bc_after_push(pictor_ctx *c);
int bc_pml(pictor_ctx *c, bool b);
int bc_final(pictor_ctx *c);
void bc_release(pictor_ctx *c);

// the faces as the restart files keep them (bc_face of vars.F90:148-158 and the
// periodic axes); the PML auxiliary fields follow in the file when pml_on
struct BcState {
    int32_t per[3], fld[6], prtl[6], pml_on;
    double pos_fld[6], pos_prtl[6], speed[6], pml_L[6];
};
void bc_get(const pictor_ctx *c, BcState &s);
int bc_set(pictor_ctx *c, const BcState &s);   // allocates the PML fields when needed

// adjoint engine owned by a context (adjoint_api.hip)
struct AdjEngineBase {
    virtual ~AdjEngineBase() {}
};

}  // namespace pictor

struct pictor_ctx {
    using GridP = pictor::GridP;
    using PhysP = pictor::PhysP;
    using TileP = pictor::TileP;
    using PartF64 = pictor::PartF64;
    using PartSplit = pictor::PartSplit;

    pictor_config cfg{};
    GridP g{};
    PhysP ph{};
    TileP tp{};
    int form = 0;
    int dev = 0;
    hipStream_t stream = nullptr;

    // fields: double (FP64 form) or float (split form); J double or int64 fixed point
    void *E[3] = {nullptr, nullptr, nullptr};
    void *B[3] = {nullptr, nullptr, nullptr};
    void *J[3] = {nullptr, nullptr, nullptr};
    // split form: the low FP32 part of E (E = hi + lo, hi in E[]), see launch_update_e_pair
    float *Elo[3] = {nullptr, nullptr, nullptr};
    // smoothen_current: passes per step and the filter's scratch current
    int curr_filter = 0;
    double *Jt[3] = {nullptr, nullptr, nullptr};

    // tile bins
    std::vector<long long> h_base;
    std::vector<int> h_cap;
    long long *d_base = nullptr;
    int *d_cap = nullptr;
    int *d_cnt[2] = {nullptr, nullptr};
    long long layout_cap = 0;   // particle slots the layout needs per buffer

    // particle buffers: [0],[1] ping-pong, spill, staging for host transfers
    PartF64 pf[2]{};
    PartSplit ps[2]{};
    long long buf_cap[2] = {0, 0};
    int cur = 0;
    bool pending_swap = false;
    PartF64 spill_f{};
    PartSplit spill_s{};
    int spill_cap = 0;
    int *d_spill_count = nullptr;
    int *d_err = nullptr;
    int *h_flags = nullptr;   // pinned: {spill count, error bits}
    // re-layout scratch, allocated with the bins: per (tile, species) group counts,
    // totals, fill cursors and offsets; the group of each spilled particle
    int *rl_gcount = nullptr, *rl_gtotal = nullptr, *rl_fill = nullptr, *rl_grp = nullptr;
    long long *rl_off = nullptr;
    int rl_grp_cap = 0;

    // diagnostics scratch
    double *d_partial = nullptr;
    int partial_len = 0;
    double *d_rho = nullptr;

    // timers: 0 push, 1 field kernels, 2 whole step.  ev[0..7]: the phase calls and the
    // cylindrical steps; ev[8..14] and ev[15..21]: the fused Cartesian step, two sets of 7
    // used alternately so that one step's events are read while the next step is being
    // enqueued; ev[22]: the spill and error flags of a phase-call push copied to the host
    hipEvent_t ev[23] = {};
    double t_ms[3] = {0, 0, 0};
    long long t_n[3] = {0, 0, 0};

    long long np = 0;
    // the particles carry the caller's tags (uploaded with pictor_upload_particles_tagged):
    // a tag array beside every particle buffer (PartF64::tag, PartSplit::id)
    bool tagged = false;
    // split form: the particles do not all carry their flavour's charge, so every particle
    // record has a charge weight (PartSplit::wq); set by the upload
    bool weighted = false;
    long long n_relayout = 0;   // re-layouts after a spill (bins overflowed) since init
    long long n_spilled = 0;    // particles those re-layouts took from the spill buffer

    // cylindrical geometry: particles in the caller's order (cyl.hip)
    bool cyl = false;
    pictor::CylP cp{};
    PartF64 pcyl{};
    int *d_cerr = nullptr;
    unsigned long long *d_nsub = nullptr;
    // azimuthal-mode cylinder (cylm.hip): set together with cyl, which routes
    // the phase calls; the modal functions then take over
    bool cylm = false;
    pictor::CylmP cq{};
    pictor::CylmState cm{};

    // boundary faces (bc.hip): face f = 0..5 is x lo, x hi, y lo, y hi, z lo, z hi;
    // PICTOR_BC_* types, global positions, speeds (updateBC_Pos) and PML thickness
    int bc_fld[6] = {0, 0, 0, 0, 0, 0}, bc_prtl[6] = {0, 0, 0, 0, 0, 0};
    double bc_pos_fld[6] = {0, 0, 0, 0, 0, 0}, bc_pos_prtl[6] = {0, 0, 0, 0, 0, 0};
    double bc_speed[6] = {0, 0, 0, 0, 0, 0}, pml_L[6] = {4, 4, 4, 4, 4, 4};
    bool bc_any = false;        // a face has been set
    bool pml_on = false;        // a PML face is set
    double *pml_psi[12] = {};   // pml_E_psi1 x,y,z, pml_E_psi2, pml_B_psi1, pml_B_psi2 (ncell each)
    double *pml_tab = nullptr;  // PML coefficient tables (bc.hip), rebuilt when a PML face moves
    bool pml_tab_stale = true;

    pictor::AdjEngineBase *adj = nullptr;   // created by pictor_adjoint_forward
    int adj_engine = 0;                      // 0 tile-sorted (FP32 split form), 1 unsorted
};
