// restart.hip -- restart files: SaveRestartData (src/io/savedata.F90:73-141, writers
// :940-1210) and the reload of src/io/reload.F90:17-338 (restart branch
// src/MAIN.F90:35-39).
//
// The reference writes three HDF5 files per rank and restart step: fld_<proc>_<t>
// (Ex..Bz, (mx,my,mz)), prtl_<proc>_<t> (qp, flvp, xp..wp and the tag fields, the
// first used_prtl_arr_size slots) and param_<proc>_<t> (grid sizes, particle counts,
// flavour tables), plus the restart time.  This image has no HDF5, so one plain
// binary file holds the same content:
//
//   StateHeader (configuration, t, particle count, moving-window state)
//   E, B       6 arrays of ncell values in the form's own precision: FP64, or the
//              split form's FP32 B and the FP32 (hi, lo) pair of E (9 arrays)
//   particles  q, x, y, z, u, v, w (FP64) and flv (int32), n values each, the
//              reference SoA; it holds the split record exactly (cell + FP32 offset
//              and FP32 momenta convert to FP64 and back without rounding)
//   faces      (version 2) the boundary faces with their current positions and
//              speeds, then the 12 PML auxiliary fields (FP64) if a PML face is set;
//              a restart takes the face positions from the file, as the reference's
//              SetBC_* skip the setup position when restarting (bc.F90:30, 58)
//
// J is not saved: every step starts by resetting it (MAIN.F90:248).  A state loads
// into a context created with the same configuration; the tile edge may differ.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pictor_hip.h"
#include "ctx.h"

using namespace pictor;

namespace {

struct StateHeader {
    char magic[8];
    int32_t version;
    int32_t field_bytes;   // 8 (FP64 fields) or 4 (split form)
    int32_t has_elo;       // split form: the low FP32 half of E follows B
    int32_t geometry, ndim, nx, ny, nz, epc, form, tile, fdtd, curr_filter;
    double c, c_ompe, g0, mi_me, dr;
    int64_t t, np, ncell;
    // modal cylinder: moving window and sort schedule
    int32_t win_on, win_nrp, win_ppc, sort_period;
    int64_t win_steps, win_shift, steps_since_sort;
    uint64_t win_seed;
    // the reference r-theta cylinder: push scheme (0 Boris, 1 Strang)
    int32_t cyl_scheme;
    int32_t has_tags;      // version 3: the particles' tags follow their flavours
    int32_t reserved[6];
};
constexpr char kMagic[8] = {'P', 'I', 'C', 'T', 'O', 'R', 'S', 'T'};
constexpr int32_t kVersion = 3;   // 1: no faces section; 2: no tags

struct File {
    FILE *f = nullptr;
    explicit File(FILE *p) : f(p) {}
    ~File() { if (f) fclose(f); }
};

int write_all(FILE *f, const void *p, size_t bytes, const char *what)
{
    if (bytes && fwrite(p, 1, bytes, f) != bytes) return fail(PICTOR_ERR_ARG, std::string("restart file: writing ") + what + " failed");
    return PICTOR_OK;
}

int read_all(FILE *f, void *p, size_t bytes, const char *what)
{
    if (bytes && fread(p, 1, bytes, f) != bytes)
        return fail(PICTOR_ERR_ARG, std::string("restart file: truncated while reading ") + what);
    return PICTOR_OK;
}

// device arrays of the state's fields, in file order, and their element size
void state_arrays(pictor_ctx *c, std::vector<void *> &arr, size_t &esize)
{
    const bool f64 = c->form == PICTOR_FORM_FP64 || c->cylm;
    esize = f64 ? sizeof(double) : sizeof(float);
    arr.assign({c->E[0], c->E[1], c->E[2], c->B[0], c->B[1], c->B[2]});
    if (!f64 && c->Elo[0]) arr.insert(arr.end(), {c->Elo[0], c->Elo[1], c->Elo[2]});
}

}  // namespace

extern "C" {

int pictor_save_state(pictor_ctx *c, const char *path, int64_t t)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    if (!path) return fail(PICTOR_ERR_ARG, "null path");
    HIPTRY(hipSetDevice(c->dev));
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "save between move_deposit and exchange_particles");
    int64_t np = 0;
    int rc = pictor_particle_count(c, &np);
    if (rc) return rc;
    std::vector<double> q(np), x(np), y(np), z(np), u(np), v(np), w(np);
    std::vector<int32_t> flv(np), tag(c->tagged ? np : 0);
    int64_t got = 0;
    if ((rc = pictor_download_particles_tagged(c, np, q.data(), x.data(), y.data(), z.data(), u.data(), v.data(),
                                               w.data(), flv.data(), nullptr, c->tagged ? tag.data() : nullptr, &got)))
        return rc;
    if (got != np) return fail(PICTOR_ERR_ARG, "restart: particle count changed during the download");

    StateHeader h{};
    memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = kVersion;
    std::vector<void *> arr;
    size_t es = 0;
    state_arrays(c, arr, es);
    h.field_bytes = (int32_t)es;
    h.has_elo = arr.size() > 6;
    const pictor_config &cf = c->cfg;
    h.geometry = cf.geometry; h.ndim = cf.ndim; h.nx = cf.nx; h.ny = cf.ny; h.nz = cf.nz; h.epc = cf.epc;
    h.form = cf.form; h.tile = cf.tile; h.fdtd = cf.fdtd; h.curr_filter = cf.curr_filter;
    h.c = cf.c; h.c_ompe = cf.c_ompe; h.g0 = cf.g0; h.mi_me = cf.mi_me; h.dr = cf.dr;
    h.t = t; h.np = np; h.ncell = c->g.ncell;
    h.has_tags = c->tagged ? 1 : 0;
    if (c->cylm) {
        const CylmState &st = c->cm;
        h.win_on = st.win_on; h.win_nrp = st.win_nrp; h.win_ppc = st.win_ppc; h.sort_period = st.sort_period;
        h.win_steps = st.win_steps; h.win_shift = st.win_shift; h.steps_since_sort = st.steps_since_sort;
        h.win_seed = st.win_seed;
    }
    if (c->cyl && !c->cylm) h.cyl_scheme = c->cp.strang;

    // fields: one host staging array, copied array by array
    const size_t n = (size_t)c->g.ncell;
    std::vector<unsigned char> buf(n * es);
    std::string tmp = std::string(path) + ".tmp";
    {
        File f(fopen(tmp.c_str(), "wb"));
        if (!f.f) return fail(PICTOR_ERR_ARG, std::string("restart: cannot create ") + tmp);
        if ((rc = write_all(f.f, &h, sizeof(h), "the header"))) return rc;
        for (void *a : arr) {
            HIPTRY(hipMemcpyAsync(buf.data(), a, n * es, hipMemcpyDeviceToHost, c->stream));
            HIPTRY(hipStreamSynchronize(c->stream));
            if ((rc = write_all(f.f, buf.data(), n * es, "a field"))) return rc;
        }
        const std::vector<double> *cols[7] = {&q, &x, &y, &z, &u, &v, &w};
        for (const auto *col : cols)
            if ((rc = write_all(f.f, col->data(), (size_t)np * sizeof(double), "particles"))) return rc;
        if ((rc = write_all(f.f, flv.data(), (size_t)np * sizeof(int32_t), "particle flavours"))) return rc;
        if (h.has_tags && (rc = write_all(f.f, tag.data(), (size_t)np * sizeof(int32_t), "particle tags"))) return rc;
        BcState bs;
        bc_get(c, bs);
        if ((rc = write_all(f.f, &bs, sizeof(bs), "the boundary faces"))) return rc;
        if (bs.pml_on) {
            std::vector<double> pb(n);
            for (int q = 0; q < 12; q++) {
                HIPTRY(hipMemcpyAsync(pb.data(), c->pml_psi[q], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                HIPTRY(hipStreamSynchronize(c->stream));
                if ((rc = write_all(f.f, pb.data(), n * sizeof(double), "a PML field"))) return rc;
            }
        }
        if (fflush(f.f) != 0) return fail(PICTOR_ERR_ARG, "restart: flush failed");
    }
    // a complete file replaces an older one of the same name only once it is whole
    if (rename(tmp.c_str(), path) != 0) return fail(PICTOR_ERR_ARG, std::string("restart: cannot rename to ") + path);
    return PICTOR_OK;
}

int pictor_load_state(pictor_ctx *c, const char *path, int64_t *t)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    if (!path) return fail(PICTOR_ERR_ARG, "null path");
    HIPTRY(hipSetDevice(c->dev));
    File f(fopen(path, "rb"));
    if (!f.f) return fail(PICTOR_ERR_ARG, std::string("restart: cannot open ") + path);
    StateHeader h{};
    int rc = read_all(f.f, &h, sizeof(h), "the header");
    if (rc) return rc;
    if (memcmp(h.magic, kMagic, sizeof(kMagic)) != 0 || h.version < 1 || h.version > kVersion)
        return fail(PICTOR_ERR_ARG, "restart: not a state file of this version");
    const pictor_config &cf = c->cfg;
    std::vector<void *> arr;
    size_t es = 0;
    state_arrays(c, arr, es);
    // the configuration decides the charge scaling (InitScale), the grid and the arithmetic
    if (h.geometry != cf.geometry || h.ndim != cf.ndim || h.nx != cf.nx || h.ny != cf.ny ||
        (cf.ndim == 3 && h.nz != cf.nz) || h.epc != cf.epc || h.form != cf.form || h.fdtd != cf.fdtd ||
        h.curr_filter != cf.curr_filter || h.c != cf.c || h.c_ompe != cf.c_ompe || h.g0 != cf.g0 ||
        h.mi_me != cf.mi_me || h.dr != cf.dr)
        return fail(PICTOR_ERR_ARG, "restart: the state was saved with another configuration");
    if (h.ncell != c->g.ncell || (size_t)h.field_bytes != es || (h.has_elo != 0) != (arr.size() > 6) || h.np < 0)
        return fail(PICTOR_ERR_ARG, "restart: inconsistent header");
    const size_t n = (size_t)c->g.ncell;
    std::vector<std::vector<unsigned char>> fld(arr.size(), std::vector<unsigned char>(n * es));
    for (auto &b : fld)
        if ((rc = read_all(f.f, b.data(), n * es, "a field"))) return rc;
    const int64_t np = h.np;
    std::vector<double> q(np), x(np), y(np), z(np), u(np), v(np), w(np);
    std::vector<int32_t> flv(np);
    std::vector<double> *cols[7] = {&q, &x, &y, &z, &u, &v, &w};
    for (auto *col : cols)
        if ((rc = read_all(f.f, col->data(), (size_t)np * sizeof(double), "particles"))) return rc;
    if ((rc = read_all(f.f, flv.data(), (size_t)np * sizeof(int32_t), "particle flavours"))) return rc;
    const bool has_tags = h.version >= 3 && h.has_tags;
    std::vector<int32_t> tag(has_tags ? np : 0);
    if (has_tags && (rc = read_all(f.f, tag.data(), (size_t)np * sizeof(int32_t), "particle tags"))) return rc;
    BcState bs;
    bc_get(c, bs);   // version 1: the context's own faces
    std::vector<std::vector<double>> psi;
    if (h.version >= 2) {
        if ((rc = read_all(f.f, &bs, sizeof(bs), "the boundary faces"))) return rc;
        if (bs.pml_on) {
            psi.assign(12, std::vector<double>(n));
            for (auto &b : psi)
                if ((rc = read_all(f.f, b.data(), n * sizeof(double), "a PML field"))) return rc;
        }
    }
    if (c->cyl && (bs.per[0] == 0 || bs.per[1] == 0 || bs.per[2] == 0))
        return fail(PICTOR_ERR_ARG, "restart: boundary faces on a cylindrical grid");

    // faces first (the upload's tile assignment does not depend on them), then the
    // particles: a rejected upload leaves the fields as they were
    if (!c->cyl && (rc = bc_set(c, bs))) return rc;
    if (has_tags) {
        if ((rc = pictor_upload_particles_tagged(c, np, q.data(), x.data(), y.data(), z.data(), u.data(), v.data(),
                                                 w.data(), flv.data(), tag.data())))
            return rc;
    } else if ((rc = pictor_upload_particles(c, np, q.data(), x.data(), y.data(), z.data(), u.data(), v.data(),
                                             w.data(), flv.data()))) {
        return rc;
    }
    for (size_t i = 0; i < arr.size(); i++)
        HIPTRY(hipMemcpyAsync(arr[i], fld[i].data(), n * es, hipMemcpyHostToDevice, c->stream));
    for (size_t q = 0; q < psi.size(); q++)
        HIPTRY(hipMemcpyAsync(c->pml_psi[q], psi[q].data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    if ((rc = pictor_reset_current(c))) return rc;
    if (c->cylm) {
        CylmState &st = c->cm;
        st.win_on = h.win_on != 0;
        c->cq.zopen = st.win_on ? 1 : 0;
        st.win_nrp = h.win_nrp; st.win_ppc = h.win_ppc; st.win_seed = h.win_seed;
        st.win_steps = h.win_steps; st.win_shift = h.win_shift;
        st.sort_period = h.sort_period; st.steps_since_sort = h.steps_since_sort;
    }
    if (c->cyl && !c->cylm) c->cp.strang = h.cyl_scheme;
    HIPTRY(hipStreamSynchronize(c->stream));
    if (t) *t = h.t;
    return PICTOR_OK;
}

}  // extern "C"
