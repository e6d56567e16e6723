// io.hip -- energy and field output files.
//
// save_total_energy (src/io/savedata.F90:856-890) writes energy_<t> with the
// datasets Benergy, Eenergy and PrtlEnergy(Nflvr); save_field_collective
// (savedata.F90:356-497) writes fld_<t> with Ex, Ey, Ez, Bx, By, Bz (and Jx,
// Jy, Jz with save_tot_curr), each the node values CollectFld
// (savedata_routines.F90:207-261) samples every fsave_ratio nodes.  The
// reference writes HDF5; this image has no HDF5 library, so each output is a
// NumPy .npz archive (an uncompressed zip of .npy arrays) with the same
// dataset names and shapes: numpy.load reads it, and the arrays are in the
// reference's Fortran order.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pictor_hip.h"
#include "ctx.h"
#include "pictor_common.h"

using namespace pictor;

namespace {

// ---------------------------------------------------------------------------
// CollectFld on the device: output node (i1, j1, k1) is grid node
// (i0 + i1 r, j0 + j1 r, k0 + k1 r); the staggered components averaged onto
// the node in the reference's expression order, rounded to FP32 (real())
// ---------------------------------------------------------------------------
struct CollectP {
    int i0, j0, k0, r, nx1, ny1, nz1;
};

template <typename R>
__device__ inline double fv(const R *F, const float *L, long long o)
{
    double v = (double)F[o];
    if (L) v = v + (double)L[o];
    return v;
}

// which: 0 Ex, 1 Ey, 2 Ez, 3 Bx, 4 By, 5 Bz, 6 a node value (J)
template <typename R>
__global__ __launch_bounds__(256) void k_collect(GridP g, CollectP cp, int which, const R *__restrict__ F,
                                                 const float *__restrict__ L, double jscale, float *__restrict__ out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)cp.nx1 * cp.ny1 * cp.nz1;
    if (t >= n) return;
    const int i1 = (int)(t % cp.nx1);
    const long long q = t / cp.nx1;
    const int j1 = (int)(q % cp.ny1), k1 = (int)(q / cp.ny1);
    const int i = cp.i0 + i1 * cp.r, j = cp.j0 + j1 * cp.r, k = cp.k0 + k1 * cp.r;
    const long long o = fidx(g, i, j, k), sx = 1, sy = g.mx, sz = (long long)g.mx * g.my;
    const bool three = g.ndim == 3;
    double v = 0.0;
    switch (which) {
    case 0: v = (fv(F, L, o - sx) + fv(F, L, o)) * 0.5; break;
    case 1: v = (fv(F, L, o - sy) + fv(F, L, o)) * 0.5; break;
    case 2: v = three ? (fv(F, L, o - sz) + fv(F, L, o)) * 0.5 : fv(F, L, o); break;
    case 3:
        v = three ? (fv(F, L, o) + fv(F, L, o - sy) + fv(F, L, o - sz) + fv(F, L, o - sy - sz)) * 0.25
                  : (fv(F, L, o) + fv(F, L, o - sy)) * 0.5;
        break;
    case 4:
        v = three ? (fv(F, L, o) + fv(F, L, o - sx) + fv(F, L, o - sz) + fv(F, L, o - sx - sz)) * 0.25
                  : (fv(F, L, o) + fv(F, L, o - sx)) * 0.5;
        break;
    case 5: v = (fv(F, L, o) + fv(F, L, o - sx) + fv(F, L, o - sy) + fv(F, L, o - sx - sy)) * 0.25; break;
    default: v = fv(F, L, o) * jscale; break;
    }
    out[t] = (float)v;
}

// ---------------------------------------------------------------------------
// .npz writer: stored (uncompressed) zip members, each a version-1.0 .npy file
// ---------------------------------------------------------------------------
uint32_t crc32_update(uint32_t crc, const unsigned char *p, size_t n)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int b = 0; b < 8; b++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

struct Member {
    std::string name;
    std::vector<unsigned char> bytes;   // the whole .npy file
};

// .npy file: magic, version 1.0, header dictionary padded to a multiple of 64 bytes
Member npy(const std::string &name, const char *descr, const std::vector<long long> &shape, const void *data,
           size_t nbytes)
{
    std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': True, 'shape': (";
    for (size_t a = 0; a < shape.size(); a++) dict += std::to_string(shape[a]) + (shape.size() == 1 || a + 1 < shape.size() ? ", " : "");
    dict += "), }";
    size_t total = 10 + dict.size() + 1;
    const size_t pad = (64 - total % 64) % 64;
    dict += std::string(pad, ' ') + "\n";
    Member m;
    m.name = name + ".npy";
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    m.bytes.insert(m.bytes.end(), magic, magic + 8);
    const uint16_t hl = (uint16_t)dict.size();
    m.bytes.push_back((unsigned char)(hl & 0xFF));
    m.bytes.push_back((unsigned char)(hl >> 8));
    m.bytes.insert(m.bytes.end(), dict.begin(), dict.end());
    const unsigned char *d = (const unsigned char *)data;
    m.bytes.insert(m.bytes.end(), d, d + nbytes);
    return m;
}

void put16(std::vector<unsigned char> &b, uint16_t v) { b.push_back(v & 0xFF); b.push_back(v >> 8); }
void put32(std::vector<unsigned char> &b, uint32_t v)
{
    for (int s = 0; s < 32; s += 8) b.push_back((v >> s) & 0xFF);
}

// the archive goes to path.tmp and is renamed when complete
int write_npz(const std::string &path, const std::vector<Member> &members)
{
    std::vector<unsigned char> central;
    std::string tmp = path + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return fail(PICTOR_ERR_ARG, "output: cannot create " + tmp);
    uint64_t off = 0;
    bool ok = true;
    for (const Member &m : members) {
        if (m.bytes.size() >= 0xFFFFFFFFull || off >= 0xFFFFFFFFull) { ok = false; break; }
        const uint32_t crc = crc32_update(0, m.bytes.data(), m.bytes.size());
        const uint32_t sz = (uint32_t)m.bytes.size();
        std::vector<unsigned char> h;
        put32(h, 0x04034b50u); put16(h, 20); put16(h, 0); put16(h, 0);   // signature, version, flags, stored
        put16(h, 0); put16(h, 0x21);                                     // time, date (1980-01-01)
        put32(h, crc); put32(h, sz); put32(h, sz);
        put16(h, (uint16_t)m.name.size()); put16(h, 0);
        h.insert(h.end(), m.name.begin(), m.name.end());
        ok = ok && fwrite(h.data(), 1, h.size(), f) == h.size() &&
             fwrite(m.bytes.data(), 1, m.bytes.size(), f) == m.bytes.size();
        put32(central, 0x02014b50u); put16(central, 20); put16(central, 20); put16(central, 0); put16(central, 0);
        put16(central, 0); put16(central, 0x21);
        put32(central, crc); put32(central, sz); put32(central, sz);
        put16(central, (uint16_t)m.name.size()); put16(central, 0); put16(central, 0); put16(central, 0);
        put16(central, 0); put32(central, 0); put32(central, (uint32_t)off);
        central.insert(central.end(), m.name.begin(), m.name.end());
        off += h.size() + m.bytes.size();
    }
    std::vector<unsigned char> end;
    put32(end, 0x06054b50u); put16(end, 0); put16(end, 0);
    put16(end, (uint16_t)members.size()); put16(end, (uint16_t)members.size());
    put32(end, (uint32_t)central.size()); put32(end, (uint32_t)off); put16(end, 0);
    ok = ok && off < 0xFFFFFFFFull && fwrite(central.data(), 1, central.size(), f) == central.size() &&
         fwrite(end.data(), 1, end.size(), f) == end.size();
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        remove(tmp.c_str());
        return fail(PICTOR_ERR_ARG, "output: writing " + tmp + " failed (or an archive above 4 GiB)");
    }
    if (rename(tmp.c_str(), path.c_str()) != 0) return fail(PICTOR_ERR_ARG, "output: cannot rename to " + path);
    return PICTOR_OK;
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

extern "C" {

int pictor_save_energy(pictor_ctx *c, const char *path)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    if (!path) return fail(PICTOR_ERR_ARG, "null path");
    double d[9];
    int rc = pictor_diagnostics(c, d);   // CalcEnergy: B, E energy, then one per flavour
    if (rc) return rc;
    const double b = d[0], e = d[1], pe[2] = {d[2], d[3]};
    std::vector<Member> m;
    m.push_back(npy("Benergy", "<f8", {1}, &b, sizeof(double)));
    m.push_back(npy("Eenergy", "<f8", {1}, &e, sizeof(double)));
    m.push_back(npy("PrtlEnergy", "<f8", {2}, pe, sizeof(pe)));
    return write_npz(path, m);
}

int pictor_save_fields(pictor_ctx *c, const char *path, int32_t res, int32_t save_current)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    if (!path) return fail(PICTOR_ERR_ARG, "null path");
    HIPTRY(hipSetDevice(c->dev));
    if (c->cyl) return fail(PICTOR_ERR_ARG, "field files: Cartesian geometry only");
    if (res < 1) return fail(PICTOR_ERR_ARG, "field files: fsave_ratio must be >= 1");
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "field file between move_deposit and exchange_particles");
    const GridP &g = c->g;
    // GetSizeofCollectFld (savedata_routines.F90:114-173) for one rank over the box
    // [0, n]: nodes 3 .. 3 + r int(n / r), the right edge included
    CollectP cp{};
    cp.r = res;
    const int n3[3] = {g.nx, g.ny, g.nz};
    int cnt[3];
    for (int a = 0; a < 3; a++) {
        const int start = 3, end = res * (n3[a] / res) + 3;
        cnt[a] = (end - start) / res + 1;
    }
    cp.i0 = 3; cp.j0 = 3; cp.k0 = g.ndim == 3 ? 3 : 1;
    cp.nx1 = cnt[0]; cp.ny1 = cnt[1]; cp.nz1 = g.ndim == 3 ? cnt[2] : 1;
    const long long n = (long long)cp.nx1 * cp.ny1 * cp.nz1;
    float *d_out = nullptr;
    HIPTRY(dalloc(&d_out, (size_t)n));
    std::vector<float> h((size_t)n);
    std::vector<Member> mem;
    const char *names[9] = {"Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz"};
    const std::vector<long long> shape = {cp.nx1, cp.ny1, cp.nz1};
    const int nout = save_current ? 9 : 6;
    int rc = PICTOR_OK;
    if (save_current) {
        // SendRecvFlds(Jx,Jy,Jz) before the current is collected (savedata.F90:423)
        double *J[3] = {(double *)c->J[0], (double *)c->J[1], (double *)c->J[2]};
        launch_ghost_copy<double>(g, J, c->stream);
    }
    const double js = c->form == PICTOR_FORM_FP64 ? 1.0 : c->ph.qi * kFixInv;
    for (int q = 0; q < nout && !rc; q++) {
        const int which = q < 6 ? q : 6;
        if (q >= 6)
            hipLaunchKernelGGL(k_collect<double>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, g, cp, which,
                               (const double *)c->J[q - 6], (const float *)nullptr, js, d_out);
        else if (c->form == PICTOR_FORM_FP64)
            hipLaunchKernelGGL(k_collect<double>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, g, cp, which,
                               (const double *)(q < 3 ? c->E[q] : c->B[q - 3]), (const float *)nullptr, 1.0, d_out);
        else
            hipLaunchKernelGGL(k_collect<float>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, g, cp, which,
                               (const float *)(q < 3 ? c->E[q] : c->B[q - 3]), q < 3 ? (const float *)c->Elo[q] : nullptr,
                               1.0, d_out);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(PICTOR_ERR_HIP, std::string("field file: ") + hipGetErrorString(e));
        else mem.push_back(npy(names[q], "<f4", shape, h.data(), (size_t)n * sizeof(float)));
    }
    dfree(d_out);
    return rc ? rc : write_npz(path, mem);
}

}  // extern "C"
