// extern "C" entry points of include/lsbflow.h.
//
// Tier A (host pointers) stages the caller's arrays through device buffers the
// context owns, runs the same kernels as tier B and synchronises before
// returning (the reference's routines are synchronous, one rank per process).

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lsb_internal.h"
#include "physical.h"
#include "spectral.h"

namespace lsb {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

int dev_alloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return LSB_E_NOMEM;
    }
    return LSB_OK;
}

int upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
    LSB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return LSB_OK;
}

// Staging buffer `slot` of at least `bytes` (grown on demand, kept for reuse).
int stage_get(lsb_ctx* c, int slot, size_t bytes, void** p) {
    if ((int)c->stage.size() <= slot) {
        c->stage.resize(slot + 1, nullptr);
        c->stage_bytes.resize(slot + 1, 0);
    }
    if ((int)c->stage_in_graph.size() <= slot) c->stage_in_graph.resize(slot + 1, 0);
    if (c->capturing) c->stage_in_graph[slot] = 1;
    if (c->stage_bytes[slot] < bytes) {
        if (c->stage_locked && c->stage_in_graph[slot]) {
            set_error("staging buffers are locked while the context's steps are captured (lsb_ensemble)");
            return LSB_E_STATE;
        }
        if (c->stage[slot]) {
            // the old buffer may still be in use on a caller's stream
            LSB_HIP(hipDeviceSynchronize());
            LSB_HIP(hipFree(c->stage[slot]));
        }
        c->stage[slot] = nullptr;
        c->stage_bytes[slot] = 0;
        LSB_TRY(dev_alloc(&c->stage[slot], bytes));
        c->stage_bytes[slot] = bytes;
    }
    *p = c->stage[slot];
    return LSB_OK;
}

int prof_mark_begin(lsb_ctx* c, int cls, hipStream_t s, int* token, double flops, double bytes) {
    *token = -1;
    if (!c->profiling || !((c->prof_mask >> cls) & 1u)) return LSB_OK;
    lsb_ctx::ProfEv ev;
    ev.cls = cls;
    ev.flops = flops;
    ev.bytes = bytes;
    // timing-only events: no system-scope fence (cache write-back and invalidation) at
    // each record; lsb_profile_end synchronises the stream before reading them
    if (hipEventCreateWithFlags(&ev.a, hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&ev.b, hipEventDisableSystemFence) != hipSuccess)
        return LSB_OK;
    (void)hipEventRecord(ev.a, s);
    c->prof.push_back(ev);
    *token = (int)c->prof.size() - 1;
    return LSB_OK;
}
int prof_mark_end(lsb_ctx* c, int token, hipStream_t s) {
    if (token < 0 || !c->profiling) return LSB_OK;
    (void)hipEventRecord(c->prof[token].b, s);
    return LSB_OK;
}

// Column-major copy with leading dimension ldm and zero columns up to a multiple of 4
// (the matrix-core contraction consumes k in steps of 4).
static void pad_matrix(const std::vector<double>& src, int nrow, int ncol, int ldm, std::vector<double>& dst) {
    dst.assign((size_t)ldm * ((ncol + 3) / 4 * 4), 0.0);
    for (int j = 0; j < ncol; ++j)
        for (int i = 0; i < nrow; ++i) dst[(size_t)j * ldm + i] = src[(size_t)j * nrow + i];
}

// Upload grid-, matrix- and parameter-dependent device arrays.
int upload_context(lsb_ctx* c) {
    const int nzp = c->nzp, nzm = c->nzm;
    hipStream_t s = c->stream;
    auto put = [&](double** dptr, const std::vector<double>& h) -> int {
        if (!*dptr) LSB_TRY(dev_alloc((void**)dptr, h.size() * sizeof(double)));
        LSB_HIP(hipMemcpyAsync(*dptr, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
        return LSB_OK;
    };
    LSB_TRY(put(&c->d_wavx, c->wavx));
    LSB_TRY(put(&c->d_wavy, c->wavy));
    LSB_TRY(put(&c->d_zpts, c->zpts));
    std::vector<double> t;
    c->ldm_z = ((nzp + 63) / 64) * 64 + 64;
    pad_matrix(c->mdz, nzp, nzp, c->ldm_z, t);
    LSB_TRY(put(&c->d_mdz, t));
    pad_matrix(c->md2, nzp, nzp, c->ldm_z, t);
    LSB_TRY(put(&c->d_md2, t));
    pad_matrix(c->d, nzp, nzp, c->ldm_z, t);
    LSB_TRY(put(&c->d_d, t));
    {   // rows 0 and nzp-1 of md2, ld 64
        std::vector<double> b((size_t)kLdmBnd * ((nzp + 3) / 4 * 4), 0.0);
        for (int j = 0; j < nzp; ++j) {
            b[(size_t)j * kLdmBnd + 0] = c->md2[(size_t)j * nzp + 0];
            b[(size_t)j * kLdmBnd + 1] = c->md2[(size_t)j * nzp + nzp - 1];
        }
        LSB_TRY(put(&c->d_md2b, b));
    }
    const int ldh = ((nzm + 63) / 64) * 64 + 64;
    pad_matrix(c->smat, nzm, nzm, ldh, t);
    LSB_TRY(put(&c->d_smat, t));
    pad_matrix(c->simat, nzm, nzm, ldh, t);
    LSB_TRY(put(&c->d_simat, t));
    LSB_TRY(put(&c->d_diag, c->diag));
    std::vector<double> lift;
    lift.insert(lift.end(), c->hbot.begin(), c->hbot.end());
    lift.insert(lift.end(), c->htop.begin(), c->htop.end());
    lift.insert(lift.end(), c->d2hbot.begin(), c->d2hbot.end());
    lift.insert(lift.end(), c->d2htop.begin(), c->d2htop.end());
    // Einv applied to the interior values of each lifting function (rows 1..nz-1)
    std::vector<double> elift((size_t)4 * nzm);
    for (int q = 0; q < 4; ++q) {
        const double* h = lift.data() + (size_t)q * nzp;
        for (int r = 0; r < nzm; ++r) {
            long double acc = 0.0L;
            for (int m = 0; m < nzm; ++m) acc += (long double)c->simat[(size_t)m * nzm + r] * h[m + 1];
            elift[(size_t)q * nzm + r] = (double)acc;
        }
    }
    lift.insert(lift.end(), elift.begin(), elift.end());
    LSB_TRY(put(&c->d_lift, lift));
    LSB_TRY(zsplit_setup(c, elift));
    std::vector<double> par;
    par.insert(par.end(), c->ubar.begin(), c->ubar.end());
    par.insert(par.end(), c->ubarp.begin(), c->ubarp.end());
    par.insert(par.end(), c->ubarpp.begin(), c->ubarpp.end());
    LSB_TRY(put(&c->d_paral, par));
    std::vector<double> lam;  // kept alive until the stream is synchronised
    if (c->cfg.forcing == 1) {  // navierstokes.f90:119-187, from the current xpts
        lam.assign(c->nxpp, 0.0);
        sponge_lambda(c->nx, c->xpts.data(), c->cfg.xlen, c->cfg.u0, c->sponge[0], c->sponge[1], c->sponge[2],
                      lam.data());
        LSB_TRY(put(&c->d_lam, lam));
    }
    std::vector<double> del;
    del.insert(del.end(), c->delta_static.begin(), c->delta_static.end());
    del.insert(del.end(), c->delta_dyn.begin(), c->delta_dyn.end());
    del.insert(del.end(), c->delta_sigma.begin(), c->delta_sigma.end());
    LSB_TRY(put(&c->d_delta, del));
    std::vector<double> fz;
    fz.insert(fz.end(), c->fz_ab.begin(), c->fz_ab.end());
    fz.insert(fz.end(), c->fz_11.begin(), c->fz_11.end());
    fz.insert(fz.end(), c->fz_16.begin(), c->fz_16.end());
    fz.insert(fz.end(), c->fz_12.begin(), c->fz_12.end());
    LSB_TRY(put(&c->d_fz, fz));
    if (!c->d_smag) LSB_TRY(dev_alloc((void**)&c->d_smag, (size_t)nzp * sizeof(double)));
    if (3 * nzp + 128 > kRedDoubles) {
        set_error("nz too large for the small-table scratch");
        return LSB_E_ARG;
    }
    if (!c->d_red) LSB_TRY(dev_alloc((void**)&c->d_red, (size_t)kRedDoubles * sizeof(double)));
    {   // wall planes of the static model (sgsterm.f90:59-81)
        std::vector<int> walls;
        for (int k = 0; k < nzp; ++k)
            if (c->zpts[k] == 0.0 || std::fabs(c->zpts[k]) == 0.5) walls.push_back(k);
        c->nwalls = (int)walls.size();
        if (!c->d_walls) LSB_TRY(dev_alloc((void**)&c->d_walls, (size_t)nzp * sizeof(int)));
        if (!walls.empty())
            LSB_HIP(hipMemcpyAsync(c->d_walls, walls.data(), walls.size() * sizeof(int), hipMemcpyHostToDevice, s));
        LSB_HIP(hipStreamSynchronize(s));
    }
    LSB_HIP(hipStreamSynchronize(s));
    return LSB_OK;
}

static int check_ctx(const lsb_ctx* c) {
    if (!c) {
        set_error("null context");
        return LSB_E_ARG;
    }
    return LSB_OK;
}

}  // namespace lsb

using namespace lsb;

// ============================================================== context
extern "C" const char* lsb_last_error(void) { return g_err.c_str(); }

extern "C" int lsb_create(const lsb_config* cfg, lsb_ctx** out) {
    if (!cfg || !out) {
        set_error("lsb_create: null argument");
        return LSB_E_ARG;
    }
    *out = nullptr;
    const lsb_config& g = *cfg;
    if (g.nx < 4 || g.nx % 2 || g.ny < 2 || g.ny % 2 || g.nz < 4) {
        set_error("lsb_create: need even nx >= 4, even ny >= 2, nz >= 4");
        return LSB_E_ARG;
    }
    if (g.forcing != 0 && g.forcing != 1) {
        set_error("lsb_create: forcing must be 0 (periodic) or 1 (sponge layer)");
        return LSB_E_ARG;
    }
    if (g.dealiasflag < 0 || g.dealiasflag > 2) {
        set_error("lsb_create: dealiasflag must be 0, 1 or 2");
        return LSB_E_UNSUPPORTED;
    }
    if (g.bctype != 0 && g.bctype != 1 && g.bctype != 3) {
        set_error("lsb_create: bctype must be 0, 1 or 3");
        return LSB_E_UNSUPPORTED;
    }
    if ((g.dct_exact != 0 && g.dct_exact != 1) || g.top_bc < 0 || g.top_bc > 2) {
        set_error("lsb_create: dct_exact must be 0 or 1, top_bc 0, 1 or 2");
        return LSB_E_ARG;
    }
    if (g.nut_ceil0 != 0 && g.nut_ceil0 != 1) {
        set_error("lsb_create: nut_ceil0 must be 0 or 1");
        return LSB_E_ARG;
    }
    if (g.nut_ceil0 == 1 && g.sgsflag == 1 && (g.sgsmodel == 1 || g.sgsmodel == 3)) {
        set_error("lsb_create: nut_ceil0 applies to the static and sigma models' filter widths "
                  "(the dynamic and interscale models compute their coefficients from theirs)");
        return LSB_E_ARG;
    }
    if (g.top_bc == 2 && g.bctype == 1) {
        set_error("lsb_create: top_bc 2 weights the ceiling by the bctype 3 blowing profile (not bctype 1)");
        return LSB_E_ARG;
    }
    lsb_ctx* c = new lsb_ctx();
    c->cfg = g;
    c->nx = g.nx;
    c->ny = g.ny;
    c->nz = g.nz;
    c->nzp = g.nz + 1;
    c->nzm = g.nz - 1;
    c->nxh = g.nx / 2;
    c->nyh = g.ny / 2;
    c->nxpp = g.nx + 2;
    c->plane_s = (size_t)g.nx * g.ny;
    c->plane_p = (size_t)c->nxpp * g.ny;
    c->nspec = c->plane_s * c->nzp;
    c->nphys = c->plane_p * c->nzp;
    c->dt = c->dt2 = c->dt3 = g.dt;
    c->device = g.device;
    int rc = build_grid(c);
    if (rc == LSB_OK) rc = build_operators(c);
    if (rc != LSB_OK) {
        delete c;
        return rc;
    }
    c->ubar.assign(c->nzp, 0.0);
    c->ubarp.assign(c->nzp, 0.0);
    c->ubarpp.assign(c->nzp, 0.0);
    c->wbar.assign(c->nzp, 0.0);
    c->wbarpp.assign(c->nzp, 0.0);
    hipError_t e = hipSetDevice(g.device);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        set_error(std::string("lsb_create: ") + hipGetErrorString(e));
        delete c;
        return LSB_E_HIP;
    }
    rc = hor_plan_create(&c->hp, c->nx, c->ny);
    if (rc == LSB_OK) rc = upload_context(c);
    if (rc == LSB_OK) rc = noalloc_resolve(g.device);
    if (rc != LSB_OK) {
        lsb_destroy(c);
        return rc;
    }
    *out = c;
    return LSB_OK;
}

extern "C" int lsb_destroy(lsb_ctx* c) {
    if (!c) return LSB_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    step_free(c);
    temp_free(c);
    double* ds[] = {c->d_wavx, c->d_wavy, c->d_zpts, c->d_mdz, c->d_md2, c->d_md2b, c->d_d, c->d_smat,
                    c->d_simat, c->d_diag, c->d_lift, c->d_paral, c->d_delta, c->d_fz, c->d_smag, c->d_red, c->d_ftab, c->d_diag_red, c->d_lam, c->d_tavg};
    for (double* p : ds)
        if (p) (void)hipFree(p);
    if (c->d_walls) (void)hipFree(c->d_walls);
    for (double* p : {c->zs_dz.d_m, c->zs_h1.d_m, c->zs_h2.d_m, c->d_helm_s})
        if (p) (void)hipFree(p);
    for (float* p : {c->zs_dz.d_m32, c->zs_h1.d_m32, c->zs_h2.d_m32})
        if (p) (void)hipFree(p);
    for (void* p : c->stage)
        if (p) (void)hipFree(p);
    if (c->st) (void)hipFree(c->st);
    hor_plan_destroy(&c->hp);
    hor_plan_destroy(&c->hp_lg);
    for (auto& ev : c->prof) {
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LSB_OK;
}

extern "C" int lsb_set_grid(lsb_ctx* c, const double* wavx, const double* wavy, const double* zpts, const double* gf,
                            const double* xpts, double dx, double dy, double xl) {
    LSB_TRY(check_ctx(c));
    if (wavx) c->wavx.assign(wavx, wavx + c->nx);
    if (wavy) c->wavy.assign(wavy, wavy + c->ny);
    if (zpts) c->zpts.assign(zpts, zpts + c->nzp);
    if (gf) c->gf.assign(gf, gf + c->nzp);
    if (xpts) c->xpts.assign(xpts, xpts + c->nx);
    c->dx = dx;
    c->dy = dy;
    c->xl = xl;
    for (int i = 0; i < c->nx; ++i) c->wavxx[i] = c->wavx[i] * c->wavx[i];
    for (int j = 0; j < c->ny; ++j) c->wavyy[j] = c->wavy[j] * c->wavy[j];
    LSB_TRY(build_derived(c));
    return upload_context(c);
}

extern "C" int lsb_set_mats(lsb_ctx* c, const double* d, const double* d2, const double* diag, const double* smat,
                            const double* simat) {
    LSB_TRY(check_ctx(c));
    const size_t n2 = (size_t)c->nzp * c->nzp, m2 = (size_t)c->nzm * c->nzm;
    if (d) c->d.assign(d, d + n2);
    if (d2) c->d2.assign(d2, d2 + n2);
    if (diag) c->diag.assign(diag, diag + c->nzm);
    if (smat) c->smat.assign(smat, smat + m2);
    if (simat) c->simat.assign(simat, simat + m2);
    LSB_TRY(build_derived(c));
    return upload_context(c);
}

extern "C" int lsb_set_paral(lsb_ctx* c, const double* ubar, const double* ubarp, const double* ubarpp,
                             const double* wbar, const double* wbarpp) {
    LSB_TRY(check_ctx(c));
    const int n = c->nzp;
    if (ubar) c->ubar.assign(ubar, ubar + n);
    if (ubarp) c->ubarp.assign(ubarp, ubarp + n);
    if (ubarpp) c->ubarpp.assign(ubarpp, ubarpp + n);
    if (wbar) c->wbar.assign(wbar, wbar + n);
    if (wbarpp) c->wbarpp.assign(wbarpp, wbarpp + n);
    lifting_functions(c);
    return upload_context(c);
}

extern "C" int lsb_get_grid(const lsb_ctx* c, double* wavx, double* wavy, double* zpts, double* gf, double* xpts,
                            double* dxdyxl) {
    LSB_TRY(check_ctx(c));
    if (wavx) std::memcpy(wavx, c->wavx.data(), c->nx * sizeof(double));
    if (wavy) std::memcpy(wavy, c->wavy.data(), c->ny * sizeof(double));
    if (zpts) std::memcpy(zpts, c->zpts.data(), c->nzp * sizeof(double));
    if (gf) std::memcpy(gf, c->gf.data(), c->nzp * sizeof(double));
    if (xpts) std::memcpy(xpts, c->xpts.data(), c->nx * sizeof(double));
    if (dxdyxl) {
        dxdyxl[0] = c->dx;
        dxdyxl[1] = c->dy;
        dxdyxl[2] = c->xl;
    }
    return LSB_OK;
}

extern "C" int lsb_get_mats(const lsb_ctx* c, double* dor, double* d, double* d2, double* mdz, double* md2,
                            double* diag, double* smat, double* simat) {
    LSB_TRY(check_ctx(c));
    const size_t n2 = (size_t)c->nzp * c->nzp * sizeof(double), m2 = (size_t)c->nzm * c->nzm * sizeof(double);
    if (dor) std::memcpy(dor, c->dor.data(), n2);
    if (d) std::memcpy(d, c->d.data(), n2);
    if (d2) std::memcpy(d2, c->d2.data(), n2);
    if (mdz) std::memcpy(mdz, c->mdz.data(), n2);
    if (md2) std::memcpy(md2, c->md2.data(), n2);
    if (diag) std::memcpy(diag, c->diag.data(), c->nzm * sizeof(double));
    if (smat) std::memcpy(smat, c->smat.data(), m2);
    if (simat) std::memcpy(simat, c->simat.data(), m2);
    return LSB_OK;
}

// ================================================= host-only setup routines
extern "C" int lsb_build_cheb_zder_matrix(int nzp, double xl, const double* gf, double* dor, double* d, double* d2) {
    if (nzp < 3 || !gf || !dor || !d || !d2) {
        set_error("build_cheb_zder_matrix: bad arguments");
        return LSB_E_ARG;
    }
    cheb_matrix(nzp, xl, gf, dor, d, d2);
    return LSB_OK;
}

extern "C" int lsb_dct_diff(int n1, int n2, int n, double xl, const double* gf, const double* y, double* dydx, int p) {
    if (n < 3 || n1 < 1 || n2 < 1 || !gf || !y || !dydx || p < 1) {
        set_error("dct_diff: bad arguments");
        return LSB_E_ARG;
    }
    std::vector<double> m((size_t)n * n);
    dct_diff_matrix(n, xl, gf, p, m.data());
    const size_t P = (size_t)n1 * n2;
    std::vector<long double> acc(n);
    for (size_t col = 0; col < P; ++col) {
        for (int k = 0; k < n; ++k) {
            long double s = 0;
            for (int j = 0; j < n; ++j) s += (long double)m[(size_t)j * n + k] * (long double)y[col + P * j];
            acc[k] = s;
        }
        for (int k = 0; k < n; ++k) dydx[col + P * k] = (double)acc[k];
    }
    return LSB_OK;
}

extern "C" int lsb_diagonalize_matrix(const double* D2, int n, double* lamda, double* e, double* ei) {
    if (!D2 || n < 1 || !lamda || !e || !ei) {
        set_error("diagonalize_matrix: bad arguments");
        return LSB_E_ARG;
    }
    return eigen_decompose(D2, n, lamda, e, ei);
}

// ================================================================ tier A
namespace {

struct Stage {
    lsb_ctx* c;
    explicit Stage(lsb_ctx* cc) : c(cc) {}
    int get(int slot, size_t nreal, double** p) { return stage_get(c, slot, nreal * sizeof(double), (void**)p); }
    int in(double* d, const double* h, size_t n) {
        LSB_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        return LSB_OK;
    }
    int out(double* h, const double* d, size_t n) {
        LSB_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        return LSB_OK;
    }
    int sync() {
        LSB_HIP(hipStreamSynchronize(c->stream));
        return LSB_OK;
    }
};

int set_device(lsb_ctx* c) {
    LSB_HIP(hipSetDevice(c->device));
    return LSB_OK;
}

// the on-the-fly time averages back to zero (no-op while averaging is off)
int tavg_zero(lsb_ctx* c) {
    if (c->d_tavg) LSB_HIP(hipMemsetAsync(c->d_tavg, 0, 5 * c->nspec * sizeof(double), c->stream));
    return LSB_OK;
}

}  // namespace

#define LSB_ENTER(c)          \
    LSB_TRY(check_ctx(c));    \
    LSB_TRY(set_device(c));   \
    Stage st(c)

extern "C" int lsb_horfft(lsb_ctx* c, double* ain, double* ainf, int is, int nfields) {
    LSB_ENTER(c);
    if (!ain || !ainf || (is != 1 && is != -1) || nfields < 1) {
        set_error("horfft: bad arguments");
        return LSB_E_ARG;
    }
    double *dp, *ds;
    LSB_TRY(st.get(0, c->nphys * nfields, &dp));
    LSB_TRY(st.get(1, c->nspec * nfields, &ds));
    for (int f0 = 0; f0 < nfields; f0 += kMaxFields) {
        const int nb = std::min(kMaxFields, nfields - f0);
        const double* pin[kMaxFields];
        double* pout[kMaxFields];
        for (int q = 0; q < nb; ++q) {
            pin[q] = (is == -1) ? dp + (size_t)(f0 + q) * c->nphys : ds + (size_t)(f0 + q) * c->nspec;
            pout[q] = (is == -1) ? ds + (size_t)(f0 + q) * c->nspec : dp + (size_t)(f0 + q) * c->nphys;
        }
        if (is == -1) {
            LSB_TRY(st.in(dp + (size_t)f0 * c->nphys, ain + (size_t)f0 * c->nphys, c->nphys * nb));
            LSB_TRY(horfft_fwd(c, c->hp, pin, pout, nb, c->nzp, c->stream));
            LSB_TRY(st.out(ainf + (size_t)f0 * c->nspec, ds + (size_t)f0 * c->nspec, c->nspec * nb));
        } else {
            LSB_TRY(st.in(ds + (size_t)f0 * c->nspec, ainf + (size_t)f0 * c->nspec, c->nspec * nb));
            LSB_TRY(horfft_inv(c, c->hp, pin, pout, nb, c->nzp, c->stream));
            LSB_TRY(st.out(ain + (size_t)f0 * c->nphys, dp + (size_t)f0 * c->nphys, c->nphys * nb));
        }
    }
    return st.sync();
}

extern "C" int lsb_horfft_dev(lsb_ctx* c, double* ain, double* ainf, int is, int nfields, void* stream) {
    LSB_TRY(check_ctx(c));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    for (int f0 = 0; f0 < nfields; f0 += kMaxFields) {
        const int nb = std::min(kMaxFields, nfields - f0);
        const double* pin[kMaxFields];
        double* pout[kMaxFields];
        for (int q = 0; q < nb; ++q) {
            if (is == -1) {
                pin[q] = ain + (size_t)(f0 + q) * c->nphys;
                pout[q] = ainf + (size_t)(f0 + q) * c->nspec;
            } else {
                pin[q] = ainf + (size_t)(f0 + q) * c->nspec;
                pout[q] = ain + (size_t)(f0 + q) * c->nphys;
            }
        }
        if (is == -1)
            LSB_TRY(horfft_fwd(c, c->hp, pin, pout, nb, c->nzp, s));
        else
            LSB_TRY(horfft_inv(c, c->hp, pin, pout, nb, c->nzp, s));
    }
    return LSB_OK;
}

extern "C" int lsb_norm(lsb_ctx* c, double* uf, double* vf, double* wf, double* tempf) {
    LSB_ENTER(c);
    double* hs[4] = {uf, vf, wf, tempf};
    double* d;
    LSB_TRY(st.get(0, c->nspec * 4, &d));
    double* dl[4];
    int n = 0;
    for (int q = 0; q < 4; ++q)
        if (hs[q]) {
            dl[n] = d + (size_t)n * c->nspec;
            LSB_TRY(st.in(dl[n], hs[q], c->nspec));
            ++n;
        }
    LSB_TRY(sp_norm(c, dl, n, c->stream));
    n = 0;
    for (int q = 0; q < 4; ++q)
        if (hs[q]) {
            LSB_TRY(st.out(hs[q], dl[n], c->nspec));
            ++n;
        }
    return st.sync();
}

// At one rank the transposes are pure reorders between the x-row layout and
// the spectral layout (horfft.f90:303-315, 333-345, 401-413, 431-443).
extern "C" int lsb_dataflip_xtoy(lsb_ctx* c, const double* x, double* xf) {
    LSB_TRY(check_ctx(c));
    for (int k = 0; k < c->nzp; ++k)
        for (int j = 0; j < c->ny; ++j)
            std::memcpy(xf + ((size_t)k * c->ny + j) * c->nx, x + ((size_t)k * c->ny + j) * c->nxpp,
                        c->nx * sizeof(double));
    return LSB_OK;
}
extern "C" int lsb_dataflip_ytox(lsb_ctx* c, double* x, const double* xf) {
    LSB_TRY(check_ctx(c));
    for (int k = 0; k < c->nzp; ++k)
        for (int j = 0; j < c->ny; ++j)
            std::memcpy(x + ((size_t)k * c->ny + j) * c->nxpp, xf + ((size_t)k * c->ny + j) * c->nx,
                        c->nx * sizeof(double));
    return LSB_OK;
}

// generic one-in one-out spectral op
template <class F>
static int spec_unary(lsb_ctx* c, const double* in, double* out, F fn) {
    LSB_ENTER(c);
    if (!in || !out) {
        set_error("null array");
        return LSB_E_ARG;
    }
    double *a, *b;
    LSB_TRY(st.get(0, c->nspec, &a));
    LSB_TRY(st.get(1, c->nspec, &b));
    LSB_TRY(st.in(a, in, c->nspec));
    LSB_TRY(fn(a, b));
    LSB_TRY(st.out(out, b, c->nspec));
    return st.sync();
}

extern "C" int lsb_dfdx(lsb_ctx* c, const double* uf, double* dudxf) {
    return spec_unary(c, uf, dudxf, [&](double* a, double* b) { return sp_dfdx(c, a, b, c->stream); });
}
extern "C" int lsb_dfdy(lsb_ctx* c, const double* uf, double* dudyf) {
    return spec_unary(c, uf, dudyf, [&](double* a, double* b) { return sp_dfdy(c, a, b, c->stream); });
}
extern "C" int lsb_dfdz(lsb_ctx* c, const double* uf, double* dudzf) {
    return spec_unary(c, uf, dudzf, [&](double* a, double* b) {
        const double* in[1] = {a};
        double* out[1] = {b};
        return zcontract(c, c->d_mdz, c->ldm_z, c->nzp, c->nzp, in, out, 1, c->stream);
    });
}
extern "C" int lsb_d2fdz2(lsb_ctx* c, const double* uf, double* d2f) {
    return spec_unary(c, uf, d2f, [&](double* a, double* b) {
        const double* in[1] = {a};
        double* out[1] = {b};
        return zcontract(c, c->d_md2, c->ldm_z, c->nzp, c->nzp, in, out, 1, c->stream);
    });
}
extern "C" int lsb_dfdz_dev(lsb_ctx* c, const double* uf, double* dudzf, int nfields, void* stream) {
    LSB_TRY(check_ctx(c));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    for (int f0 = 0; f0 < nfields; f0 += kMaxFields) {
        const int nb = std::min(kMaxFields, nfields - f0);
        const double* in[kMaxFields];
        double* out[kMaxFields];
        for (int q = 0; q < nb; ++q) {
            in[q] = uf + (size_t)(f0 + q) * c->nspec;
            out[q] = dudzf + (size_t)(f0 + q) * c->nspec;
        }
        LSB_TRY(zcontract(c, c->d_mdz, c->ldm_z, c->nzp, c->nzp, in, out, nb, s));
    }
    return LSB_OK;
}

// derivatives.f90:135-158 with n = nzp: dct_diff (n > 31) or matmul(d, .) once (any p).
extern "C" int lsb_dpdz(lsb_ctx* c, int n1, int n2, const double* u, double* dudz, int p) {
    LSB_ENTER(c);
    if (!u || !dudz || n1 < 1 || n2 < 1 || (p != 1 && p != 2)) {
        set_error("dpdz: bad arguments");
        return LSB_E_ARG;
    }
    const size_t P = (size_t)n1 * n2, n = P * c->nzp;
    double *a, *b;
    LSB_TRY(st.get(0, n, &a));
    LSB_TRY(st.get(1, n, &b));
    LSB_TRY(st.in(a, u, n));
    const double* M = (c->nzp > 31) ? (p == 1 ? c->d_mdz : c->d_md2) : c->d_d;
    // contraction over planes of P columns: reuse the kernel with P = n1*n2
    const size_t save = c->plane_s;
    c->plane_s = P;
    const double* in[1] = {a};
    double* out[1] = {b};
    int rc = zcontract(c, M, c->ldm_z, c->nzp, c->nzp, in, out, 1, c->stream);
    c->plane_s = save;
    LSB_TRY(rc);
    LSB_TRY(st.out(dudz, b, n));
    return st.sync();
}

extern "C" int lsb_divuv(lsb_ctx* c, const double* uf, const double* vf, double* divf) {
    LSB_ENTER(c);
    double *a, *b, *o;
    LSB_TRY(st.get(0, c->nspec, &a));
    LSB_TRY(st.get(1, c->nspec, &b));
    LSB_TRY(st.get(2, c->nspec, &o));
    LSB_TRY(st.in(a, uf, c->nspec));
    LSB_TRY(st.in(b, vf, c->nspec));
    LSB_TRY(sp_divuv(c, a, b, o, c->stream));
    LSB_TRY(st.out(divf, o, c->nspec));
    return st.sync();
}

__global__ void k_add(const double* a, const double* b, double* o, long n) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        o[t] = a[t] + b[t];
}

extern "C" int lsb_div(lsb_ctx* c, const double* uf, const double* vf, const double* wf, double* divf) {
    LSB_ENTER(c);
    double *a, *b, *w, *o, *t;
    LSB_TRY(st.get(0, c->nspec, &a));
    LSB_TRY(st.get(1, c->nspec, &b));
    LSB_TRY(st.get(2, c->nspec, &w));
    LSB_TRY(st.get(3, c->nspec, &o));
    LSB_TRY(st.get(4, c->nspec, &t));
    LSB_TRY(st.in(a, uf, c->nspec));
    LSB_TRY(st.in(b, vf, c->nspec));
    LSB_TRY(st.in(w, wf, c->nspec));
    const double* in[1] = {w};
    double* out[1] = {o};
    LSB_TRY(zcontract(c, c->d_mdz, c->ldm_z, c->nzp, c->nzp, in, out, 1, c->stream));  // divf = dfdz(w)
    LSB_TRY(sp_divuv(c, a, b, t, c->stream));                                          // dqf
    const long n = (long)c->nspec;
    hipLaunchKernelGGL(k_add, dim3((unsigned)std::min<long>((n + 255) / 256, 65535L * 8)), dim3(256), 0, c->stream, t,
                       o, o, n);  // divf = dqf + divf
    LSB_HIP(hipGetLastError());
    LSB_TRY(st.out(divf, o, c->nspec));
    return st.sync();
}

extern "C" int lsb_lapl(lsb_ctx* c, const double* uf, double* lapluf) {
    LSB_ENTER(c);
    double *a, *b, *o;
    LSB_TRY(st.get(0, c->nspec, &a));
    LSB_TRY(st.get(1, c->nspec, &b));
    LSB_TRY(st.get(2, c->nspec, &o));
    LSB_TRY(st.in(a, uf, c->nspec));
    const double* in[1] = {a};
    double* out[1] = {b};
    LSB_TRY(zcontract(c, c->d_md2, c->ldm_z, c->nzp, c->nzp, in, out, 1, c->stream));
    LSB_TRY(sp_lapl_combine(c, b, a, o, c->stream));
    LSB_TRY(st.out(lapluf, o, c->nspec));
    return st.sync();
}

__global__ void k_abs_sum(const double* __restrict__ p, int nx, int nxpp, long nrows, double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    const long n = nrows * nx;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long r = t / nx;
        const int i = (int)(t - r * nx);
        s += fabs(p[r * nxpp + i]);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

extern "C" int lsb_divcheck(lsb_ctx* c, const double* uf, const double* vf, const double* wf, double* diverr) {
    LSB_ENTER(c);
    if (!diverr) return LSB_E_ARG;
    double *a, *b, *w, *t1, *t2, *ph;
    LSB_TRY(st.get(0, c->nspec, &a));
    LSB_TRY(st.get(1, c->nspec, &b));
    LSB_TRY(st.get(2, c->nspec, &w));
    LSB_TRY(st.get(3, c->nspec, &t1));
    LSB_TRY(st.get(4, c->nspec, &t2));
    LSB_TRY(st.get(5, c->nphys, &ph));
    LSB_TRY(st.in(a, uf, c->nspec));
    LSB_TRY(st.in(b, vf, c->nspec));
    LSB_TRY(st.in(w, wf, c->nspec));
    LSB_TRY(sp_dfdx(c, a, t1, c->stream));
    LSB_TRY(sp_dfdy(c, b, t2, c->stream));
    const long n = (long)c->nspec;
    const dim3 g((unsigned)std::min<long>((n + 255) / 256, 65535L * 8));
    hipLaunchKernelGGL(k_add, g, dim3(256), 0, c->stream, t1, t2, t1, n);
    const double* in[1] = {w};
    double* out[1] = {t2};
    LSB_TRY(zcontract(c, c->d_mdz, c->ldm_z, c->nzp, c->nzp, in, out, 1, c->stream));
    hipLaunchKernelGGL(k_add, g, dim3(256), 0, c->stream, t1, t2, t1, n);
    const double* si[1] = {t1};
    double* so[1] = {ph};
    LSB_TRY(horfft_inv(c, c->hp, si, so, 1, c->nzp, c->stream));
    const int nb = 256;
    hipLaunchKernelGGL(k_abs_sum, dim3(nb), dim3(256), 0, c->stream, ph, c->nx, c->nxpp, (long)c->nzp * c->ny,
                       c->d_red);
    std::vector<double> part(nb);
    LSB_HIP(hipMemcpyAsync(part.data(), c->d_red, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LSB_TRY(st.sync());
    double s = 0.0;
    for (double v : part) s += v;
    *diverr = s / ((double)c->nx * c->ny * c->nzp);
    return LSB_OK;
}

// ------------------------------------------------------------- helmholtz
static int check_mats(lsb_ctx* c, int n, const double* diag, const double* smat, const double* simat) {
    if (n != c->nzm) {
        set_error("solve_helmholtz: n must be nzm");
        return LSB_E_ARG;
    }
    const size_t m2 = (size_t)n * n;
    bool same = (!diag || std::memcmp(diag, c->diag.data(), n * sizeof(double)) == 0) &&
                (!smat || std::memcmp(smat, c->smat.data(), m2 * sizeof(double)) == 0) &&
                (!simat || std::memcmp(simat, c->simat.data(), m2 * sizeof(double)) == 0);
    if (!same) return lsb_set_mats(c, nullptr, nullptr, diag, smat, simat);
    return LSB_OK;
}

extern "C" int lsb_solve_helmholtz(lsb_ctx* c, double* rhs, double dtnui, int n, const double* diag,
                                   const double* smat, const double* simat) {
    LSB_ENTER(c);
    LSB_TRY(check_mats(c, n, diag, smat, simat));
    double* a;
    LSB_TRY(st.get(0, c->nspec, &a));
    LSB_TRY(st.in(a, rhs, c->nspec));
    double* r[1] = {a};
    LSB_TRY(helmholtz(c, r, nullptr, nullptr, 1, dtnui, c->stream));
    LSB_TRY(st.out(rhs, a, c->nspec));
    return st.sync();
}

extern "C" int lsb_solve_helmholtz_dev(lsb_ctx* c, double* rhs, double dtnui, int nfields, void* stream) {
    LSB_TRY(check_ctx(c));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    for (int f0 = 0; f0 < nfields; f0 += kMaxFields) {
        const int nb = std::min(kMaxFields, nfields - f0);
        double* r[kMaxFields];
        for (int q = 0; q < nb; ++q) r[q] = rhs + (size_t)(f0 + q) * c->nspec;
        LSB_TRY(helmholtz(c, r, nullptr, nullptr, nb, dtnui, s));
    }
    return LSB_OK;
}

extern "C" int lsb_solve_nhom_helmholtz(lsb_ctx* c, double* rhs, const double* bc, double dtnui, const double* diag,
                                        const double* smat, const double* simat) {
    LSB_ENTER(c);
    LSB_TRY(check_mats(c, c->nzm, diag, smat, simat));
    double *a, *b;
    LSB_TRY(st.get(0, c->nspec, &a));
    LSB_TRY(st.get(1, c->nspec, &b));
    LSB_TRY(st.in(a, rhs, c->nspec));
    LSB_TRY(st.in(b, bc, c->nspec));
    double* r[1] = {a};
    const double* bt[1] = {b};
    const double* bb[1] = {b + (size_t)(c->nzp - 1) * c->plane_s};
    LSB_TRY(helmholtz(c, r, bt, bb, 1, dtnui, c->stream));
    LSB_TRY(st.out(rhs, a, c->nspec));
    return st.sync();
}

// ------------------------------------------------------------ sgs_models

static int tensor_in(Stage& st, lsb_ctx* c, int slot, const double* h, int nf, double** d) {
    LSB_TRY(st.get(slot, c->nphys * nf, d));
    if (h) LSB_TRY(st.in(*d, h, c->nphys * nf));
    return LSB_OK;
}

extern "C" int lsb_sigma_model(lsb_ctx* c, const double* duidxj, double* tauij) {
    LSB_ENTER(c);
    double *g, *nlt, *u, *t;
    LSB_TRY(tensor_in(st, c, 0, duidxj, 9, &g));
    LSB_TRY(st.get(1, c->nphys * 3, &nlt));
    LSB_TRY(st.get(2, c->nphys * 3, &u));
    LSB_TRY(st.get(3, c->nphys * 9, &t));
    LSB_HIP(hipMemsetAsync(u, 0, c->nphys * 3 * sizeof(double), c->stream));
    // zero base flow and u: h = 1/2 * 0 + tau = tau
    double* zero;
    LSB_TRY(st.get(4, 3 * (size_t)c->nzp, &zero));
    LSB_HIP(hipMemsetAsync(zero, 0, 3 * (size_t)c->nzp * sizeof(double), c->stream));
    NlArgs a{};
    for (int l = 0; l < 3; ++l) {
        a.ui[l] = u + l * c->nphys;
        a.nlt[l] = nlt + l * c->nphys;
    }
    for (int q = 0; q < 9; ++q) {
        a.g[q] = g + q * c->nphys;
        a.h[q] = t + q * c->nphys;
    }
    a.ubar = zero;
    a.ubarp = zero;
    a.delta = c->d_delta + 2 * c->nzp;
    LSB_TRY(ph_nonlinear(c, NL_SIGMA, a, c->stream));
    double* tl[9];
    for (int q = 0; q < 9; ++q) tl[q] = t + q * c->nphys;
    LSB_TRY(ph_zero_pads(c, tl, 9, c->stream));
    LSB_TRY(st.out(tauij, t, c->nphys * 9));
    return st.sync();
}

extern "C" int lsb_static_smagorinsky(lsb_ctx* c, double cs, const double* duidxj, double* tauij) {
    LSB_ENTER(c);
    double *g, *nlt, *u, *t, *zero;
    LSB_TRY(tensor_in(st, c, 0, duidxj, 9, &g));
    LSB_TRY(st.get(1, c->nphys * 3, &nlt));
    LSB_TRY(st.get(2, c->nphys * 3, &u));
    LSB_TRY(st.get(3, c->nphys * 9, &t));
    LSB_TRY(st.get(4, 3 * (size_t)c->nzp, &zero));
    LSB_HIP(hipMemsetAsync(u, 0, c->nphys * 3 * sizeof(double), c->stream));
    LSB_HIP(hipMemsetAsync(zero, 0, 3 * (size_t)c->nzp * sizeof(double), c->stream));
    LSB_TRY(ph_smag_factor(c, g + 6 * c->nphys, nullptr, cs, c->stream));
    NlArgs a{};
    for (int l = 0; l < 3; ++l) {
        a.ui[l] = u + l * c->nphys;
        a.nlt[l] = nlt + l * c->nphys;
    }
    for (int q = 0; q < 9; ++q) {
        a.g[q] = g + q * c->nphys;
        a.h[q] = t + q * c->nphys;
    }
    a.ubar = zero;
    a.ubarp = zero;
    a.smag = c->d_smag;
    LSB_TRY(ph_nonlinear(c, NL_STATIC, a, c->stream));
    double* tl[9];
    for (int q = 0; q < 9; ++q) tl[q] = t + q * c->nphys;
    LSB_TRY(ph_zero_pads(c, tl, 9, c->stream));
    LSB_TRY(st.out(tauij, t, c->nphys * 9));
    return st.sync();
}

extern "C" int lsb_dynamic_smagorinsky(lsb_ctx* c, const double* ui, const double* uiuj, const double* duidxj,
                                       double* tauij) {
    LSB_ENTER(c);
    double *u, *uu, *g, *t, *w;
    LSB_TRY(tensor_in(st, c, 0, ui, 3, &u));
    LSB_TRY(tensor_in(st, c, 1, uiuj, 9, &uu));
    LSB_TRY(tensor_in(st, c, 2, duidxj, 9, &g));
    LSB_TRY(st.get(3, c->nphys * 9, &t));
    LSB_TRY(st.get(4, c->nphys * (1 + 9 * 5 + 3 + 4 + 16), &w));
    LSB_TRY(dyn_full(c, u, uu, g, t, w, c->stream));
    double* tl[9];
    for (int q = 0; q < 9; ++q) tl[q] = t + q * c->nphys;
    LSB_TRY(ph_zero_pads(c, tl, 9, c->stream));
    LSB_TRY(st.out(tauij, t, c->nphys * 9));
    return st.sync();
}

extern "C" int lsb_interscale_model(lsb_ctx* c, const double* ui, double* tauij) {
    LSB_ENTER(c);
    double *u, *t, *w;
    LSB_TRY(tensor_in(st, c, 0, ui, 3, &u));
    LSB_TRY(st.get(3, c->nphys * 9, &t));
    LSB_TRY(st.get(4, c->nphys * 40, &w));
    LSB_TRY(isc_full(c, u, t, w, c->stream));
    double* tl[9];
    for (int q = 0; q < 9; ++q) tl[q] = t + q * c->nphys;
    LSB_TRY(ph_zero_pads(c, tl, 9, c->stream));
    LSB_TRY(st.out(tauij, t, c->nphys * 9));
    return st.sync();
}

extern "C" int lsb_tau_sgs(lsb_ctx* c, const double* ui, const double* uiuj, const double* duidxj, double* tauij) {
    LSB_TRY(check_ctx(c));
    switch (c->cfg.sgsmodel) {
        case 0: return lsb_static_smagorinsky(c, c->cfg.cs, duidxj, tauij);
        case 1: return lsb_dynamic_smagorinsky(c, ui, uiuj, duidxj, tauij);
        case 2: return lsb_sigma_model(c, duidxj, tauij);
        case 3: return lsb_interscale_model(c, ui, tauij);
        case 4:  // TNS: no case in sgsterm.f90:21-35, tauij keeps the caller's zero
            if (!tauij) return LSB_E_ARG;
            std::memset(tauij, 0, c->nphys * 9 * sizeof(double));
            return LSB_OK;
        default: set_error("tau_sgs: sgsmodel must be 0..4"); return LSB_E_ARG;
    }
}

// --------------------------------------------------------------- filters
extern "C" int lsb_filter_q_3pt(lsb_ctx* c, const double* q, double* q_filt, int n4, int n5, double a, double b,
                                int interpflag) {
    LSB_ENTER(c);
    const int nf = n4 * n5;
    if (!q || !q_filt || nf < 1) {
        set_error("filter_q_3pt: bad arguments");
        return LSB_E_ARG;
    }
    double *din, *dout, *tmp;
    LSB_TRY(tensor_in(st, c, 0, q, nf, &din));
    LSB_TRY(st.get(1, c->nphys * nf, &dout));
    LSB_TRY(st.get(2, c->nphys * std::min(nf, (int)kMaxFields), &tmp));
    std::vector<const double*> pi(nf);
    std::vector<double*> po(nf);
    for (int f = 0; f < nf; ++f) {
        pi[f] = din + (size_t)f * c->nphys;
        po[f] = dout + (size_t)f * c->nphys;
    }
    LSB_TRY(ph_filter3(c, pi.data(), po.data(), nf, a, b, interpflag, tmp, c->stream));
    LSB_TRY(ph_zero_pads(c, po.data(), nf, c->stream));
    LSB_TRY(st.out(q_filt, dout, c->nphys * nf));
    return st.sync();
}

extern "C" int lsb_specfilter_exp_xy(lsb_ctx* c, const double* uf, double* uff, double kc, double p, double amp,
                                     int filtery) {
    LSB_ENTER(c);
    double *a, *b;
    LSB_TRY(st.get(0, c->nspec, &a));
    LSB_TRY(st.get(1, c->nspec, &b));
    LSB_TRY(st.in(a, uf, c->nspec));
    const double* in[1] = {a};
    double* out[1] = {b};
    LSB_TRY(sp_specfilter(c, in, out, 1, kc, p, amp, filtery, c->stream));
    LSB_TRY(st.out(uff, b, c->nspec));
    return st.sync();
}

extern "C" int lsb_filterqspec(lsb_ctx* c, double* qf, int nq) {
    LSB_ENTER(c);
    if (!qf || nq < 1 || nq > kMaxFields) {
        set_error("filterqspec: 1..16 fields");
        return LSB_E_ARG;
    }
    double *a, *b;
    LSB_TRY(st.get(0, c->nspec * nq, &a));
    LSB_TRY(st.get(1, c->nspec * nq, &b));
    LSB_TRY(st.in(a, qf, c->nspec * nq));
    const double* in[kMaxFields];
    double* out[kMaxFields];
    for (int f = 0; f < nq; ++f) {
        in[f] = a + (size_t)f * c->nspec;
        out[f] = b + (size_t)f * c->nspec;
    }
    LSB_TRY(sp_specfilter(c, in, out, nq, c->cfg.kc, c->cfg.filt_order, c->cfg.filt_amp, 1, c->stream));
    LSB_TRY(st.out(qf, b, c->nspec * nq));
    return st.sync();
}

// --------------------------------------------------------------- dealias

extern "C" int lsb_setup_dealiasing(lsb_ctx* c, int* nx_lg, int* ny_lg) {
    LSB_ENTER(c);
    LSB_TRY(dealias_setup(c));
    if (nx_lg) *nx_lg = c->nx_lg;
    if (ny_lg) *ny_lg = c->ny_lg;
    return LSB_OK;
}

extern "C" int lsb_dealiasing_2d_fftw(lsb_ctx* c, const double* a, const double* b, double* ab) {
    LSB_ENTER(c);
    LSB_TRY(dealias_setup(c));
    double *da, *db, *dab;
    LSB_TRY(st.get(0, c->nphys, &da));
    LSB_TRY(st.get(1, c->nphys, &db));
    LSB_TRY(st.get(2, c->nphys, &dab));
    LSB_TRY(st.in(da, a, c->nphys));
    LSB_TRY(st.in(db, b, c->nphys));
    LSB_TRY(dealias_product(c, da, db, dab, c->stream));
    LSB_TRY(st.out(ab, dab, c->nphys));
    return st.sync();
}

extern "C" int lsb_dealiasf_2d_fftw(lsb_ctx* c, const double* af, const double* bf, double* abf) {
    LSB_ENTER(c);
    if (!af || !bf || !abf) {
        set_error("dealiasf_2d_fftw: null array");
        return LSB_E_ARG;
    }
    LSB_TRY(dealias_setup(c));
    double *da, *db, *dab;
    LSB_TRY(st.get(0, c->nspec, &da));
    LSB_TRY(st.get(1, c->nspec, &db));
    LSB_TRY(st.get(2, c->nspec, &dab));
    LSB_TRY(st.in(da, af, c->nspec));
    LSB_TRY(st.in(db, bf, c->nspec));
    LSB_TRY(dealias_product_spec(c, da, db, dab, c->stream));
    LSB_TRY(st.out(abf, dab, c->nspec));
    return st.sync();
}

extern "C" int lsb_test_fft(lsb_ctx* c, const double* u, double* avgerror) {
    LSB_ENTER(c);
    if (!u || !avgerror) {
        set_error("test_fft: null argument");
        return LSB_E_ARG;
    }
    double *dq, *dqf;
    LSB_TRY(st.get(0, c->nphys, &dq));
    LSB_TRY(st.get(1, c->nspec, &dqf));
    LSB_TRY(st.in(dq, u, c->nphys));
    {
        const double* in[1] = {dq};
        double* out[1] = {dqf};
        LSB_TRY(horfft_fwd(c, c->hp, in, out, 1, c->nzp, c->stream));
    }
    {
        const double* in[1] = {dqf};
        double* out[1] = {dq};
        LSB_TRY(horfft_inv(c, c->hp, in, out, 1, c->nzp, c->stream));
    }
    std::vector<double> q(c->nphys);
    LSB_TRY(st.out(q.data(), dq, c->nphys));
    LSB_TRY(st.sync());
    // horfft.f90:26-33: the mean absolute round-trip error over the nx physical points
    const double npts = (double)c->nx * c->ny * c->nzp;
    double e = 0.0;
    for (int k = 0; k < c->nzp; ++k)
        for (int j = 0; j < c->ny; ++j)
            for (int i = 0; i < c->nx; ++i) {
                const size_t o = ((size_t)k * c->ny + j) * c->nxpp + i;
                e = e + std::fabs(u[o] - q[o]) / npts;
            }
    *avgerror = e;
    return LSB_OK;
}

extern "C" int lsb_zinteg(lsb_ctx* c, const double* tmp, double* tint, int n1, int n2, int mapping, int per,
                          int rev) {
    LSB_ENTER(c);
    if (!tmp || !tint || n1 < 1 || n2 < 1 || (per != 0 && per != 1) || (rev != 0 && rev != 1)) {
        set_error("zinteg: need arrays (n1, n2, nzp), n1, n2 >= 1, per and rev 0 or 1");
        return LSB_E_ARG;
    }
    const size_t n = (size_t)n1 * n2 * c->nzp;
    double *dq, *dt, *dw = nullptr;
    LSB_TRY(st.get(0, n, &dq));
    LSB_TRY(st.get(1, n, &dt));
    LSB_TRY(st.in(dq, tmp, n));
    std::vector<double> w;
    if (per == 1) {
        // derivatives.f90:25-35: interior weights wts(j), j = 1..nzm, times the change of
        // variable 0.5 xl / gf(k) at point k = j + 1
        const int nz = c->nz, nzm = c->nzm;
        const double pi = 4.0 * std::atan(1.0);
        w.assign(c->nzp, 0.0);
        for (int j = 1; j <= nzm; ++j) {
            const double tj = pi * (double)j / (double)nz;
            double s = 0.0;
            for (int m = 1; m <= nzm; ++m) s = s + std::sin((double)m * tj) * (1.0 - std::cos(pi * (double)m)) / (double)m;
            const double wts = 2.00 / (double)nz * std::sin(tj) * s;
            w[j] = wts * 0.5 * c->xl / c->gf[j];
        }
        LSB_TRY(st.get(2, c->nzp, &dw));
        LSB_TRY(st.in(dw, w.data(), c->nzp));
    }
    LSB_TRY(zinteg_dev(c, dq, dt, n1, n2, mapping, per, rev, dw, c->stream));
    LSB_TRY(st.out(tint, dt, n));
    return st.sync();
}

// ------------------------------------------------------------ step API
extern "C" int lsb_nlterm_skewsym(lsb_ctx* c, const double* uif, double* nltermif) {
    LSB_ENTER(c);
    if (c->cfg.dealiasflag == 1) {
        set_error("nlterm_skewsym: the 3/2-rule products are only available through dealiasing_2d_fftw");
        return LSB_E_UNSUPPORTED;
    }
    double *u, *n;
    LSB_TRY(st.get(0, c->nspec * 3, &u));
    LSB_TRY(st.get(1, c->nspec * 3, &n));
    LSB_TRY(st.in(u, uif, c->nspec * 3));
    const double* ui[3] = {u, u + c->nspec, u + 2 * c->nspec};
    double* no[3] = {n, n + c->nspec, n + 2 * c->nspec};
    LSB_TRY(nlterm_dev(c, ui, no, c->stream));
    LSB_TRY(st.out(nltermif, n, c->nspec * 3));
    return st.sync();
}

extern "C" int lsb_state_set(lsb_ctx* c, const double* uf, const double* vf, const double* wf, const double* pf,
                             const double* pnf, const double* unif, const double* unmif, const double* nlnif,
                             const double* nlnmif, int istep, double t) {
    LSB_ENTER(c);
    if (c->cfg.dealiasflag == 1) {
        set_error("timestep: dealiasflag 1 runs through tier A (dealiasing_2d_fftw) only");
        return LSB_E_UNSUPPORTED;
    }
    LSB_TRY(step_alloc(c));
    const size_t n = c->nspec;
    auto put = [&](int which, const double* h) -> int {
        if (step_hist32(c, which)) {  // FP32 N history: convert on the device
            double* tmp = nullptr;
            if (h) {
                LSB_TRY(st.get(0, n, &tmp));
                LSB_TRY(st.in(tmp, h, n));
            }
            LSB_TRY(step_hist_set(c, which, tmp, c->stream));
            return st.sync();
        }
        double* d = step_field(c, which);
        if (h)
            LSB_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        else
            LSB_HIP(hipMemsetAsync(d, 0, n * sizeof(double), c->stream));
        return LSB_OK;
    };
    LSB_TRY(put(0, uf));
    LSB_TRY(put(1, vf));
    LSB_TRY(put(2, wf));
    for (int l = 0; l < 3; ++l) {
        LSB_TRY(put(3 + l, unif ? unif + l * n : nullptr));
        LSB_TRY(put(6 + l, unmif ? unmif + l * n : nullptr));
        LSB_TRY(put(9 + l, nlnif ? nlnif + l * n : nullptr));
        LSB_TRY(put(12 + l, nlnmif ? nlnmif + l * n : nullptr));
    }
    LSB_TRY(put(15, pf));
    LSB_TRY(put(16, pnf));
    LSB_TRY(tavg_zero(c));  // a new run starts its averages at zero (init.f90:112-119)
    c->istep = istep;
    c->isum = istep;
    c->coef_valid = false;
    c->t = t;
    c->dt = c->dt2 = c->dt3 = c->cfg.dt;
    LSB_TRY(step_prepare(c));
    c->state_ready = true;
    return st.sync();
}

extern "C" int lsb_state_get(lsb_ctx* c, double* uf, double* vf, double* wf, double* pf, double* pnf, double* unif,
                             double* unmif, double* nlnif, double* nlnmif, int* istep, double* t) {
    LSB_ENTER(c);
    if (!c->state_ready) {
        set_error("state_get: no state");
        return LSB_E_STATE;
    }
    const size_t n = c->nspec;
    auto get = [&](int which, double* h) -> int {
        if (!h) return LSB_OK;
        if (step_hist32(c, which)) {  // FP32 N history: widen on the device
            double* tmp = nullptr;
            LSB_TRY(st.get(0, n, &tmp));
            LSB_TRY(step_hist_get(c, which, tmp, c->stream));
            LSB_TRY(st.out(h, tmp, n));
            return st.sync();
        }
        LSB_HIP(hipMemcpyAsync(h, step_field(c, which), n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        return LSB_OK;
    };
    LSB_TRY(get(0, uf));
    LSB_TRY(get(1, vf));
    LSB_TRY(get(2, wf));
    for (int l = 0; l < 3; ++l) {
        LSB_TRY(get(3 + l, unif ? unif + l * n : nullptr));
        LSB_TRY(get(6 + l, unmif ? unmif + l * n : nullptr));
        LSB_TRY(get(9 + l, nlnif ? nlnif + l * n : nullptr));
        LSB_TRY(get(12 + l, nlnmif ? nlnmif + l * n : nullptr));
    }
    LSB_TRY(get(15, pf));
    LSB_TRY(get(16, pnf));
    if (istep) *istep = c->isum;
    if (t) *t = c->t;
    return st.sync();
}

extern "C" int lsb_timestep(lsb_ctx* c, int nsteps) {
    LSB_TRY(check_ctx(c));
    LSB_TRY(set_device(c));
    LSB_TRY(step_run(c, nsteps, true));
    return LSB_OK;
}

extern "C" int lsb_navierstokes_update(lsb_ctx* c) {
    LSB_TRY(check_ctx(c));
    LSB_TRY(set_device(c));
    return step_run(c, 1, false);
}

extern "C" int lsb_state_device_ptrs(lsb_ctx* c, void** uf, void** vf, void** wf, void** pf) {
    LSB_TRY(check_ctx(c));
    if (!c->state_ready) return LSB_E_STATE;
    if (uf) *uf = step_field(c, 0);
    if (vf) *vf = step_field(c, 1);
    if (wf) *wf = step_field(c, 2);
    if (pf) *pf = step_field(c, 15);
    return LSB_OK;
}

extern "C" void* lsb_stream(lsb_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int lsb_profile_begin_classes(lsb_ctx* c, unsigned class_mask) {
    LSB_TRY(check_ctx(c));
    for (auto& ev : c->prof) {
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    c->prof.clear();
    c->profiling = true;
    c->prof_mask = class_mask;
    return LSB_OK;
}

extern "C" int lsb_profile_begin(lsb_ctx* c) { return lsb_profile_begin_classes(c, ~0u); }

extern "C" int lsb_synchronize(lsb_ctx* c) {
    LSB_TRY(check_ctx(c));
    LSB_HIP(hipStreamSynchronize(c->stream));
    return LSB_OK;
}

// Launch-shape autotuner: times every variant of each tunable kernel family on this
// context's grid (scratch fields allocated for the duration of the call) and keeps the
// fastest in the context; later launches of the family use it.
extern "C" int lsb_autotune(lsb_ctx* c) {
    LSB_TRY(check_ctx(c));
    LSB_HIP(hipSetDevice(c->device));
    constexpr int nf = 8;
    void* work = nullptr;
    // the filter's 2 nf fields; the y passes' right-hand-side variant needs 21 (of 24)
    const size_t bytes = 24 * (size_t)c->nphys * sizeof(double);
    LSB_TRY(dev_alloc(&work, bytes));
    int rc = LSB_OK;
    if (hipMemsetAsync(work, 0, bytes, c->stream) != hipSuccess) rc = LSB_E_HIP;
    if (rc == LSB_OK) rc = ph_tune_filt3(c, work, nf, c->stream);
    if (rc == LSB_OK) rc = fft_tune_y(c, work, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(work);
    return rc;
}

extern "C" int lsb_autotune_get(const lsb_ctx* c, int* variant, float* ms, int n) {
    LSB_TRY(check_ctx(c));
    for (int i = 0; i < n && i < lsb::TUNE_SLOTS; ++i) {
        if (variant) variant[i] = c->tune[i];
        if (ms) ms[i] = c->tune_ms[i];
    }
    return LSB_OK;
}

extern "C" int lsb_autotune_set(lsb_ctx* c, const int* variant, int n) {
    LSB_TRY(check_ctx(c));
    for (int i = 0; i < n && i < lsb::TUNE_SLOTS; ++i) c->tune[i] = variant[i];
    return LSB_OK;
}

extern "C" int lsb_profile_end(lsb_ctx* c, double* ms, int* count, double* flops, double* bytes) {
    LSB_TRY(check_ctx(c));
    LSB_HIP(hipStreamSynchronize(c->stream));
    double t_ms[kProfClasses] = {}, fl[kProfClasses] = {}, by[kProfClasses] = {};
    int cnt[kProfClasses] = {};
    for (auto& ev : c->prof) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, ev.a, ev.b) == hipSuccess) {
            t_ms[ev.cls] += t;
            cnt[ev.cls] += 1;
            fl[ev.cls] += ev.flops;
            by[ev.cls] += ev.bytes;
        }
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    c->prof.clear();
    c->profiling = false;
    for (int i = 0; i < kProfClasses; ++i) {
        if (ms) ms[i] = t_ms[i];
        if (count) count[i] = cnt[i];
        if (flops) flops[i] = fl[i];
        if (bytes) bytes[i] = by[i];
    }
    return LSB_OK;
}

// ======================================= snapshots, restart files, diagnostics
namespace {

constexpr int kSnapHead = 7;      // mpi_io.f90:408
constexpr int kRestartHead = 9;   // mpi_io.f90:44
constexpr int kRestartFields = 20;

// mpi_io.f90:400-404: an existing file of that name becomes prev_<name> (same folder).
int write_record(const char* path, const float* head, size_t nh, const float* body, size_t nb) {
    std::string p(path);
    FILE* probe = std::fopen(path, "rb");
    if (probe) {
        std::fclose(probe);
        const size_t cut = p.find_last_of('/');
        const std::string prev = (cut == std::string::npos) ? "prev_" + p : p.substr(0, cut + 1) + "prev_" +
                                                                                  p.substr(cut + 1);
        if (std::rename(path, prev.c_str()) != 0) {
            set_error("snapshot: cannot rename " + p + " to " + prev);
            return LSB_E_ARG;
        }
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        set_error("cannot open " + p + " for writing");
        return LSB_E_ARG;
    }
    const bool ok = std::fwrite(head, sizeof(float), nh, f) == nh && std::fwrite(body, sizeof(float), nb, f) == nb;
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) {
        set_error("short write on " + p);
        return LSB_E_ARG;
    }
    return LSB_OK;
}

int read_record(const char* path, std::vector<float>* rec, size_t expect) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_error(std::string("cannot open ") + path);
        return LSB_E_ARG;
    }
    rec->resize(expect);
    const size_t got = std::fread(rec->data(), sizeof(float), expect, f);
    std::fclose(f);
    if (got != expect) {
        set_error(std::string("short record in ") + path);
        return LSB_E_ARG;
    }
    return LSB_OK;
}

void snap_header(const lsb_ctx* c, int isum, double t, double dt, float* h) {
    const double v[kSnapHead] = {(double)isum, t, dt, (double)c->nx, (double)c->ny, (double)c->nzp, (double)c->ny};
    for (int q = 0; q < kSnapHead; ++q) h[q] = (float)v[q];
}

// Inverse transforms of the present spectral fields (device) into physical stage slots
// 1.. and the (x, z, y) FP32 record body in stage slot 0.
int snapshot_body(lsb_ctx* c, Stage& st, const double* const* spec, int name, const double* tbar_dev,
                  float** body) {
    double* dp;
    LSB_TRY(st.get(1, c->nphys * 5, &dp));
    const double* si[5];
    double* so[5];
    const double* phys[5];
    int n = 0;
    for (int l = 0; l < 5; ++l) {
        phys[l] = nullptr;
        if (!spec[l]) continue;
        si[n] = spec[l];
        so[n] = dp + (size_t)l * c->nphys;
        phys[l] = so[n];
        ++n;
    }
    LSB_TRY(horfft_inv(c, c->hp, si, so, n, c->nzp, c->stream));
    double* b;
    LSB_TRY(st.get(0, (c->nspec * 5 + 1) / 2, &b));
    *body = (float*)b;
    return snap_pack(c, phys, name != 2 ? c->d_paral : nullptr, name != 2 ? tbar_dev : nullptr, *body, c->stream);
}

// navierstokes.f90:452-467: the CFL limits and the 4/5 (or 5/4) change of dt.
bool maxcalc_decide(const double* s, double dtmax, double* dt, double* dt2) {
    const double cflx = s[8], cfly = s[9], cflz = s[10];
    const bool exceed = cflz > 0.45 || cflx > 0.25 || cfly > 0.25;
    const bool below = cflz < 0.25 && cflx < 0.10 && cfly < 0.12 && *dt < dtmax;
    if (!(exceed || below)) return false;
    const double cdt = exceed ? 1.25 : 1.0 / 1.25;
    *dt2 = *dt;
    *dt = *dt / cdt;
    return true;
}

// maxcalc on three (four) spectral fields on the device: inverse transforms into stage
// slots 5.., then the reduction.  stats[8..10] become the CFL numbers.
int maxcalc_fields(lsb_ctx* c, Stage& st, const double* u, const double* v, const double* w, const double* temp,
                   const double* tbar, double dt, double t, double fsize, double* stats) {
    double* dp;
    const int nf = temp ? 4 : 3;
    LSB_TRY(st.get(5, c->nphys * nf, &dp));
    const double* si[4] = {u, v, w, temp};
    double* so[4] = {dp, dp + c->nphys, dp + 2 * c->nphys, dp + 3 * c->nphys};
    LSB_TRY(horfft_inv(c, c->hp, si, so, nf, c->nzp, c->stream));
    int findex = c->nx;
    if (t > c->cfg.istartbound) {  // navierstokes.f90:442-447
        findex = (int)((1.0 - 0.8 * fsize) * c->nx);
        findex = std::max(0, std::min(c->nx, findex));
    }
    LSB_TRY(maxcalc_reduce(c, so[0], so[1], so[2], temp ? so[3] : nullptr, tbar, dt, findex, stats, c->stream));
    stats[8] = (stats[8] * dt) / c->dx;  // maxval(abs(u))*dt/dx (:439)
    stats[9] = (stats[9] * dt) / c->dy;
    return LSB_OK;
}

// cf(x) from the spanwise mean of u + ubar at z rows nz and nzp (1-based), and the
// separation / reattachment points (run/compare_results.m:63,76-79).
int wall_shear_field(lsb_ctx* c, Stage& st, const double* uf_dev, double* cf, double* x_sep, double* x_reatt) {
    double* ph;
    LSB_TRY(st.get(9, c->plane_p * 2, &ph));
    const double* si[1] = {uf_dev + (size_t)(c->nz - 1) * c->plane_s};
    double* so[1] = {ph};
    LSB_TRY(horfft_inv(c, c->hp, si, so, 1, 2, c->stream));
    std::vector<double> mean(2 * (size_t)c->nx);
    LSB_TRY(wall_mean(c, ph, mean.data(), c->stream));
    const int nz = c->nz, nx = c->nx;
    const double dz = c->zpts[nz - 1] - c->zpts[nz];
    const double den = 0.5 * c->cfg.u0 * c->cfg.u0;
    std::vector<double> cfv(nx);
    for (int i = 0; i < nx; ++i) {
        const double a = mean[i] + c->ubar[nz - 1], b = mean[nx + i] + c->ubar[nz];
        cfv[i] = ((c->cfg.xnu * (a - b)) / dz) / den;
    }
    if (cf) std::memcpy(cf, cfv.data(), nx * sizeof(double));
    double sep = NAN, reatt = NAN;
    int i0 = -1;
    for (int i = 1; i < nx; ++i)
        if (cfv[i - 1] >= 0.0 && cfv[i] < 0.0) {
            sep = c->xpts[i - 1] + (c->xpts[i] - c->xpts[i - 1]) * cfv[i - 1] / (cfv[i - 1] - cfv[i]);
            i0 = i;
            break;
        }
    if (i0 > 0)
        for (int i = i0 + 1; i < nx; ++i)
            if (cfv[i - 1] < 0.0 && cfv[i] >= 0.0) {
                reatt = c->xpts[i - 1] + (c->xpts[i] - c->xpts[i - 1]) * cfv[i - 1] / (cfv[i - 1] - cfv[i]);
                break;
            }
    if (x_sep) *x_sep = sep;
    if (x_reatt) *x_reatt = reatt;
    return LSB_OK;
}

// FP64 device view of state field `which` (a converted copy in stage slot 8 for the
// FP32 history of the mixed form).
int state_field_f64(lsb_ctx* c, Stage& st, int which, const double** p) {
    if (!step_hist32(c, which)) {
        *p = step_field(c, which);
        return LSB_OK;
    }
    double* tmp;
    LSB_TRY(st.get(8, c->nspec, &tmp));
    LSB_TRY(step_hist_get(c, which, tmp, c->stream));
    *p = tmp;
    return LSB_OK;
}

int need_state(lsb_ctx* c, const char* who) {
    if (!c->state_ready) {
        set_error(std::string(who) + ": no state (lsb_state_set or lsb_read_restart first)");
        return LSB_E_STATE;
    }
    return LSB_OK;
}

}  // namespace

extern "C" int lsb_snapshot_record(lsb_ctx* c, const double* uf, const double* vf, const double* wf,
                                   const double* tempf, const double* pf, const double* tbar, int name, int isum,
                                   double t, double dt, float* rec) {
    LSB_ENTER(c);
    if (!uf || !vf || !wf || !pf || !rec) {
        set_error("snapshot_record: null array");
        return LSB_E_ARG;
    }
    const double* hs[5] = {uf, vf, wf, tempf, pf};
    double *ds, *dtb = nullptr;
    LSB_TRY(st.get(2, c->nspec * 5, &ds));
    const double* spec[5];
    for (int l = 0; l < 5; ++l) {
        spec[l] = nullptr;
        if (!hs[l]) continue;
        double* d = ds + (size_t)l * c->nspec;
        LSB_TRY(st.in(d, hs[l], c->nspec));
        spec[l] = d;
    }
    if (tbar) {
        LSB_TRY(st.get(3, c->nzp, &dtb));
        LSB_TRY(st.in(dtb, tbar, c->nzp));
    }
    float* body;
    LSB_TRY(snapshot_body(c, st, spec, name, dtb, &body));
    snap_header(c, isum, t, dt, rec);
    LSB_HIP(hipMemcpyAsync(rec + kSnapHead, body, c->nspec * 5 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return st.sync();
}

extern "C" int lsb_state_snapshot(lsb_ctx* c, int name, const char* path, float* rec) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "state_snapshot"));
    const double* spec[5] = {step_field(c, 0), step_field(c, 1), step_field(c, 2),
                             temp_on(c) ? temp_field(c, 0) : nullptr, step_field(c, 15)};
    float* body;
    LSB_TRY(snapshot_body(c, st, spec, name, temp_tbar_dev(c), &body));
    const size_t nb = c->nspec * 5;
    std::vector<float> own;
    float* hb = rec ? rec + kSnapHead : nullptr;
    if (!hb) {
        own.resize(nb);
        hb = own.data();
    }
    float head[kSnapHead];
    snap_header(c, c->isum, c->t, c->dt, head);
    LSB_HIP(hipMemcpyAsync(hb, body, nb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    LSB_TRY(st.sync());
    if (rec) std::memcpy(rec, head, sizeof(head));
    if (path) LSB_TRY(write_record(path, head, kSnapHead, hb, nb));
    return LSB_OK;
}

extern "C" int lsb_read_snapshot(lsb_ctx* c, const char* path, double* q, int* isum, double* t, double* dt) {
    LSB_TRY(check_ctx(c));
    if (!path) {
        set_error("read_snapshot: null path");
        return LSB_E_ARG;
    }
    const size_t res = c->nspec;  // nx*ny*nzp
    std::vector<float> rec;
    LSB_TRY(read_record(path, &rec, kSnapHead));
    const int nx_in = (int)std::lround(rec[3]), ny_in = (int)std::lround(rec[4]), nzp_in = (int)std::lround(rec[5]);
    if (nx_in != c->nx || ny_in != c->ny || nzp_in != c->nzp) {  // mpi_io.f90:575-583
        set_error("read_snapshot: dimension mismatch (file " + std::to_string(nx_in) + " " + std::to_string(ny_in) +
                  " " + std::to_string(nzp_in) + ")");
        return LSB_E_ARG;
    }
    LSB_TRY(read_record(path, &rec, kSnapHead + 5 * res));
    if (isum) *isum = (int)std::lround(rec[0]);
    if (t) *t = rec[1];
    if (dt) *dt = rec[2];
    if (q) {  // q(i, j, k, l) = buf(ii, l), ii over x, z, y (mpi_io.f90:623-636)
        const int nx = c->nx, ny = c->ny, nzp = c->nzp;
        for (int l = 0; l < 5; ++l) {
            const float* b = rec.data() + kSnapHead + (size_t)l * res;
            double* o = q + (size_t)l * res;
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nzp; ++k)
                    for (int i = 0; i < nx; ++i)
                        o[((size_t)k * ny + j) * nx + i] = b[((size_t)j * nzp + k) * nx + i];
        }
    }
    return LSB_OK;
}

extern "C" int lsb_write_restart(lsb_ctx* c, const char* path, float* rec) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "write_restart"));
    double* dbody;
    LSB_TRY(st.get(0, c->nspec * kRestartFields / 2, &dbody));
    float* body = (float*)dbody;
    // file slot -> state field; -2..-6: tempf, tnf, tnmf, nltnf, nltnmf (zero without the
    // temperature path); slots 14-16 hold nlnmif
    const int which[kRestartFields] = {0, 1, 2, -2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, -3, -4, -5, -6};
    for (int f = 0; f < kRestartFields; ++f) {
        const double* src[1] = {nullptr};
        if (which[f] >= 0) LSB_TRY(state_field_f64(c, st, which[f], &src[0]));
        else if (temp_on(c)) src[0] = temp_field(c, -2 - which[f]);
        LSB_TRY(restart_pack(c, src, 1, body + (size_t)f * c->nspec, c->stream));
    }
    const size_t nb = c->nspec * kRestartFields;
    std::vector<float> own;
    float* hb = rec ? rec + kRestartHead : nullptr;
    if (!hb) {
        own.resize(nb);
        hb = own.data();
    }
    const double hv[kRestartHead] = {(double)c->isum, c->t, c->dt, c->dt2, c->dt3,
                                     (double)c->nx, (double)c->ny, (double)c->nzp, (double)c->nx};
    float head[kRestartHead];
    for (int q = 0; q < kRestartHead; ++q) head[q] = (float)hv[q];
    LSB_HIP(hipMemcpyAsync(hb, body, nb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    LSB_TRY(st.sync());
    if (rec) std::memcpy(rec, head, sizeof(head));
    if (path) LSB_TRY(write_record(path, head, kRestartHead, hb, nb));
    return LSB_OK;
}

extern "C" int lsb_read_restart(lsb_ctx* c, const char* path, const float* rec) {
    LSB_ENTER(c);
    if (!path && !rec) {
        set_error("read_restart: no file and no record");
        return LSB_E_ARG;
    }
    if (c->cfg.dealiasflag == 1) {
        set_error("read_restart: dealiasflag 1 runs through tier A (dealiasing_2d_fftw) only");
        return LSB_E_UNSUPPORTED;
    }
    const size_t nb = c->nspec * kRestartFields;
    std::vector<float> own;
    const float* r = rec;
    if (!r) {
        LSB_TRY(read_record(path, &own, kRestartHead));
        r = own.data();
    }
    const int nx_in = (int)std::lround(r[5]), ny_in = (int)std::lround(r[6]), nzp_in = (int)std::lround(r[7]);
    if (nx_in != c->nx || ny_in != c->ny || nzp_in != c->nzp) {  // mpi_io.f90:243-252
        set_error("read_restart: dimension mismatch (file " + std::to_string(nx_in) + " " + std::to_string(ny_in) +
                  " " + std::to_string(nzp_in) + ")");
        return LSB_E_ARG;
    }
    if (!rec) {
        LSB_TRY(read_record(path, &own, kRestartHead + nb));
        r = own.data();
    }
    // temperature slots: loaded into the temperature path (FP64 form), which a nonzero
    // temperature switches on
    const int tslots[5] = {3, 16, 17, 18, 19};
    bool hot = false;
    for (int f : tslots) {
        const float* b = r + kRestartHead + (size_t)f * c->nspec;
        for (size_t q = 0; q < c->nspec && !hot; ++q) hot = b[q] != 0.0f;
    }
    if (hot) LSB_TRY(temp_enable(c));
    const double hv[5] = {r[0], r[1], r[2], r[3], r[4]};
    LSB_TRY(step_alloc(c));
    const bool fresh = !c->state_ready;
    double* dbody;
    LSB_TRY(st.get(0, nb / 2, &dbody));
    float* body = (float*)dbody;
    LSB_HIP(hipMemcpyAsync(body, r + kRestartHead, nb * sizeof(float), hipMemcpyHostToDevice, c->stream));
    const int which[kRestartFields] = {0, 1, 2, -1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, -1, -1, -1, -1};
    double* tmp;
    LSB_TRY(st.get(8, c->nspec, &tmp));
    for (int f = 0; f < kRestartFields; ++f) {
        const int w = which[f];
        if (w < 0) continue;
        const bool h32 = step_hist32(c, w);
        double* dst[1] = {h32 ? tmp : step_field(c, w)};
        LSB_TRY(restart_unpack(c, body + (size_t)f * c->nspec, dst, 1, c->stream));
        if (h32) LSB_TRY(step_hist_set(c, w, tmp, c->stream));
    }
    if (temp_on(c))
        for (int q = 0; q < 5; ++q) {
            double* dst[1] = {temp_field(c, q)};
            LSB_TRY(restart_unpack(c, body + (size_t)tslots[q] * c->nspec, dst, 1, c->stream));
        }
    if (fresh) {  // pf, pnf are not in the file; the reference keeps them as they were
        LSB_HIP(hipMemsetAsync(step_field(c, 15), 0, c->nspec * sizeof(double), c->stream));
        LSB_HIP(hipMemsetAsync(step_field(c, 16), 0, c->nspec * sizeof(double), c->stream));
    }
    // isum = nint(prebuf(1)) (mpi_io.f90:231); the restarted run's own step counter
    // starts again at 1 (main.f90:61), so the start-up rule of nstart applies
    LSB_TRY(tavg_zero(c));  // the restarted run starts its averages at zero (init.f90:112-119)
    c->isum = (int)std::lround(hv[0]);
    c->istep = 0;
    c->coef_valid = false;
    c->t = hv[1];
    c->dt = hv[2];
    c->dt2 = hv[3];
    c->dt3 = hv[4];
    LSB_TRY(step_prepare(c));
    c->state_ready = true;
    return st.sync();
}

extern "C" int lsb_maxcalc(lsb_ctx* c, const double* uf, const double* vf, const double* wf, const double* tempf,
                           const double* tbar, double dtmax, double t, double fsize, double* dt, double* dt2,
                           double* dt3, double* stats, int* change_dt) {
    LSB_ENTER(c);
    if (!uf || !vf || !wf || !dt || !dt2 || !dt3) {
        set_error("maxcalc: null argument");
        return LSB_E_ARG;
    }
    double* ds;
    const int nf = tempf ? 4 : 3;
    LSB_TRY(st.get(4, c->nspec * nf, &ds));
    const double* hs[4] = {uf, vf, wf, tempf};
    for (int l = 0; l < nf; ++l) LSB_TRY(st.in(ds + (size_t)l * c->nspec, hs[l], c->nspec));
    double s[11];
    LSB_TRY(maxcalc_fields(c, st, ds, ds + c->nspec, ds + 2 * c->nspec, tempf ? ds + 3 * c->nspec : nullptr, tbar,
                           *dt, t, fsize, s));
    const bool ch = maxcalc_decide(s, dtmax, dt, dt2);
    (void)dt3;  // dt3 is passed through unchanged (navierstokes.f90:462-464)
    if (stats) std::memcpy(stats, s, sizeof(s));
    if (change_dt) *change_dt = ch ? 1 : 0;
    return LSB_OK;
}

extern "C" int lsb_state_maxcalc(lsb_ctx* c, double dtmax, double fsize, double* stats, int* change_dt) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "state_maxcalc"));
    double s[11];
    LSB_TRY(maxcalc_fields(c, st, step_field(c, 0), step_field(c, 1), step_field(c, 2),
                           temp_on(c) ? temp_field(c, 0) : nullptr, temp_tbar_host(c), c->dt, c->t, fsize, s));
    const bool ch = maxcalc_decide(s, dtmax, &c->dt, &c->dt2);
    if (stats) std::memcpy(stats, s, sizeof(s));
    if (change_dt) *change_dt = ch ? 1 : 0;
    return LSB_OK;
}

extern "C" int lsb_nancheck(lsb_ctx* c, const double* uf, const double* vf, const double* wf, const double* tempf,
                            int* has_nan) {
    LSB_ENTER(c);
    if (!has_nan) return LSB_E_ARG;
    const double* hs[4] = {uf, vf, wf, tempf};
    double* ds;
    LSB_TRY(st.get(4, c->nspec * 4, &ds));
    const double* f[4];
    int n = 0;
    for (int l = 0; l < 4; ++l) {
        if (!hs[l]) continue;
        double* d = ds + (size_t)n * c->nspec;
        LSB_TRY(st.in(d, hs[l], c->nspec));
        f[n++] = d;
    }
    bool found = false;
    if (n) LSB_TRY(nan_scan(c, f, n, c->nspec, &found, c->stream));
    *has_nan = found ? 1 : 0;
    return LSB_OK;
}

extern "C" int lsb_state_nancheck(lsb_ctx* c, int* has_nan) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "state_nancheck"));
    if (!has_nan) return LSB_E_ARG;
    const double* f[3] = {step_field(c, 0), step_field(c, 1), step_field(c, 2)};
    bool found = false;
    LSB_TRY(nan_scan(c, f, 3, c->nspec, &found, c->stream));
    *has_nan = found ? 1 : 0;
    return LSB_OK;
}

extern "C" int lsb_wall_shear(lsb_ctx* c, const double* uf, double* cf, double* x_sep, double* x_reatt) {
    LSB_ENTER(c);
    if (!uf) return LSB_E_ARG;
    double* d;
    LSB_TRY(st.get(4, c->nspec, &d));
    LSB_TRY(st.in(d, uf, c->nspec));
    return wall_shear_field(c, st, d, cf, x_sep, x_reatt);
}

extern "C" int lsb_state_wall_shear(lsb_ctx* c, double* cf, double* x_sep, double* x_reatt) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "state_wall_shear"));
    return wall_shear_field(c, st, step_field(c, 0), cf, x_sep, x_reatt);
}

extern "C" int lsb_set_adaptive_dt(lsb_ctx* c, int dtadapt, int iomod, double fsize) {
    LSB_TRY(check_ctx(c));
    if ((dtadapt != 0 && dtadapt != 1) || iomod < 1 || !(fsize >= 0.0 && fsize <= 1.25)) {
        set_error("set_adaptive_dt: dtadapt 0/1, iomod >= 1, 0 <= fsize <= 1.25");
        return LSB_E_ARG;
    }
    c->dtadapt = dtadapt;
    c->iomod = iomod;
    c->fsize = fsize;
    return LSB_OK;
}

extern "C" int lsb_set_time_average(lsb_ctx* c, int otfavg, int istartavg, int iomod) {
    LSB_ENTER(c);
    if ((otfavg != 0 && otfavg != 1) || iomod < 1) {
        set_error("set_time_average: otfavg 0/1, iomod >= 1");
        return LSB_E_ARG;
    }
    c->otfavg = otfavg;
    c->istartavg = istartavg;
    c->avg_iomod = iomod;
    if (otfavg == 1) {  // init.f90:112-119: allocated and zeroed
        if (!c->d_tavg) LSB_TRY(dev_alloc((void**)&c->d_tavg, 5 * c->nspec * sizeof(double)));
        LSB_TRY(tavg_zero(c));
    } else if (c->d_tavg) {
        LSB_HIP(hipStreamSynchronize(c->stream));
        LSB_HIP(hipFree(c->d_tavg));
        c->d_tavg = nullptr;
    }
    return st.sync();
}

extern "C" int lsb_state_get_time_average(lsb_ctx* c, double* uf, double* vf, double* wf, double* tempf, double* pf) {
    LSB_ENTER(c);
    if (!c->d_tavg) {
        set_error("state_get_time_average: time averaging is off (lsb_set_time_average)");
        return LSB_E_STATE;
    }
    double* hs[5] = {uf, vf, wf, tempf, pf};
    for (int l = 0; l < 5; ++l)
        if (hs[l])
            LSB_HIP(hipMemcpyAsync(hs[l], c->d_tavg + (size_t)l * c->nspec, c->nspec * sizeof(double),
                                   hipMemcpyDeviceToHost, c->stream));
    return st.sync();
}

extern "C" int lsb_state_snapshot_average(lsb_ctx* c, const char* path, float* rec) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "state_snapshot_average"));
    if (!c->d_tavg) {
        set_error("state_snapshot_average: time averaging is off (lsb_set_time_average)");
        return LSB_E_STATE;
    }
    const double* spec[5];
    for (int l = 0; l < 5; ++l) spec[l] = c->d_tavg + (size_t)l * c->nspec;
    float* body;
    LSB_TRY(snapshot_body(c, st, spec, 1, temp_tbar_dev(c), &body));
    const size_t nb = c->nspec * 5;
    std::vector<float> own;
    float* hb = rec ? rec + kSnapHead : nullptr;
    if (!hb) {
        own.resize(nb);
        hb = own.data();
    }
    float head[kSnapHead];
    snap_header(c, c->isum, c->t, c->dt, head);
    LSB_HIP(hipMemcpyAsync(hb, body, nb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    LSB_TRY(tavg_zero(c));  // main.f90:104-106, after the averaged snapshot
    LSB_TRY(st.sync());
    if (rec) std::memcpy(rec, head, sizeof(head));
    if (path) LSB_TRY(write_record(path, head, kSnapHead, hb, nb));
    return LSB_OK;
}

namespace lsb {
// main.f90:69-76, after timestep() (whose time update has counted the step in its):
// q_tavg = q_tavg + q/iomod for uf, vf, wf, tempf, pf once its > istartavg.  tempf is
// zero while the temperature path is off, so its average is left alone then.
int state_time_average(lsb_ctx* c) {
    if (c->otfavg != 1 || !c->d_tavg || c->isum <= c->istartavg) return LSB_OK;
    const double* f[5];
    double* a[5];
    int n = 0;
    const int which[5] = {0, 1, 2, -1, 15};
    for (int l = 0; l < 5; ++l) {
        if (which[l] < 0 && !temp_on(c)) continue;
        f[n] = which[l] < 0 ? temp_field(c, 0) : step_field(c, which[l]);
        a[n] = c->d_tavg + (size_t)l * c->nspec;
        ++n;
    }
    return tavg_accum(c, f, a, n, c->avg_iomod, c->stream);
}

// navierstokes.f90:74-87, after the time update of step istep: every 5th (or iomod-th)
// step a NaN check (abort) and maxcalc(..., dtmax = 1.0, dt, dt2, dt3).
int state_adapt_dt(lsb_ctx* c) {
    if (c->dtadapt != 1 || (c->istep % c->iomod != 0 && c->istep % 5 != 0)) return LSB_OK;
    int bad = 0;
    LSB_TRY(lsb_state_nancheck(c, &bad));
    if (bad) {
        set_error("NaN detected at t=" + std::to_string(c->t) + " iteration #: " + std::to_string(c->isum));
        return LSB_E_NUMERIC;
    }
    return lsb_state_maxcalc(c, 1.0, c->fsize, nullptr, nullptr);
}
}  // namespace lsb

extern "C" int lsb_sponge_lambda(int nx, const double* xpts, double xlen, double u0, double fsize, double lfsize,
                                 double sfsize, double* lam) {
    if (nx < 4 || !xpts || !lam || !(fsize > 0.0) || !(lfsize > 0.0) || !(sfsize > 0.0) ||
        !(sfsize + lfsize < 1.0)) {
        set_error("sponge_lambda: need nx >= 4, fsize, lfsize, sfsize > 0 and lfsize + sfsize < 1");
        return LSB_E_ARG;
    }
    sponge_lambda(nx, xpts, xlen, u0, fsize, lfsize, sfsize, lam);
    return LSB_OK;
}

extern "C" int lsb_set_sponge(lsb_ctx* c, double fsize, double lfsize, double sfsize) {
    LSB_ENTER(c);
    std::vector<double> lam(c->nxpp, 0.0);
    LSB_TRY(lsb_sponge_lambda(c->nx, c->xpts.data(), c->cfg.xlen, c->cfg.u0, fsize, lfsize, sfsize, lam.data()));
    c->sponge[0] = fsize;
    c->sponge[1] = lfsize;
    c->sponge[2] = sfsize;
    c->fsize = fsize;
    if (c->cfg.forcing == 1) {
        if (!c->d_lam) LSB_TRY(dev_alloc((void**)&c->d_lam, (size_t)c->nxpp * sizeof(double)));
        LSB_HIP(hipMemcpyAsync(c->d_lam, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    return st.sync();
}

// ================================================ TNS / mode-cut filtering (filters.f90)
namespace {

constexpr double kModecut[6] = {6., -15., 20., -15., 6., -1.};  // filters.f90:637

// filters.f90:617-654 modecutFC: qb = sum_l coef_l F^l q on nf physical device fields, F the
// 3-pt filter (a, b) with interpolating z weights.  A, B: nf-field ping-pong buffers, tmp: the
// filter's own scratch.
int modecut(lsb_ctx* c, const double* const* q, double* const* qb, int nf, double a, double b, double* A, double* B,
            double* tmp) {
    std::vector<const double*> src(q, q + nf);
    std::vector<double*> dst(nf);
    for (int l = 0; l < 6; ++l) {
        double* base = (l % 2 == 0) ? A : B;
        for (int f = 0; f < nf; ++f) dst[f] = base + (size_t)f * c->nphys;
        LSB_TRY(ph_filter3(c, src.data(), dst.data(), nf, a, b, 1, tmp, c->stream));
        LSB_TRY(ph_axpy(c, dst.data(), qb, nf, kModecut[l], l == 0, c->stream));
        for (int f = 0; f < nf; ++f) src[f] = dst[f];
    }
    return LSB_OK;
}

struct TnsBufs {
    double *q, *A, *B, *tmp, *qb1, *qb2;
};

int tns_bufs(lsb_ctx* c, Stage& st, int nf, TnsBufs* b) {
    const size_t n = c->nphys * (size_t)nf;
    LSB_TRY(st.get(10, n, &b->q));
    LSB_TRY(st.get(11, n, &b->A));
    LSB_TRY(st.get(12, n, &b->B));
    LSB_TRY(st.get(13, n, &b->tmp));
    LSB_TRY(st.get(14, n, &b->qb1));
    LSB_TRY(st.get(15, n, &b->qb2));
    return LSB_OK;
}

std::vector<double*> field_ptrs(double* base, int nf, size_t stride) {
    std::vector<double*> p(nf);
    for (int f = 0; f < nf; ++f) p[f] = base + (size_t)f * stride;
    return p;
}

// filters.f90:9-49 filterall3pt on nf spectral device fields, in place.
int filterall_dev(lsb_ctx* c, Stage& st, double* const* spec, int nf) {
    TnsBufs b;
    LSB_TRY(tns_bufs(c, st, nf, &b));
    auto q = field_ptrs(b.q, nf, c->nphys), qb = field_ptrs(b.qb1, nf, c->nphys);
    std::vector<const double*> sp(spec, spec + nf), qc(q.begin(), q.end()), qbc(qb.begin(), qb.end());
    LSB_TRY(horfft_inv(c, c->hp, sp.data(), q.data(), nf, c->nzp, c->stream));
    LSB_TRY(modecut(c, qc.data(), qb.data(), nf, c->cfg.a, c->cfg.b, b.A, b.B, b.tmp));
    return horfft_fwd(c, c->hp, qbc.data(), spec, nf, c->nzp, c->stream);
}

// filters.f90:52-161 tns_automatic_filtering on u, v, w, p (spectral device fields, in place).
int tns_dev(lsb_ctx* c, Stage& st, double* const* spec, double* ih, double* idelta, int* applied) {
    const int nf = 4;
    TnsBufs b;
    LSB_TRY(tns_bufs(c, st, nf, &b));
    auto q = field_ptrs(b.q, nf, c->nphys), q2 = field_ptrs(b.qb1, nf, c->nphys), qf = field_ptrs(b.qb2, nf, c->nphys);
    std::vector<const double*> sp(spec, spec + nf), qc(q.begin(), q.end()), q2c(q2.begin(), q2.end()),
        qfc(qf.begin(), qf.end());
    LSB_TRY(horfft_inv(c, c->hp, sp.data(), q.data(), nf, c->nzp, c->stream));
    LSB_TRY(modecut(c, qc.data(), q2.data(), nf, 1.0, 6.0, b.A, b.B, b.tmp));  // 1/8 3/4 1/8
    LSB_TRY(modecut(c, qc.data(), qf.data(), nf, 1.0, 2.0, b.A, b.B, b.tmp));  // 1/4 1/2 1/4
    std::vector<double> e_delta(c->nzp), e_h(c->nzp);
    LSB_TRY(ph_plane_energy(c, qc.data(), qfc.data(), 3, e_delta.data(), c->stream));
    LSB_TRY(ph_plane_energy(c, qc.data(), q2c.data(), 3, e_h.data(), c->stream));
    const double area = (double)c->ny * (double)c->nx;
    double Ih = 0.0, Id = 0.0;
    for (int k = c->nz - 1; k >= 0; --k) {  // filters.f90:113-117, k = nz .. 1
        const double dz = c->zpts[k] - c->zpts[k + 1];
        Ih = Ih + 0.5 * (e_h[k] / area + e_h[k + 1] / area) * dz;
        Id = Id + 0.5 * (e_delta[k] / area + e_delta[k + 1] / area) * dz;
    }
    const bool apply = Ih / Id > 0.007;  // ekin_ratio_limit (Tantikul & Domaradzki 2011)
    if (apply) LSB_TRY(horfft_fwd(c, c->hp, q2c.data(), spec, nf, c->nzp, c->stream));
    if (ih) *ih = Ih;
    if (idelta) *idelta = Id;
    if (applied) *applied = apply ? 1 : 0;
    return LSB_OK;
}

}  // namespace

extern "C" int lsb_filterall3pt(lsb_ctx* c, double* uf, double* vf, double* wf, double* tempf, double* pf) {
    LSB_ENTER(c);
    if (!uf || !vf || !wf || !pf) {
        set_error("filterall3pt: null array");
        return LSB_E_ARG;
    }
    bool hot = false;  // filters.f90:34: temperature only when some |T| > 1e-10
    if (tempf)
        for (size_t q = 0; q < c->nspec && !hot; ++q) hot = std::fabs(tempf[q]) > 1.0e-10;
    double* hs[5] = {uf, vf, wf, pf, tempf};
    const int nf = hot ? 5 : 4;
    double* d;
    LSB_TRY(st.get(0, c->nspec * nf, &d));
    std::vector<double*> sp(nf);
    for (int f = 0; f < nf; ++f) {
        sp[f] = d + (size_t)f * c->nspec;
        LSB_TRY(st.in(sp[f], hs[f], c->nspec));
    }
    LSB_TRY(filterall_dev(c, st, sp.data(), nf));
    for (int f = 0; f < nf; ++f) LSB_TRY(st.out(hs[f], sp[f], c->nspec));
    return st.sync();
}

extern "C" int lsb_tns_automatic_filtering(lsb_ctx* c, double* uf, double* vf, double* wf, double* tempf, double* pf,
                                           double* ih, double* idelta, int* applied) {
    LSB_ENTER(c);
    (void)tempf;  // the temperature is not filtered here (filters.f90:120-124)
    if (!uf || !vf || !wf || !pf) {
        set_error("tns_automatic_filtering: null array");
        return LSB_E_ARG;
    }
    double* hs[4] = {uf, vf, wf, pf};
    double* d;
    LSB_TRY(st.get(0, c->nspec * 4, &d));
    std::vector<double*> sp(4);
    for (int f = 0; f < 4; ++f) {
        sp[f] = d + (size_t)f * c->nspec;
        LSB_TRY(st.in(sp[f], hs[f], c->nspec));
    }
    int ap = 0;
    LSB_TRY(tns_dev(c, st, sp.data(), ih, idelta, &ap));
    if (ap)
        for (int f = 0; f < 4; ++f) LSB_TRY(st.out(hs[f], sp[f], c->nspec));
    if (applied) *applied = ap;
    return st.sync();
}

extern "C" int lsb_state_filterall3pt(lsb_ctx* c) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "state_filterall3pt"));
    double* sp[4] = {step_field(c, 0), step_field(c, 1), step_field(c, 2), step_field(c, 15)};
    return filterall_dev(c, st, sp, 4);
}

extern "C" int lsb_state_tns_filter(lsb_ctx* c, double* ih, double* idelta, int* applied) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "state_tns_filter"));
    double* sp[4] = {step_field(c, 0), step_field(c, 1), step_field(c, 2), step_field(c, 15)};
    return tns_dev(c, st, sp, ih, idelta, applied);
}

extern "C" int lsb_set_filtering(lsb_ctx* c, int ifilt, int modfilt) {
    LSB_TRY(check_ctx(c));
    if (ifilt < 0 || (ifilt >= 1 && modfilt < 1)) {
        set_error("set_filtering: ifilt >= 0 and modfilt >= 1 when filtering");
        return LSB_E_ARG;
    }
    c->ifilt = ifilt;
    c->modfilt = modfilt;
    return LSB_OK;
}

namespace lsb {
// navierstokes.f90:31-39: before the update of step istep, every modfilt steps.
int state_filter_hook(lsb_ctx* c) {
    if (c->ifilt < 1 || c->istep % c->modfilt != 0) return LSB_OK;
    if (c->cfg.sgsflag == 1 && c->cfg.sgsmodel == 4) return lsb_state_tns_filter(c, nullptr, nullptr, nullptr);
    return lsb_state_filterall3pt(c);
}
}  // namespace lsb

// =========================================== linear-stability energy (orrsomm.f90)
namespace {

// orrsomm.f90:625-666 ener on three spectral device fields (physical in stage slot 16)
int ener_dev(lsb_ctx* c, Stage& st, const double* uf, const double* vf, const double* wf, int itot, double* en) {
    double* ph;
    LSB_TRY(st.get(16, c->nphys * 3, &ph));
    const double* si[3] = {uf, vf, wf};
    double* so[3] = {ph, ph + c->nphys, ph + 2 * c->nphys};
    LSB_TRY(horfft_inv(c, c->hp, si, so, 3, c->nzp, c->stream));
    std::vector<double> ee(c->nzp);
    LSB_TRY(plane_ke(c, so[0], so[1], so[2], itot == 1, ee.data(), c->stream));
    double e = 0.0;
    for (int k = c->nz - 1; k >= 1; --k)  // Fortran k = nz .. k0 = 2
        e = e + 0.5 * (ee[k + 1] + ee[k]) * (c->zpts[k] - c->zpts[k + 1]);
    *en = e;
    return LSB_OK;
}

// orrsomm.f90:566-622 on device fields; istep is 1-based, the arrays hold at least istep values
int lintest_dev(lsb_ctx* c, Stage& st, const double* uf, const double* vf, const double* wf, int istep, double dt,
                double* enr1, double* enr2, double* grow1, double* grow2) {
    const int i = istep - 1;
    LSB_TRY(ener_dev(c, st, uf, vf, wf, 1, &enr1[i]));
    double* cp;
    LSB_TRY(st.get(17, c->nspec * 3, &cp));
    const double* src[3] = {uf, vf, wf};
    double* dst[3] = {cp, cp + c->nspec, cp + 2 * c->nspec};
    for (int l = 0; l < 3; ++l)
        LSB_HIP(hipMemcpyAsync(dst[l], src[l], c->nspec * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    LSB_TRY(zero_mean_modes(c, dst, 3, c->stream));
    LSB_TRY(ener_dev(c, st, dst[0], dst[1], dst[2], 0, &enr2[i]));
    if (istep > 1) {
        grow1[i] = 1. / (2.0 * enr1[i]) * (enr1[i] - enr1[i - 1]) / dt;
        grow2[i] = 1. / (2.0 * enr2[i]) * (enr2[i] - enr2[i - 1]) / dt;
    } else {
        grow1[i] = 0.0;
        grow2[i] = 0.0;
    }
    return LSB_OK;
}

}  // namespace

extern "C" int lsb_ener(lsb_ctx* c, const double* uf, const double* vf, const double* wf, int itot, double* en) {
    LSB_ENTER(c);
    if (!uf || !vf || !wf || !en) return LSB_E_ARG;
    double* d;
    LSB_TRY(st.get(0, c->nspec * 3, &d));
    LSB_TRY(st.in(d, uf, c->nspec));
    LSB_TRY(st.in(d + c->nspec, vf, c->nspec));
    LSB_TRY(st.in(d + 2 * c->nspec, wf, c->nspec));
    return ener_dev(c, st, d, d + c->nspec, d + 2 * c->nspec, itot, en);
}

extern "C" int lsb_lintest(lsb_ctx* c, const double* uf, const double* vf, const double* wf, int istep, double dt,
                           double* enr1, double* enr2, double* grow1, double* grow2) {
    LSB_ENTER(c);
    if (!uf || !vf || !wf || !enr1 || !enr2 || !grow1 || !grow2 || istep < 1) {
        set_error("lintest: null array or istep < 1");
        return LSB_E_ARG;
    }
    double* d;
    LSB_TRY(st.get(0, c->nspec * 3, &d));
    LSB_TRY(st.in(d, uf, c->nspec));
    LSB_TRY(st.in(d + c->nspec, vf, c->nspec));
    LSB_TRY(st.in(d + 2 * c->nspec, wf, c->nspec));
    return lintest_dev(c, st, d, d + c->nspec, d + 2 * c->nspec, istep, dt, enr1, enr2, grow1, grow2);
}

extern "C" int lsb_state_lintest(lsb_ctx* c, double* enr1, double* enr2, double* grow1, double* grow2) {
    LSB_ENTER(c);
    LSB_TRY(need_state(c, "state_lintest"));
    if (!enr1 || !enr2 || !grow1 || !grow2 || c->istep < 1) {
        set_error("state_lintest: null array or no step taken in this run");
        return LSB_E_ARG;
    }
    // main.f90:65 calls lintest after timestep(): the time module's dt at that point
    return lintest_dev(c, st, step_field(c, 0), step_field(c, 1), step_field(c, 2), c->istep, c->dt, enr1, enr2,
                       grow1, grow2);
}

// ================================================ temperature state (common.f90:162-175)
extern "C" int lsb_set_temperature(lsb_ctx* c, const double* tbar, const double* tbarp, const double* tbarpp,
                                   double coex, double gx, double gy, double gz) {
    LSB_ENTER(c);
    return temp_set_params(c, tbar, tbarp, tbarpp, coex, gx, gy, gz);
}

extern "C" int lsb_state_set_temperature(lsb_ctx* c, const double* tempf, const double* tnf, const double* tnmf,
                                         const double* nltnf, const double* nltnmf) {
    LSB_ENTER(c);
    LSB_TRY(temp_enable(c));
    const double* hs[5] = {tempf, tnf, tnmf, nltnf, nltnmf};
    for (int q = 0; q < 5; ++q) {
        double* d = temp_field(c, q);
        if (hs[q])
            LSB_TRY(st.in(d, hs[q], c->nspec));
        else
            LSB_HIP(hipMemsetAsync(d, 0, c->nspec * sizeof(double), c->stream));
    }
    return st.sync();
}

extern "C" int lsb_state_get_temperature(lsb_ctx* c, double* tempf, double* tnf, double* tnmf, double* nltnf,
                                         double* nltnmf) {
    LSB_ENTER(c);
    double* hs[5] = {tempf, tnf, tnmf, nltnf, nltnmf};
    for (int q = 0; q < 5; ++q) {
        if (!hs[q]) continue;
        if (temp_on(c))
            LSB_TRY(st.out(hs[q], temp_field(c, q), c->nspec));
        else
            std::memset(hs[q], 0, c->nspec * sizeof(double));
    }
    return st.sync();
}

extern "C" int lsb_solve_orrsomm(int nzp, const double* d, const double* d2, const double* ubar, const double* ubarp,
                                 const double* ubarpp, const double* zpts, double xnu, double eps, double alphax,
                                 double betax, double* omega, int iglb, double* uvec, double* vvec, double* wvec) {
    if (nzp < 4 || !d || !d2 || !ubar || !ubarp || !ubarpp || !zpts || !omega || !uvec || !vvec || !wvec ||
        !(alphax != 0.0) || !(xnu > 0.0)) {
        set_error("solve_orrsomm: bad arguments");
        return LSB_E_ARG;
    }
    if (iglb != 0) {
        set_error("solve_orrsomm: only iglb = 0 is restated (iglb < 0 adds the global zggevx step, iglb > 0 "
                  "skips the solve); pass the omega guess alpha*cguess*u0 with iglb = 0");
        return LSB_E_UNSUPPORTED;
    }
    return orrsomm_solve(nzp, d, d2, ubar, ubarp, ubarpp, zpts, xnu, eps, alphax, betax, omega, uvec, vvec, wvec);
}
