// ls_api.hip — the C ABI (include/lsgpu.h): context lifetime, host<->device
// transfers, the coupled step (src/main.cpp:303-432) and snapshot/rollback.
#include <cmath>
#include <cstring>
#include <utility>
#include <new>
#include <vector>

#include "ls_context.h"

using namespace ls;

namespace {

// Successive arrays start at different offsets modulo 64 KB (k x 4352 B for
// the k-th allocation, k mod 16), so that arrays streamed together (faces, p,
// q) do not address the same HBM channels in step.  The fused Schur apply at
// 512^3, timed alone over four fresh contexts per setting: 0.463 ms against
// 0.471 ms without the offsets (the plain stencils did not change).
#ifndef LS_ALLOC_STAGGER
#define LS_ALLOC_STAGGER 4352
#endif
template <typename T>
int dalloc(ls_ctx* c, T** p, size_t n) {
    void* q = nullptr;
    const size_t stagger = LS_ALLOC_STAGGER ? (c->allocations.size() % 16) * LS_ALLOC_STAGGER : 0;
    hipError_t e = hipMalloc(&q, n * sizeof(T) + 256 + 16 * (size_t)LS_ALLOC_STAGGER);
    if (e != hipSuccess) {
        c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        *p = nullptr;
        return -12;
    }
    c->allocations.push_back(q);
    *p = (T*)((char*)q + stagger);
    (n >= c->N ? c->grid_bytes : c->other_bytes) += n * sizeof(T) + 256 + 16 * (size_t)LS_ALLOC_STAGGER;
    return 0;
}



struct Timer {
    ls_ctx* c;
    int n = 0;
    void mark() {
        if (n < 14) (void)hipEventRecord(c->ev[n++], c->stream);
    }
    double between(int a, int b) {
        float f = 0.0f;
        if (a < n && b < n) (void)hipEventElapsedTime(&f, c->ev[a], c->ev[b]);
        return f;
    }
};

void reset_prof(ls_ctx* c) {
    c->prof_ms[0] = c->prof_ms[1] = c->prof_ms[2] = c->prof_ms[3] = 0.0;
    c->prof_n[0] = c->prof_n[1] = c->prof_n[2] = c->prof_n[3] = 0;
    c->prof_pending = 0;
}

void report_prof(ls_ctx* c, ls_step_report* r) {
    r->prof_stencil64_ms = c->prof_ms[0];
    r->prof_stencil32_ms = c->prof_ms[1];
    r->prof_stencil64_n = c->prof_n[0];
    r->prof_stencil32_n = c->prof_n[1];
    r->prof_schur_ms = c->prof_ms[2];
    r->prof_schur_n = c->prof_n[2];
    r->prof_schur_bytes = c->prof_schur_bytes;
    r->prof_tred_ms = c->prof_ms[3];
    r->prof_tred_n = c->prof_n[3];
    r->prof_tred_bytes = c->prof_tred_bytes;
}

// Offsets of the pool's buffers: each starts on a 64 KiB boundary plus the
// allocation stagger of dalloc, so buffers streamed together do not address
// the HBM channels in step.
struct PoolCut {
    size_t off = 0;
    int idx = 0;
    size_t take(size_t bytes) {
        off = (off + 65535) & ~(size_t)65535;
        off += (size_t)(idx++ % 16) * LS_ALLOC_STAGGER;
        const size_t o = off;
        off += bytes;
        return o;
    }
};

// Pressure scratch vectors the parameters need: the Schur residual pair,
// the projection basis with its images, the learned solver's FP32 vectors
// (two per vector) and its iterate backup.
int pool_pw_count(const ls_params& P) {
    int n = 2;
    if (P.warm_start != LS_WARM_PREVIOUS) {
        const int m = std::max(1, std::min(P.warm_history, LS_WARM_MAX)) +
                      (P.warm_start == LS_WARM_LEARNED ? 1 : 0);
        n = std::max(n, 2 * m);
    }
    if (P.solver == LS_SOLVER_LEARNED) n = std::max(n, 5);
    return n;
}

}  // namespace

namespace ls {

int copy_d2d(ls_ctx* c, void* dst, const void* src, size_t bytes) {
    LS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

// The snapshot's p and frac that the step left in the fields themselves,
// copied into the backups before anything else writes the fields.
int materialize_backups(ls_ctx* c) {
    int st = 0;
    if (c->bp_stale) st |= copy_d2d(c, c->b_p, c->p, c->N * 8);
    if (c->bfrac_stale) st |= copy_d2d(c, c->b_frac, c->frac, c->N * 4);
    c->bp_stale = c->bfrac_stale = false;
    return st;
}

// The phase-scoped pool: the pressure phase's and the transport phase's
// buffers over the same bytes.  Cut again (and reallocated) when the
// parameters change the layout; the pool holds nothing between phases.
int pool_setup(ls_ctx* c) {
    bool stored = false, schur = false, mixed = false;
    transport_form(c, &stored, &schur, &mixed);
    const int npw = pool_pw_count(c->P);
    const int key = (stored ? 1 : 0) | (schur ? 2 : 0) | (c->P.implicit_sat ? 4 : 0) |
                    (mixed ? 8 : 0) | (npw << 4);
    if (key == c->pool_key && c->pool) return 0;
    const size_t N = c->N, Nh = N / 2, d = sizeof(double);
    // pressure phase
    PoolCut pc;
    const size_t o_cls = pc.take(N);
    const size_t o_pD = pc.take(N * d), o_sd = pc.take(N * d), o_bhat = pc.take(N * d);
    size_t o_c64[3], o_c32[3];
    for (int a = 0; a < 3; ++a) o_c64[a] = pc.take(N * d);
    for (int a = 0; a < 3; ++a) o_c32[a] = pc.take(N * sizeof(float));
    size_t o_v64[4], o_v32[4], o_pw[2 * LS_WARM_MAX + 2];
    for (int a = 0; a < 4; ++a) o_v64[a] = pc.take(N * d);
    for (int a = 0; a < 4; ++a) o_v32[a] = pc.take(N * sizeof(float));
    for (int a = 0; a < npw; ++a) o_pw[a] = pc.take(N * d);
    // transport phase
    PoolCut tc;
    const size_t o_rows = stored ? tc.take(transport_rows_doubles(c) * d) : 0;
    const size_t o_tb0 = tc.take(N * d), o_tb1 = tc.take(N * d);
    const size_t nv = schur ? Nh : N;
    // the Schur form keeps its Jacobi scaling inside the operator: no ph, sh
    const int ntv = schur ? 6 : 8;
    size_t o_tv[2][8] = {}, o_xs[2] = {0, 0}, o_wr[2] = {0, 0}, o_tinvd = 0;
    // the mixed form lays two FP32 vectors over each of conc's rh, p, v, s,
    // the second LS_ALLOC_STAGGER bytes past the first one's end
    const size_t vpad = mixed ? 2 * LS_ALLOC_STAGGER : 0;
    for (int q = 0; q < ntv; ++q) o_tv[0][q] = tc.take(nv * d + vpad);
    // the Schur form cuts acid's six vectors as one block the size of the
    // twelve FP32 off-diagonal arrays (each vector two arrays' stride apart):
    // the mixed form keeps those arrays there while acid is known zero
    size_t o_s1 = 0;
    if (schur) {
        o_s1 = tc.take(transport_rows32_floats(c) * sizeof(float));
        const size_t vstride = 2 * (transport_rows32_floats(c) / 12) * sizeof(float);
        for (int q = 0; q < ntv; ++q) o_tv[1][q] = o_s1 + q * vstride;
    } else {
        for (int q = 0; q < ntv; ++q) o_tv[1][q] = tc.take(nv * d);
    }
    if (schur) {
        for (int f = 0; f < 2; ++f) {
            o_xs[f] = tc.take(Nh * d);
            o_wr[f] = tc.take(Nh * d);
        }
    } else {
        o_tinvd = tc.take(N * d);
    }
    size_t o_aux[2] = {0, 0};
    if (c->P.implicit_sat)
        for (int f = 0; f < 2; ++f) o_aux[f] = tc.take(N * d);
    const size_t bytes = std::max(pc.off, tc.off) + 256;
    if (c->pool) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(c->pool);
        c->pool = nullptr;
        c->grid_bytes -= c->pool_bytes;
        c->pool_bytes = 0;
    }
    c->pool_key = -1;
    c->press_assembled = false;
    c->trans_valid = false;
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c->err = std::string("hipMalloc of the phase pool (") + std::to_string(bytes >> 20) +
                 " MiB): " + hipGetErrorString(e);
        return -12;
    }
    c->pool = q;
    c->pool_bytes = bytes;
    c->grid_bytes += bytes;
    c->pool_press_bytes = pc.off;
    c->pool_trans_bytes = tc.off;
    c->pool_key = key;
    char* b = (char*)q;
    c->cls = (int8_t*)(b + o_cls);
    c->pD = (double*)(b + o_pD);
    c->sd = (double*)(b + o_sd);
    c->bhat = (double*)(b + o_bhat);
    for (int a = 0; a < 3; ++a) {
        c->c64[a] = (double*)(b + o_c64[a]);
        c->c32[a] = (float*)(b + o_c32[a]);
    }
    c->x64 = (double*)(b + o_v64[0]);
    c->r64 = (double*)(b + o_v64[1]);
    c->p64 = (double*)(b + o_v64[2]);
    c->ap64 = (double*)(b + o_v64[3]);
    c->x32 = (float*)(b + o_v32[0]);
    c->r32 = (float*)(b + o_v32[1]);
    c->p32 = (float*)(b + o_v32[2]);
    c->ap32 = (float*)(b + o_v32[3]);
    c->n_pw = npw;
    for (int a = 0; a < 2 * LS_WARM_MAX + 2; ++a) c->pw[a] = a < npw ? (double*)(b + o_pw[a]) : nullptr;
    c->t_stored = stored;
    c->t_schur = schur;
    c->arows = stored ? (double*)(b + o_rows) : nullptr;
    c->tb[0] = (double*)(b + o_tb0);
    c->tb[1] = (double*)(b + o_tb1);
    for (int f = 0; f < 2; ++f) {
        for (int k = 0; k < 8; ++k) c->tv[f][k] = k < ntv ? (double*)(b + o_tv[f][k]) : nullptr;
        c->txs[f] = schur ? (double*)(b + o_xs[f]) : nullptr;
        c->twr[f] = schur ? (double*)(b + o_wr[f]) : nullptr;
        c->taux[f] = c->P.implicit_sat ? (double*)(b + o_aux[f]) : nullptr;
    }
    c->tinvd = schur ? nullptr : (double*)(b + o_tinvd);
    c->t_mixed = mixed;
    c->trows32 = mixed ? (float*)(b + o_s1) : nullptr;
    for (int k = 0; k < 8; ++k)
        c->tf[0][k] = mixed ? (float*)((char*)c->tv[0][1 + k / 2] +
                                       (k & 1) * (Nh * sizeof(float) + LS_ALLOC_STAGGER))
                            : nullptr;
    return 0;
}

}  // namespace ls

extern "C" {

void ls_default_params(ls_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->D = 1.0;
    p->dx = 1.0;
    p->Dpore_fac = 1.0;
    p->kext = 1.0;
    p->kreac = 1.0;
    p->theta = 1.0;
    p->c_sat = 0.95;
    p->evap_flux = 0.0;
    p->top_pressure = 10000.0;
    p->max_cap_grad = 50000.0;
    p->vg_alpha = 1.0;
    p->vg_n = 2.0;
    p->s_res = 0.05;
    p->porosity = 0.1;
    p->solid_density = 1.0;
    p->time_scalar = 1.0;
    p->alpha_transfer = 1e-5;
    p->instant_precip = 0;
    p->bc_mode = LS_BC_REFERENCE;
    p->solver = LS_SOLVER_PLAIN64;
    p->warm_start = LS_WARM_PREVIOUS;
    p->press_rtol = 1e-8;
    p->press_maxit = 10000;
    p->adr_rtol = 1e-8;
    p->adr_maxit = 10000;
    p->warm_history = 1;
    p->learned_race = 1.0;
    p->implicit_sat = 0;
    p->mixed_schur = 1;
    p->adr_schur = 1;
    p->adr_extrapolate = 1;
    p->adr_mixed = 1;
    p->adr_mixed_inner_rtol = 1e-6;
    p->mixed_inner_rtol = 0.1;
}

int ls_abi_info(size_t* sp, size_t* sr) {
    if (sp) *sp = sizeof(ls_params);
    if (sr) *sr = sizeof(ls_step_report);
    return LSGPU_ABI_VERSION;
}

static int validate_params(ls_ctx* c, const ls_params* p) {
    if (!(p->dx > 0.0) || !(p->porosity > 0.0) || !(p->vg_n > 1.0)) {
        c->err = "invalid parameters: dx, porosity must be > 0 and vg_n > 1";
        return -3;
    }
    if (p->solver < 0 || p->solver > 2 || p->bc_mode < 0 || p->bc_mode > 1) {
        c->err = "invalid solver or boundary mode";
        return -3;
    }
    if (p->warm_start < 0 || p->warm_start > 2 ||
        (p->warm_start != LS_WARM_PREVIOUS &&
         (p->warm_history < 1 || p->warm_history > LS_WARM_MAX ||
          (p->warm_start == LS_WARM_LEARNED && p->warm_history > LS_WARM_MAX - 1)))) {
        c->err = "invalid warm start or warm_history (1..8, 1..7 with the network)";
        return -3;
    }
    if (!(p->learned_race >= 0.0)) {
        c->err = "learned_race must be >= 0";
        return -3;
    }
    if (p->mixed_schur < 0 || p->mixed_schur > 3 || p->adr_schur < 0 || p->adr_schur > 1 ||
        p->adr_extrapolate < 0 || p->adr_extrapolate > 1 || p->adr_mixed < 0 ||
        p->adr_mixed > 1 || !(p->adr_mixed_inner_rtol > 0.0 && p->adr_mixed_inner_rtol < 1.0)) {
        c->err = "invalid mixed_schur (0..3), adr_schur (0..1), adr_extrapolate (0..1), "
                 "adr_mixed (0..1) or adr_mixed_inner_rtol (0 < x < 1)";
        return -3;
    }
    return 0;
}

int ls_create(ls_ctx** out, int ni, int nj, int nk, int label_bytes, const ls_params* params) {
    if (!out || !params) return -1;
    *out = nullptr;
    ls_ctx* c = new (std::nothrow) ls_ctx();
    if (!c) return -1;
    if (ni < 2 || nj < 2 || nk < 2 || (label_bytes != 1 && label_bytes != 2) ||
        (double)ni * nj * nk >= 4294967295.0) {
        delete c;
        return -2;
    }
    c->ni = ni; c->nj = nj; c->nk = nk;
    c->N = (size_t)ni * nj * nk;
    c->label_bytes = label_bytes;
    if (int st = validate_params(c, params)) { delete c; return st; }
    c->P = *params;
    c->log_k_rock = std::log10(1.0e-18);  // src/Simulation.cpp:204-205
    c->log_k_pore = std::log10(1.0e-12);
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        c->err = "no HIP device";
        *out = c;
        return -4;
    }
    for (int i = 0; i < 18; ++i) (void)hipEventCreate(&c->ev[i]);
    const size_t N = c->N;
    int st = 0;
    st |= dalloc(c, &c->img, N);
    st |= dalloc(c, &c->img2, N);
    st |= dalloc(c, &c->frac, N);
    for (double** f : {&c->inv, &c->s, &c->p, &c->conc, &c->acid, &c->qx, &c->qy, &c->qz})
        st |= dalloc(c, f, N);
    st |= dalloc(c, &c->b_img, N);
    st |= dalloc(c, &c->b_frac, N);
    for (double** f : {&c->b_inv, &c->b_s, &c->b_p, &c->b_conc, &c->b_acid, &c->b_qx, &c->b_qy,
                       &c->b_qz})
        st |= dalloc(c, f, N);
    st |= dalloc(c, &c->perm, N);
    st |= dalloc(c, &c->pc, N);
    st |= dalloc(c, &c->partials, (size_t)MAX_PARTIALS * 8);
    st |= dalloc(c, &c->cg, 1);
    st |= dalloc(c, &c->bicg, 1);
    st |= dalloc(c, &c->bicg32, 1);
    st |= dalloc(c, &c->ss, 1);
    st |= dalloc(c, &c->fcg, 1);
    if (st || (st = pool_setup(c))) {
        *out = c;
        return -12;
    }
    if (hipHostMalloc((void**)&c->h_cg, sizeof(CgScalars)) != hipSuccess ||
        hipHostMalloc((void**)&c->h_bicg, sizeof(BicgScalars)) != hipSuccess ||
        hipHostMalloc((void**)&c->h_bicg32, sizeof(BicgScalars)) != hipSuccess ||
        hipHostMalloc((void**)&c->h_ss, sizeof(StepScalars)) != hipSuccess ||
        hipHostMalloc((void**)&c->h_fcg, sizeof(FcgScalars)) != hipSuccess) {
        c->err = "hipHostMalloc failed";
        *out = c;
        return -12;
    }
    // a fresh context holds Rock everywhere and zero fields
    (void)hipMemsetAsync(c->img, LS_ROCK, N, c->stream);
    (void)hipMemsetAsync(c->frac, 0, N * sizeof(float), c->stream);
    for (double* f : {c->inv, c->s, c->p, c->conc, c->acid, c->qx, c->qy, c->qz})
        (void)hipMemsetAsync(f, 0, N * sizeof(double), c->stream);
    (void)hipMemsetAsync(c->cg, 0, sizeof(CgScalars), c->stream);
    (void)hipMemsetAsync(c->bicg, 0, sizeof(BicgScalars), c->stream);
    (void)hipMemsetAsync(c->bicg32, 0, sizeof(BicgScalars), c->stream);
    (void)hipMemsetAsync(c->ss, 0, sizeof(StepScalars), c->stream);
    (void)hipMemsetAsync(c->fcg, 0, sizeof(FcgScalars), c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "initialisation failed";
        *out = c;
        return -10;
    }
    *out = c;
    return 0;
}

void ls_destroy(ls_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    learned_free(c);
    for (void* q : c->allocations) (void)hipFree(q);
    if (c->pool) (void)hipFree(c->pool);
    if (c->h_cg) (void)hipHostFree(c->h_cg);
    if (c->h_bicg) (void)hipHostFree(c->h_bicg);
    if (c->h_bicg32) (void)hipHostFree(c->h_bicg32);
    if (c->h_ss) (void)hipHostFree(c->h_ss);
    if (c->h_fcg) (void)hipHostFree(c->h_fcg);
    if (c->h_wsum) (void)hipHostFree(c->h_wsum);
    if (c->stream) {
        for (int i = 0; i < 18; ++i) (void)hipEventDestroy(c->ev[i]);
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

const char* ls_last_error(const ls_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ls_get_params(const ls_ctx* c, ls_params* out) {
    if (!c || !out) return -1;
    *out = c->P;
    return 0;
}

int ls_set_params(ls_ctx* c, const ls_params* p) {
    if (!c || !p) return -1;
    if (int st = validate_params(c, p)) return st;
    const ls_params old = c->P;
    c->P = *p;
    c->tprev_n = 0;   // the transport's path may change with the settings
    if (int st = pool_setup(c)) {
        // the old layout back (its bytes were freed first)
        const std::string why = c->err;
        c->P = old;
        (void)pool_setup(c);
        c->err = why;
        return st;
    }
    return 0;
}

static void* field_ptr(ls_ctx* c, int field, size_t* eb) {
    *eb = 8;
    switch (field) {
        case LS_F_FRAC: *eb = 4; return c->frac;
        case LS_F_INVENTORY: return c->inv;
        case LS_F_SATURATION: return c->s;
        case LS_F_PRESSURE: return c->p;
        case LS_F_CONC: return c->conc;
        case LS_F_ACID: return c->acid;
        case LS_F_QX: return c->qx;
        case LS_F_QY: return c->qy;
        case LS_F_QZ: return c->qz;
        default: return nullptr;
    }
}

int ls_upload(ls_ctx* c, int field, const void* host, size_t elem_bytes) {
    if (!c || !host) return -1;
    const size_t N = c->N;
    if (int st = materialize_backups(c)) return st;
    if (field == LS_F_IMG) {
        if (elem_bytes != (size_t)c->label_bytes) {
            c->err = "label element size does not match the context";
            return -3;
        }
        std::vector<uint8_t> lab(N);
        for (size_t v = 0; v < N; ++v) {
            unsigned x = elem_bytes == 1 ? ((const uint8_t*)host)[v] : ((const uint16_t*)host)[v];
            if (x < 1 || x > 5) {
                c->err = "label outside 1..5 at voxel " + std::to_string(v);
                return -2;
            }
            lab[v] = (uint8_t)x;
        }
        LS_HIP(hipMemcpyAsync(c->img, lab.data(), N, hipMemcpyHostToDevice, c->stream));
        LS_HIP(hipStreamSynchronize(c->stream));
        c->press_assembled = false;
        return 0;
    }
    size_t eb;
    void* d = field_ptr(c, field, &eb);
    if (!d || eb != elem_bytes) {
        c->err = "field cannot be uploaded or has another element size";
        return -3;
    }
    LS_HIP(hipMemcpyAsync(d, host, N * eb, hipMemcpyHostToDevice, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    if (field == LS_F_PRESSURE) c->n_hist = 0;   // history no longer matches p
    if (field == LS_F_CONC) c->tprev_n = 0;
    if (field == LS_F_ACID) {                    // +0.0 bit patterns only
        const uint64_t* a = (const uint64_t*)host;
        bool z = true;
        for (size_t v = 0; v < N && z; ++v) z = a[v] == 0;
        c->acid_zero = z;
    }
    if (field == LS_F_FRAC) {
        const float* f = (const float*)host;
        bool clean = true;
        for (size_t v = 0; v < N && clean; ++v) clean = !(f[v] > 0.0f && (double)f[v] <= 1e-4);
        c->frac_clean = clean;
    }
    return 0;
}

int ls_download(ls_ctx* c, int field, void* host, size_t elem_bytes) {
    if (!c || !host) return -1;
    const size_t N = c->N;
    if (field == LS_F_IMG) {
        if (elem_bytes != 1 && elem_bytes != 2) return -3;
        std::vector<uint8_t> lab(N);
        LS_HIP(hipMemcpyAsync(lab.data(), c->img, N, hipMemcpyDeviceToHost, c->stream));
        LS_HIP(hipStreamSynchronize(c->stream));
        for (size_t v = 0; v < N; ++v) {
            if (elem_bytes == 1) ((uint8_t*)host)[v] = lab[v];
            else ((uint16_t*)host)[v] = lab[v];
        }
        return 0;
    }
    if (field >= LS_F_KREL && field <= LS_F_MU) {
        if (elem_bytes != 8) return -3;
        // derived from the current state, as the reference recomputes them
        // before writing VTK output (src/main.cpp:491-494)
        c->trans_valid = false;
        if (int st = derived_field(c, field, c->pw[0])) return st;
        LS_HIP(hipMemcpyAsync(host, c->pw[0], N * 8, hipMemcpyDeviceToHost, c->stream));
        LS_HIP(hipStreamSynchronize(c->stream));
        return 0;
    }
    size_t eb;
    void* d = field_ptr(c, field, &eb);
    if (!d || eb != elem_bytes) {
        c->err = "unknown field or element size";
        return -3;
    }
    LS_HIP(hipMemcpyAsync(host, d, N * eb, hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int ls_init_state(ls_ctx* c, long long* nsul) {
    if (!c) return -1;
    if (int st = materialize_backups(c)) return st;
    int st = init_state(c, nsul);
    if (st) return st;
    LS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int ls_snapshot(ls_ctx* c) {
    if (!c) return -1;
    const size_t N = c->N;
    int st = 0;
    st |= copy_d2d(c, c->b_img, c->img, N);
    st |= copy_d2d(c, c->b_frac, c->frac, N * 4);
    st |= copy_d2d(c, c->b_inv, c->inv, N * 8);
    st |= copy_d2d(c, c->b_s, c->s, N * 8);
    st |= copy_d2d(c, c->b_p, c->p, N * 8);
    st |= copy_d2d(c, c->b_conc, c->conc, N * 8);
    st |= copy_d2d(c, c->b_acid, c->acid, N * 8);
    st |= copy_d2d(c, c->b_qx, c->qx, N * 8);
    st |= copy_d2d(c, c->b_qy, c->qy, N * 8);
    st |= copy_d2d(c, c->b_qz, c->qz, N * 8);
    c->b_acid_zero = c->acid_zero;
    c->b_frac_clean = c->frac_clean;
    c->bp_stale = c->bfrac_stale = false;
    return st;
}

// The step's own saveState: the flux fields are rewritten in full by the
// flux phase before anything in the step reads them, so instead of copying
// them the step swaps each flux field with its backup (the backup then holds
// the old fluxes; the flux phase fills the other buffer).  A step that stops
// before its flux phase swaps back and makes the copy after all
// (step_unswap_fluxes), so the fields and the backup are then exactly what a
// copying saveState leaves.
static int step_snapshot(ls_ctx* c) {
    const size_t N = c->N;
    int st = 0;
    st |= copy_d2d(c, c->b_img, c->img, N);
    // frac only changes in a leaching sweep, which acid known zero and a
    // clean frac skip; p is swapped with its backup by the pressure phase
    c->bfrac_stale = c->acid_zero && c->frac_clean;
    if (!c->bfrac_stale) st |= copy_d2d(c, c->b_frac, c->frac, N * 4);
    st |= copy_d2d(c, c->b_inv, c->inv, N * 8);
    st |= copy_d2d(c, c->b_s, c->s, N * 8);
    c->bp_stale = true;
    st |= copy_d2d(c, c->b_conc, c->conc, N * 8);
    // acid known zero is restored by zeroing it (b_acid_zero below)
    if (!c->acid_zero) st |= copy_d2d(c, c->b_acid, c->acid, N * 8);
    std::swap(c->qx, c->b_qx);
    std::swap(c->qy, c->b_qy);
    std::swap(c->qz, c->b_qz);
    c->b_acid_zero = c->acid_zero;
    c->b_frac_clean = c->frac_clean;
    return st;
}

static int step_unswap_fluxes(ls_ctx* c) {
    std::swap(c->qx, c->b_qx);
    std::swap(c->qy, c->b_qy);
    std::swap(c->qz, c->b_qz);
    const size_t N = c->N;
    int st = 0;
    st |= copy_d2d(c, c->b_qx, c->qx, N * 8);
    st |= copy_d2d(c, c->b_qy, c->qy, N * 8);
    st |= copy_d2d(c, c->b_qz, c->qz, N * 8);
    return st;
}

int ls_rollback(ls_ctx* c) {
    if (!c) return -1;
    const size_t N = c->N;
    int st = 0;
    st |= copy_d2d(c, c->img, c->b_img, N);
    if (!c->bfrac_stale) st |= copy_d2d(c, c->frac, c->b_frac, N * 4);
    st |= copy_d2d(c, c->inv, c->b_inv, N * 8);
    st |= copy_d2d(c, c->s, c->b_s, N * 8);
    if (!c->bp_stale) st |= copy_d2d(c, c->p, c->b_p, N * 8);
    st |= copy_d2d(c, c->conc, c->b_conc, N * 8);
    if (c->b_acid_zero)
        LS_HIP(hipMemsetAsync(c->acid, 0, N * 8, c->stream));
    else
        st |= copy_d2d(c, c->acid, c->b_acid, N * 8);
    st |= copy_d2d(c, c->qx, c->b_qx, N * 8);
    st |= copy_d2d(c, c->qy, c->b_qy, N * 8);
    st |= copy_d2d(c, c->qz, c->b_qz, N * 8);
    if (st) return st;
    c->acid_zero = c->b_acid_zero;
    c->frac_clean = c->b_frac_clean;
    c->tprev_n = 0;   // the last transport solve belonged to the failed attempt
    c->n_hist = 0;   // the pressure history no longer belongs to this state
    // rebuild the derived properties (src/Simulation.cpp:2347-2352)
    st = phase_properties(c);
    if (st) return st;
    LS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

static void fill_summary(ls_ctx* c, ls_step_report* r) {
    const StepScalars& h = *c->h_ss;
    r->p_min = h.sums[0];
    r->p_max = h.sums[1];
    r->conc_min = h.sums[2];
    r->conc_max = h.sums[3];
    r->acid_min = h.sums[4];
    r->acid_max = h.sums[5];
    r->mass_conc = h.sums[6];
    r->mass_acid = h.sums[7];
    r->mass_water = h.sums[8] * c->P.porosity;
    r->sat_substeps = h.sat_nsub;
    r->sat_max_divq = h.sat_max;
}

static int finish(ls_ctx* c, ls_step_report* rep) {
    int st = summarize(c, rep);
    if (st) return st;
    LS_HIP(hipMemcpyAsync(c->h_ss, c->ss, sizeof(StepScalars), hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    fill_summary(c, rep);
    return 0;
}

int ls_phase(ls_ctx* c, int phase, double t, double dt, ls_step_report* report) {
    if (!c) return -1;
    ls_step_report local;
    ls_step_report* rep = report ? report : &local;
    std::memset(rep, 0, sizeof(*rep));
    reset_prof(c);
    int st = 0;
    switch (phase) {
        case LS_PH_PROPERTIES: st = phase_properties(c); break;
        case LS_PH_PRESSURE: st = solve_pressure(c, t, rep); break;
        case LS_PH_FLUX: st = phase_flux(c); break;
        case LS_PH_SATURATION:
            st = c->P.implicit_sat ? phase_saturation_implicit(c, dt, rep)
                                   : phase_saturation(c, dt, rep);
            break;
        case LS_PH_SURFACE: st = phase_surface(c); break;
        case LS_PH_TRANSPORT: st = phase_transport(c, dt, rep); break;
        case LS_PH_PRECIPITATION: st = phase_precipitation(c); break;
        case LS_PH_LEACHING: st = phase_leaching(c, dt, rep); break;
        default: c->err = "unknown phase"; return -1;
    }
    if (st < 0) return st;
    int st2 = finish(c, rep);
    if (st2) return st2;
    report_prof(c, rep);
    rep->status = st;
    return st;
}

int ls_step(ls_ctx* c, double t, double dt, ls_step_report* report) {
    if (!c) return -1;
    ls_step_report local;
    ls_step_report* rep = report ? report : &local;
    std::memset(rep, 0, sizeof(*rep));
    if (!(dt > 0.0)) {
        c->err = "dt must be > 0";
        return -3;
    }
    reset_prof(c);
    Timer tm{c};
    tm.mark();                                                  // 0
    int st = step_snapshot(c);                                  // saveState
    if (st) return st;
    bool fluxes_swapped = true;     // until the flux phase has written them
    auto stop = [&](int code) {     // a fatal error before the flux phase
        if (fluxes_swapped) (void)step_unswap_fluxes(c);
        return code;
    };
    tm.mark();                                                  // 1
    // A. hydrodynamics: one Picard pass with the explicit saturation, three
    // with the implicit one (src/main.cpp:312-370); the phase marks are those
    // of the last pass, earlier passes count into t_props_ms
    bool ok = true;
    const int passes = c->P.implicit_sat ? 3 : 1;
    for (int pass = 0; pass < passes && ok; ++pass) {
        if (pass) tm.n = 1;
        if ((st = phase_properties(c))) return stop(st);
        tm.mark();                                              // 2
        st = solve_pressure(c, t, rep);
        if (st < 0) return stop(st);
        tm.mark();                                              // 3
        ok = st == 0;
        if (!ok) break;
        if ((st = phase_flux(c))) return stop(st);
        fluxes_swapped = false;
        tm.mark();                                              // 4
        if (c->P.implicit_sat) {
            st = phase_saturation_implicit(c, dt, rep);
            if (st < 0) return st;
            ok = st == 0;
        } else if ((st = phase_saturation(c, dt, rep))) {
            return st;
        }
        tm.mark();                                              // 5
    }
    // a pressure solve that failed before any flux phase: fields and backup
    // as a copying saveState leaves them
    if (fluxes_swapped && (st = step_unswap_fluxes(c))) return st;
    // an unphysical implicit saturation after the last pass forces a retry
    if (ok && c->P.implicit_sat && !c->sat_valid) {
        c->err = "implicit saturation: unphysical state after the Picard passes";
        ok = false;
    }
    if (ok) {
        // B. post-Picard physics
        if ((st = phase_surface(c))) return st;
        tm.mark();                                              // 6
        st = phase_transport(c, dt, rep);
        if (st < 0) return st;
        tm.mark();                                              // 7
        ok = st == 0;
    }
    if (ok) {
        if ((st = phase_precipitation(c))) return st;
        tm.mark();                                              // 8
        if ((st = phase_leaching(c, dt, rep))) return st;
        tm.mark();                                              // 9
    }
    if ((st = finish(c, rep))) return st;
    tm.mark();
    LS_HIP(hipEventSynchronize(c->ev[tm.n - 1]));
    rep->t_snapshot_ms = tm.between(0, 1);
    rep->t_props_ms = tm.between(1, 2);
    rep->t_pressure_ms = tm.between(2, 3);
    if (ok) {
        rep->t_flux_ms = tm.between(3, 4);
        rep->t_sat_ms = tm.between(4, 5);
        rep->t_surface_ms = tm.between(5, 6);
        rep->t_transport_ms = tm.between(6, 7);
        rep->t_precip_ms = tm.between(7, 8);
        rep->t_leach_ms = tm.between(8, 9);
    }
    rep->t_total_ms = tm.between(0, tm.n - 1);
    report_prof(c, rep);
    rep->status = ok ? 0 : 1;
    return rep->status;
}

int ls_set_network(ls_ctx* c, const float* w, size_t count) {
    if (!c || !w) return -1;
    learned_free(c);
    c->net_weights.assign(w, w + count);
    return 0;
}

int ls_debug_learned_pack(ls_ctx* c, const float* r_host, float* out_host, size_t out_count) {
    if (c) c->trans_valid = false;
    if (!c || !r_host || !out_host) return -1;
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    int st = learned_setup(c);
    if (st) return st;
    const size_t cn = (size_t)c->net_dims[0] * c->net_dims[1] * c->net_dims[2];
    if (out_count != cn * 16) {
        c->err = "output size must be 16 x the coarse network grid";
        return -3;
    }
    float *rd = nullptr, *od = nullptr;
    if (hipMalloc(&rd, c->N * sizeof(float)) != hipSuccess ||
        hipMalloc(&od, cn * 16 * sizeof(float)) != hipSuccess) {
        if (rd) (void)hipFree(rd);
        c->err = "hipMalloc failed";
        return -12;
    }
    if (hipMemcpyAsync(rd, r_host, c->N * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        st = -10;
    if (!st) st = learned_pack_debug(c, rd, od);
    if (!st && hipMemcpyAsync(out_host, od, cn * 16 * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        st = -10;
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = -10;
    (void)hipFree(rd);
    (void)hipFree(od);
    return st;
}

int ls_debug_pressure_system(ls_ctx* c, double* sd, double* bhat, int8_t* cls) {
    if (!c || !sd || !bhat || !cls) return -1;
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    LS_HIP(hipMemcpyAsync(sd, c->sd, c->N * 8, hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipMemcpyAsync(bhat, c->bhat, c->N * 8, hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipMemcpyAsync(cls, c->cls, c->N, hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int ls_debug_pressure_faces(ls_ctx* c, double* faces) {
    if (!c || !faces) return -1;
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    for (int a = 0; a < 3; ++a)
        LS_HIP(hipMemcpyAsync(faces + a * c->N, c->c64[a], c->N * 8, hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int ls_network_grid(ls_ctx* c, int* dims) {
    if (!c || !dims) return -1;
    auto up16 = [](int n) { int h = (n + 1) / 2; return ((h + 15) / 16) * 16; };
    dims[0] = up16(c->ni);
    dims[1] = up16(c->nj);
    dims[2] = up16(c->nk);
    return 0;
}

int ls_set_profiling(ls_ctx* c, int on) {
    if (!c) return -1;
    c->profiling = on ? 1 : 0;
    return 0;
}

// Training data of the learned coarse correction (no reference
// counterpart): the Galerkin operator of the assembled pressure system on
// the network grid, and one (input, P^T r, e.A e) sample from the current
// pressure field and a reference solution.
int ls_debug_learned_galerkin(ls_ctx* c, float* ac_host, size_t count) {
    if (c) c->trans_valid = false;
    if (!c || !ac_host) return -1;
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    int st = learned_setup(c);
    if (st) return st;
    const size_t cn = (size_t)c->net_dims[0] * c->net_dims[1] * c->net_dims[2];
    if (count != 4 * cn) {
        c->err = "output size must be 4 x the coarse network grid";
        return -3;
    }
    float* ad = nullptr;
    if (hipMalloc(&ad, 4 * cn * sizeof(float)) != hipSuccess) {
        c->err = "hipMalloc failed";
        return -12;
    }
    st = learned_galerkin_debug(c, ad);
    if (!st && hipMemcpyAsync(ac_host, ad, 4 * cn * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream) != hipSuccess)
        st = -10;
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = -10;
    (void)hipFree(ad);
    return st;
}

int ls_debug_learned_sample(ls_ctx* c, const double* pstar_host, float* x_host, size_t x_count,
                            float* rc_host, double* eae) {
    if (c) c->trans_valid = false;
    if (!c || !pstar_host || !x_host || !rc_host || !eae) return -1;
    if (!c->press_assembled) {
        c->err = "no pressure system assembled";
        return -1;
    }
    int st = learned_setup(c);
    if (st) return st;
    const size_t cn = (size_t)c->net_dims[0] * c->net_dims[1] * c->net_dims[2];
    if (x_count != cn * 16) {
        c->err = "input size must be 16 x the coarse network grid";
        return -3;
    }
    double* pd = nullptr;
    float *xd = nullptr, *rcd = nullptr;
    if (hipMalloc(&pd, c->N * sizeof(double)) != hipSuccess ||
        hipMalloc(&xd, cn * 16 * sizeof(float)) != hipSuccess ||
        hipMalloc(&rcd, cn * sizeof(float)) != hipSuccess) {
        if (pd) (void)hipFree(pd);
        if (xd) (void)hipFree(xd);
        c->err = "hipMalloc failed";
        return -12;
    }
    if (hipMemcpyAsync(pd, pstar_host, c->N * sizeof(double), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        st = -10;
    if (!st) st = learned_sample_debug(c, pd, xd, rcd, eae);
    if (!st && (hipMemcpyAsync(x_host, xd, cn * 16 * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream) != hipSuccess ||
                hipMemcpyAsync(rc_host, rcd, cn * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream) != hipSuccess))
        st = -10;
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = -10;
    (void)hipFree(pd);
    (void)hipFree(xd);
    (void)hipFree(rcd);
    return st;
}

int ls_debug_apply_pressure(ls_ctx* c, const double* x, double* y) {
    if (!c || !x || !y) return -1;
    const size_t N = c->N;
    c->trans_valid = false;
    LS_HIP(hipMemcpyAsync(c->pw[0], x, N * 8, hipMemcpyHostToDevice, c->stream));
    int st = apply_pressure_unscaled(c, c->pw[0], c->pw[1]);
    if (st) return st;
    LS_HIP(hipMemcpyAsync(y, c->pw[1], N * 8, hipMemcpyDeviceToHost, c->stream));
    LS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int ls_debug_schur_apply(ls_ctx* c, int form, const float* p, float* q, double* dot) {
    if (c) c->trans_valid = false;
    if (!c || !p || !q || !dot) return -1;
    if (form != 1 && form != 2 && form != 4) {
        c->err = "ls_debug_schur_apply: form must be 1 (fused), 2 (two passes) or 4 (two "
                 "passes on full-length vectors)";
        return -3;
    }
    return debug_schur_apply(c, form, p, q, dot);
}

int ls_time_stencil(ls_ctx* c, int precision, int reps, double* ms, double* bytes) {
    if (!c || !ms || !bytes || reps < 1) return -1;
    if (precision == LS_TIME_TRED || precision == LS_TIME_TBLACK ||
        precision == LS_TIME_TRED32 || precision == LS_TIME_TBLACK32)
        return time_transport_pass(c, precision, reps, ms, bytes);
    if (precision != 32 && precision != 64 && precision != LS_TIME_RB_GATHER &&
        precision != LS_TIME_RB_APPLY && precision != LS_TIME_SCHUR) {
        c->err = "ls_time_stencil: precision must be 64, 32, LS_TIME_RB_GATHER, LS_TIME_RB_APPLY, "
                 "LS_TIME_SCHUR, LS_TIME_TRED, LS_TIME_TBLACK, LS_TIME_TRED32 or "
                 "LS_TIME_TBLACK32";
        return -3;
    }
    c->trans_valid = false;     // pressure scratch of the pool written
    return time_stencil(c, precision, reps, ms, bytes);
}

__global__ void k_copy(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x)
        b[v] = a[v];
}

int ls_device_bytes(const ls_ctx* c, unsigned long long* grid_bytes,
                    unsigned long long* other_bytes) {
    if (!c || !grid_bytes || !other_bytes) return -1;
    *grid_bytes = c->grid_bytes;
    *other_bytes = c->other_bytes;
    return 0;
}

int ls_time_copy(ls_ctx* c, size_t bytes, int reps, double* ms) {
    if (!c || !ms || reps < 1) return -1;
    // copies between two pressure scratch vectors of the pool (8 N bytes
    // each; the assembled pressure system stays)
    size_t cap = c->N * 8;
    if (bytes > cap) bytes = cap;
    size_t n4 = bytes / 32;
    c->trans_valid = false;
    hipEvent_t e0 = c->ev[14], e1 = c->ev[15];
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, c->stream, (const double4*)c->pw[0],
                       (double4*)c->pw[1], n4);
    LS_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, c->stream,
                           (const double4*)c->pw[0], (double4*)c->pw[1], n4);
    LS_LAUNCH_CHECK();
    LS_HIP(hipEventRecord(e1, c->stream));
    LS_HIP(hipEventSynchronize(e1));
    float f = 0.0f;
    LS_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = (double)f / reps;
    return 0;
}

}  // extern "C"
