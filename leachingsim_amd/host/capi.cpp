// capi.cpp — C entry points of the host file formats (liblshost.so), so the
// .raw reader, the VTK writer and the checkpoint format can be checked
// without a GPU.  Returns 0 on success, -1 with lshost_last_error() set.
#include <cstring>

#include "ls_host.h"

namespace {
thread_local std::string g_err;

lshost::HostState make_state(int ni, int nj, int nk, int label_bytes) {
    lshost::HostState s;
    s.g.ni = ni; s.g.nj = nj; s.g.nk = nk;
    s.label_bytes = label_bytes;
    return s;
}

template <typename T>
void fill(std::vector<T>& v, const T* src, size_t n) {
    v.assign(src, src + n);
}
}  // namespace

extern "C" {

const char* lshost_last_error(void) { return g_err.c_str(); }

// The reference's log lines for one step (written to buf, NUL-terminated;
// returns the full length, so a short buffer can be retried).
int lshost_step_log(const ls_step_report* r, double dt, const ls_params* P, char* buf,
                    size_t n) {
    const std::string s = lshost::step_log(*r, dt, *P);
    if (buf && n > 0) {
        const size_t m = s.size() < n - 1 ? s.size() : n - 1;
        std::memcpy(buf, s.data(), m);
        buf[m] = 0;
    }
    return (int)s.size();
}

int lshost_read_raw(const char* path, int ni, int nj, int nk, int label_bytes, size_t header,
                    void* out) {
    try {
        lshost::Grid g{ni, nj, nk};
        const auto v = lshost::read_raw(path, g, label_bytes, header);
        std::memcpy(out, v.data(), v.size());
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// fields: img (label_bytes each), frac (float), then double arrays in the
// order inventory, saturation, pressure, conc, acid, qx, qy, qz, krel, pc,
// perm, mu (any may be null when its flag is off).  Returns the number of
// non-finite pressure values replaced by 0, or -1.
long long lshost_write_vtk(const char* root, size_t tstep, int ni, int nj, int nk,
                           int label_bytes, unsigned flags, const void* img, const float* frac,
                           const double* const* f64) {
    try {
        auto s = make_state(ni, nj, nk, label_bytes);
        const size_t n = s.g.n();
        if (img) fill(s.img, (const uint8_t*)img, n * (size_t)label_bytes);
        if (frac) fill(s.frac, frac, n);
        std::vector<double>* dst[12] = {&s.inventory, &s.saturation, &s.pressure, &s.conc, &s.acid,
                                        &s.qx, &s.qy, &s.qz, &s.krel, &s.pc, &s.perm, &s.mu};
        for (int q = 0; q < 12; ++q)
            if (f64 && f64[q]) fill(*dst[q], f64[q], n);
            else dst[q]->assign(n, 0.0);
        if (s.img.empty()) s.img.assign(n * (size_t)label_bytes, 0);
        if (s.frac.empty()) s.frac.assign(n, 0.0f);
        return (long long)lshost::write_vtk(root, tstep, s, flags);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int lshost_write_pvd(const char* path, const double* t, const size_t* steps, size_t count,
                     const char* root_basename) {
    try {
        lshost::write_pvd(path, std::vector<double>(t, t + count),
                          std::vector<size_t>(steps, steps + count), root_basename);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// primary arrays: img, frac, then inventory, saturation, pressure, conc, acid
int lshost_write_checkpoint(const char* dir, int step, double t, double dt, int ni, int nj, int nk,
                            int label_bytes, const void* img, const float* frac,
                            const double* const* f64) {
    try {
        auto s = make_state(ni, nj, nk, label_bytes);
        const size_t n = s.g.n();
        fill(s.img, (const uint8_t*)img, n * (size_t)label_bytes);
        fill(s.frac, frac, n);
        std::vector<double>* dst[5] = {&s.inventory, &s.saturation, &s.pressure, &s.conc, &s.acid};
        for (int q = 0; q < 5; ++q) fill(*dst[q], f64[q], n);
        lshost::write_checkpoint(dir, step, t, dt, s);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Returns 1 (own format, t and dt set), 0 (the reference's one-rank file) or -1.
int lshost_load_checkpoint(const char* dir, int step, int ni, int nj, int nk, int label_bytes,
                           void* img, float* frac, double* const* f64, double* t, double* dt) {
    try {
        auto s = make_state(ni, nj, nk, label_bytes);
        const bool own = lshost::load_checkpoint(dir, step, s, t, dt);
        std::memcpy(img, s.img.data(), s.img.size());
        std::memcpy(frac, s.frac.data(), s.frac.size() * sizeof(float));
        const std::vector<double>* src[5] = {&s.inventory, &s.saturation, &s.pressure, &s.conc, &s.acid};
        for (int q = 0; q < 5; ++q) std::memcpy(f64[q], src[q]->data(), src[q]->size() * sizeof(double));
        return own ? 1 : 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
