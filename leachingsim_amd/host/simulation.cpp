// simulation.cpp — class Simulation over the C ABI (include/lsgpu.h).
#include "ls_host.h"

namespace lshost {

Simulation::Simulation(const Grid& g, int label_bytes, const ls_params& p)
    : g_(g), lb_(label_bytes) {
    const int st = ls_create(&ctx_, g.ni, g.nj, g.nk, label_bytes, &p);
    if (st != 0) {
        const std::string msg = ctx_ ? ls_last_error(ctx_) : "ls_create failed";
        ls_destroy(ctx_);
        ctx_ = nullptr;
        throw std::runtime_error(msg + " (status " + std::to_string(st) + ")");
    }
}

Simulation::~Simulation() { ls_destroy(ctx_); }

void Simulation::check(int st) const {
    if (st < 0) throw std::runtime_error(ls_last_error(ctx_));
}

int Simulation::phase(int id, double t, double dt) {
    const int st = ls_phase(ctx_, id, t, dt, &last_);
    check(st);
    return st;
}

void Simulation::readRAW(const std::string& fname, size_t header) {
    setLabels(read_raw(fname, g_, lb_, header));
}

void Simulation::setLabels(const std::vector<uint8_t>& img) {
    if (img.size() != g_.n() * (size_t)lb_) throw std::runtime_error("label array size differs from the grid");
    check(ls_upload(ctx_, LS_F_IMG, img.data(), (size_t)lb_));
}

unsigned long long Simulation::initialise() {
    long long nsul = 0;
    check(ls_init_state(ctx_, &nsul));
    initial_sulphide_count = (unsigned long long)nsul;
    total_leached_count = 0;
    return initial_sulphide_count;
}

void Simulation::initProperties() { phase(LS_PH_PROPERTIES, 0.0, 0.0); }
bool Simulation::solvePressure() { return phase(LS_PH_PRESSURE, t_press_, 0.0) == 0; }
void Simulation::calculateFlux() { phase(LS_PH_FLUX, 0.0, 0.0); }
void Simulation::updateSaturation(double dt) { phase(LS_PH_SATURATION, 0.0, dt); }
void Simulation::handleSurfaceEffects() { phase(LS_PH_SURFACE, 0.0, 0.0); }
bool Simulation::solveConc() {
    transport_ok_ = phase(LS_PH_TRANSPORT, 0.0, dt_conc_) == 0;
    return transport_ok_;
}
void Simulation::handlePrecipitation() { phase(LS_PH_PRECIPITATION, 0.0, 0.0); }
int Simulation::updateFrac(double dt) {
    phase(LS_PH_LEACHING, 0.0, dt);
    return last_.leached;
}
void Simulation::saveState() { check(ls_snapshot(ctx_)); }
void Simulation::restoreState() { check(ls_rollback(ctx_)); }

bool Simulation::step(double t, double dt) {
    const int st = ls_step(ctx_, t, dt, &last_);
    check(st);
    return st == 0;
}

HostState Simulation::download(bool derived) const {
    HostState s;
    s.g = g_;
    s.label_bytes = lb_;
    const size_t n = g_.n();
    s.img.resize(n * (size_t)lb_);
    s.frac.resize(n);
    check(ls_download(ctx_, LS_F_IMG, s.img.data(), (size_t)lb_));
    check(ls_download(ctx_, LS_F_FRAC, s.frac.data(), sizeof(float)));
    struct F { int id; std::vector<double>* v; };
    std::vector<F> fs = {{LS_F_INVENTORY, &s.inventory}, {LS_F_SATURATION, &s.saturation},
                         {LS_F_PRESSURE, &s.pressure},   {LS_F_CONC, &s.conc},
                         {LS_F_ACID, &s.acid}};
    if (derived) {
        fs.insert(fs.end(), {{LS_F_QX, &s.qx}, {LS_F_QY, &s.qy}, {LS_F_QZ, &s.qz},
                             {LS_F_KREL, &s.krel}, {LS_F_PC, &s.pc}, {LS_F_PERM, &s.perm},
                             {LS_F_MU, &s.mu}});
    }
    for (const F& f : fs) {
        f.v->resize(n);
        check(ls_download(ctx_, f.id, f.v->data(), sizeof(double)));
    }
    return s;
}

void Simulation::upload(const HostState& s) {
    if (s.g.ni != g_.ni || s.g.nj != g_.nj || s.g.nk != g_.nk || s.label_bytes != lb_)
        throw std::runtime_error("host state grid differs from the simulation");
    check(ls_upload(ctx_, LS_F_IMG, s.img.data(), (size_t)lb_));
    check(ls_upload(ctx_, LS_F_FRAC, s.frac.data(), sizeof(float)));
    const std::pair<int, const std::vector<double>*> fs[] = {
        {LS_F_INVENTORY, &s.inventory}, {LS_F_SATURATION, &s.saturation},
        {LS_F_PRESSURE, &s.pressure},   {LS_F_CONC, &s.conc}, {LS_F_ACID, &s.acid}};
    for (const auto& f : fs) check(ls_upload(ctx_, f.first, f.second->data(), sizeof(double)));
}

void Simulation::writeVTKFile(const std::string& root, size_t tstep, unsigned flags) {
    HostState s = download(true);
    // the reference's writer zeroes non-finite pressure in the simulation's
    // own field as a side effect (src/Simulation.cpp:1487-1512)
    if (write_vtk(root, tstep, s, flags) > 0)
        check(ls_upload(ctx_, LS_F_PRESSURE, s.pressure.data(), sizeof(double)));
}

void Simulation::writeCheckpoint(const std::string& dir, int step, double t, double dt) const {
    write_checkpoint(dir, step, t, dt, download(false));
}

bool Simulation::loadCheckpoint(const std::string& dir, int step, double* t, double* dt) {
    HostState s;
    s.g = g_;
    s.label_bytes = lb_;
    const bool own = load_checkpoint(dir, step, s, t, dt);
    upload(s);
    // the derived properties from the loaded state (src/main.cpp:163-176)
    initProperties();
    return own;
}

void Simulation::setNetwork(const std::vector<float>& w) {
    check(ls_set_network(ctx_, w.data(), w.size()));
}

}  // namespace lshost
