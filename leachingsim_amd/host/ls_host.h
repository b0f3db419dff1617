// ls_host.h — C++ host side of the GPU hot path: the reference's Simulation
// interface over the C ABI (include/lsgpu.h), .raw input, VTK image output
// written without the VTK library, and checkpoints.
//
// Mirrors class Simulation of javalosp/LeachingSim (src/Simulation.h:29-186):
// the same method names, argument meaning and error behaviour (solver verdicts
// as bool, fatal errors as std::runtime_error).  All state is device-resident;
// host copies are made only for output and checkpoints.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lsgpu.h"

namespace lshost {

struct Grid {
    int ni = 0, nj = 0, nk = 0;   // internal (i, j, k) = CLI (zext, yext, xext)
    size_t n() const { return (size_t)ni * (size_t)nj * (size_t)nk; }
};

// ---- the reference's per-solve log lines of one step (steplog.cpp) --------
std::string step_log(const ls_step_report& r, double dt, const ls_params& P);

// ---- .raw input (src/MPIRawLoader.h:48-82) --------------------------------
// Headerless (or header-skipped) voxel file, k fastest, label_bytes per voxel
// (1: uint8 build, 2: uint16 build).  Throws if the file is short.
std::vector<uint8_t> read_raw(const std::string& path, const Grid& g, int label_bytes,
                              size_t header);
void write_raw(const std::string& path, const void* data, size_t bytes);

// ---- host copy of the state -------------------------------------------------
struct HostState {
    Grid g;
    int label_bytes = 1;
    std::vector<uint8_t> img;          // n * label_bytes
    std::vector<float> frac;
    std::vector<double> inventory, saturation, pressure, conc, acid;   // primary
    std::vector<double> qx, qy, qz, krel, pc, perm, mu;               // derived
    double label(size_t v) const {
        return label_bytes == 1 ? img[v] : (double)((const uint16_t*)img.data())[v];
    }
};

// ---- VTK image output (src/Simulation.cpp:1306-1566, src/main.cpp:496-526) --
// The reference's output flags, with its values (VTKOutput,
// src/Simulation.h:15-27), and its array names; files: <root>_<step>.pvti,
// <root>_0_<step>.vti (one piece, the whole grid), and
// <dir>/time_mapping.pvd.  Arrays are written as raw appended binary (UInt64
// block headers), points ordered i fastest as VTK's x axis, the reference's
// extent convention.
enum VtkFlags : unsigned {
    VTK_CONC = 0x001,
    VTK_FRAC = 0x004,
    VTK_TYPE = 0x008,
    VTK_PRESSURE = 0x010,
    VTK_PERMEABILITY = 0x020,
    VTK_SATURATION = 0x040,
    VTK_FLUX_VEC = 0x080,
    VTK_CAP_PRESSURE = 0x100,
    VTK_REL_PERMEABILITY = 0x200,
    VTK_ACID_CONC = 0x400,
    VTK_ALL = 0x7fd
};
// Writes the .vti piece and the .pvti master; returns the number of
// non-finite pressure values replaced by 0 (the reference's scrub, which
// also changes the simulation's pressure).
size_t write_vtk(const std::string& root, size_t tstep, HostState& s, unsigned flags);
void write_pvd(const std::string& path, const std::vector<double>& t,
               const std::vector<size_t>& steps, const std::string& root_basename);

// ---- checkpoints (src/Simulation.cpp:2235-2307) ------------------------------
// Own format <dir>/checkpoint_<step>.lsck: a header (magic, version, grid,
// label bytes, step, t, dt) then the reference's 7 primary arrays in its
// order (labels, frac, inventory, saturation, pressure, conc, acid).
// load_checkpoint also reads the reference's file of a one-rank run,
// <dir>/checkpoint_rank0_<step>.bin (the same 7 arrays, no header).
void write_checkpoint(const std::string& dir, int step, double t, double dt, const HostState& s);
// Returns true for the own format (t, dt filled), false for the reference
// layout (t, dt untouched).  Throws if neither file exists or sizes differ.
bool load_checkpoint(const std::string& dir, int step, HostState& s, double* t, double* dt);

// ---- Simulation over the C ABI ------------------------------------------------
class Simulation {
public:
    Simulation(const Grid& g, int label_bytes, const ls_params& p);
    ~Simulation();
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    // input and initial state (src/Simulation.cpp:68-74; src/main.cpp:183-199)
    void readRAW(const std::string& fname, size_t header);
    void setLabels(const std::vector<uint8_t>& img);
    // initFields + initFrac + initSaturation + initAcid + setupReactionRates
    // and the first properties; returns the initial Sulphide count
    unsigned long long initialise();

    // one step's phases (src/main.cpp:312-432), same names as the reference
    void updateRelativePermeability() {}   // folded into initProperties
    void updateCapillaryPressure() {}      // folded into initProperties
    void initProperties();
    void setupPressureEqns(double t) { t_press_ = t; }
    bool solvePressure();
    void calculateFlux();
    void updateSaturation(double dt);
    void handleSurfaceEffects();
    void setupConcentrationEqns(double dt) { dt_conc_ = dt; }
    bool solveConc();                      // conc and acid: one operator, two right-hand sides
    bool solveAcid() { return transport_ok_; }
    void handlePrecipitation();
    int updateFrac(double dt);
    void saveState();
    void restoreState();

    // one accepted step on the device (the body of src/main.cpp:312-432)
    bool step(double t, double dt);

    // host copies and files
    HostState download(bool derived = true) const;
    void upload(const HostState& s);
    void writeVTKFile(const std::string& root, size_t tstep, unsigned flags);
    void writeCheckpoint(const std::string& dir, int step, double t, double dt) const;
    bool loadCheckpoint(const std::string& dir, int step, double* t, double* dt);
    void setNetwork(const std::vector<float>& w);

    const ls_step_report& report() const { return last_; }
    const Grid& grid() const { return g_; }
    int labelBytes() const { return lb_; }

    unsigned long long initial_sulphide_count = 0;
    unsigned long long total_leached_count = 0;

private:
    int phase(int id, double t, double dt);
    void check(int st) const;
    Grid g_;
    int lb_;
    ls_ctx* ctx_ = nullptr;
    ls_step_report last_{};
    double t_press_ = 0.0, dt_conc_ = 0.0;
    bool transport_ok_ = true;
};

}  // namespace lshost
