// lsim.cpp — the simulator's command line and time loop over the GPU path.
//
// Same options, defaults and loop as src/main.cpp of javalosp/LeachingSim
// (option table :39; setup :69-148; initialisation or restart :151-199; the
// adaptive step loop with retry manager, output cadence, .pvd tracking,
// checkpoints and dt ramp :293-541).  Options of the GPU path (solver
// choice, boundary mode, label width) are listed under --help.  Unknown
// options (the reference passes PETSc options through) are ignored with a
// notice.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sys/stat.h>

#include "ls_host.h"

namespace {

struct Opt {
    std::string value;
    std::string help;
    bool required = false;
    bool set = false;
};

class Options {
public:
    void add(const std::string& name, const std::string& def, const std::string& help,
             bool required = false) {
        Opt o;
        o.value = def;
        o.help = help;
        o.required = required;
        order_.push_back(name);
        opts_[name] = o;
    }
    // --name value | --name=value; returns false for --help
    bool parse(int argc, char** argv) {
        for (int a = 1; a < argc; ++a) {
            std::string s = argv[a];
            if (s == "--help" || s == "-h") return false;
            if (s.rfind("--", 0) != 0) {
                std::cerr << "[lsim] ignoring argument " << s << std::endl;
                continue;
            }
            std::string name = s.substr(2), val;
            const size_t eq = name.find('=');
            bool have = false;
            if (eq != std::string::npos) {
                val = name.substr(eq + 1);
                name = name.substr(0, eq);
                have = true;
            }
            auto it = opts_.find(name);
            if (it == opts_.end()) {
                std::cerr << "[lsim] ignoring unregistered option --" << name << std::endl;
                if (!have && a + 1 < argc && std::string(argv[a + 1]).rfind("--", 0) != 0) ++a;
                continue;
            }
            if (!have) {
                if (a + 1 >= argc) throw std::runtime_error("option --" + name + " needs a value");
                val = argv[++a];
            }
            it->second.value = val;
            it->second.set = true;
        }
        for (const auto& n : order_)
            if (opts_[n].required && !opts_[n].set)
                throw std::runtime_error("the option '--" + n + "' is required but missing");
        return true;
    }
    void usage(std::ostream& os) const {
        os << "Command line arguments:\n";
        for (const auto& n : order_) {
            const Opt& o = opts_.at(n);
            os << "  --" << std::left << std::setw(22) << n << o.help;
            if (!o.required) os << " (default " << o.value << ")";
            os << "\n";
        }
    }
    std::string str(const std::string& n) const { return opts_.at(n).value; }
    double num(const std::string& n) const { return std::stod(opts_.at(n).value); }
    long long integer(const std::string& n) const { return std::stoll(opts_.at(n).value); }
    bool flag(const std::string& n) const {
        std::string v = opts_.at(n).value;
        for (auto& ch : v) ch = (char)std::tolower(ch);
        if (v == "true" || v == "1" || v == "yes" || v == "on") return true;
        if (v == "false" || v == "0" || v == "no" || v == "off") return false;
        throw std::runtime_error("option --" + n + ": not a boolean: " + opts_.at(n).value);
    }

private:
    std::vector<std::string> order_;
    std::map<std::string, Opt> opts_;
};

void make_dirs(const std::string& path) {
    std::string cur;
    for (size_t q = 0; q <= path.size(); ++q) {
        if (q == path.size() || path[q] == '/') {
            if (!cur.empty()) mkdir(cur.c_str(), 0755);
        }
        if (q < path.size()) cur += path[q];
    }
}

std::vector<float> read_weights(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f.is_open()) throw std::runtime_error("cannot open network weights " + path);
    const size_t bytes = (size_t)f.tellg();
    if (bytes != ls_net_num_weights() * sizeof(float))
        throw std::runtime_error("network weights " + path + ": wrong size");
    std::vector<float> w(bytes / sizeof(float));
    f.seekg(0);
    f.read((char*)w.data(), (std::streamsize)bytes);
    return w;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        Options o;
        // the reference's options and defaults (src/main.cpp:39)
        o.add("raw-file", "", "Input RAW file specifying the domain.", true);
        o.add("xext", "", "The x extent of the domain", true);
        o.add("yext", "", "The y extent of the domain", true);
        o.add("zext", "", "The z extent of the domain", true);
        o.add("out_dir", "./output", "The output directory");
        o.add("header_size", "0", "RAW file header size in bytes.");
        o.add("voxel_size", "1", "Size of voxels (m).");
        o.add("D", "1", "Diffusion constant (m^2/s).");
        o.add("kext", "1", "Mass transfer to exterior.");
        o.add("kreac", "1", "Reaction transfer from sulphide grains.");
        o.add("Dpore_fac", "1", "Pore diffusivity enhancement factor");
        o.add("dt", "1", "Simulation time step (s).");
        o.add("tmax", "10", "Maximum simulation time (s).");
        o.add("nout", "1", "Output every N-th time step.");
        o.add("seed", "42", "Pseudo-RNG seed value.");
        o.add("csat", "0.95", "Saturation concentration limit.");
        o.add("evap_flux", "0.0", "Evaporative flux (m/s)");
        o.add("theta", "1.0", "Time scheme (0=Explicit, 0.5=Crank-Nicolson, 1.0=Implicit)");
        o.add("instant_precip", "false", "Enable instant pore blinding upon supersaturation");
        o.add("porosity", "0.1", "Material porosity");
        o.add("top_pressure", "10000.0", "Applied pressure head at the top boundary (Pa)");
        o.add("max_cap_grad", "50000.0", "Maximum capillary pressure gradient (Pa/m)");
        o.add("implicit_sat", "false", "Toggle: True = Implicit upwind solve in 3 Picard passes, False = Explicit Sub-stepping");
        o.add("restart_step", "0", "Step to restart from (0 = fresh start)");
        o.add("chk_freq", "1000", "Number of steps between saving binary checkpoints");
        o.add("debug", "false", "Enable detailed runtime debugging");
        o.add("use_out_time", "false", "Output based on physical time instead of iterations");
        o.add("out_time", "3600.0", "Physical time interval between VTK outputs (seconds)");
        o.add("time_scalar", "1.0", "Physical time accelerator for solid mass transfer");
        o.add("enable_microporosity", "false", "Acid diffusion through the rock matrix (read only by code the reference has commented out: no effect)");
        o.add("rock_D_fac", "1e-4", "Diffusion multiplier inside the rock matrix (no effect, as above)");
        o.add("alpha_transfer", "1e-5", "Mass transfer coefficient for acid entering the rock matrix");
        // options of the GPU path
        o.add("label_bytes", "1", "Bytes per label in the RAW file: 1 (uint8) or 2 (uint16)");
        o.add("bc", "reference", "Pressure boundary: reference (Air voxels) | zfaces (inlet/outlet planes)");
        o.add("solver", "mixed", "Pressure solver: plain | mixed | learned");
        o.add("warm", "projected", "Pressure warm start: previous | projected | learned");
        o.add("weights", "", "U-Net weights for the learned solver / warm start");
        o.add("max_steps", "0", "Stop after this many accepted steps (0 = run to tmax)");
        if (!o.parse(argc, argv)) {
            std::cout << "Dynamic Advection-Diffusion Leaching Simulator (MI355X hot path)\n";
            o.usage(std::cout);
            return 0;
        }
        (void)o.flag("enable_microporosity");   // validated; inert (see its help text)

        ls_params p;
        ls_default_params(&p);
        p.D = o.num("D");
        p.dx = o.num("voxel_size");
        p.kext = o.num("kext");
        p.kreac = o.num("kreac");
        p.Dpore_fac = o.num("Dpore_fac");
        p.c_sat = o.num("csat");
        p.evap_flux = o.num("evap_flux");
        p.top_pressure = o.num("top_pressure");
        p.max_cap_grad = o.num("max_cap_grad");
        p.theta = o.num("theta");
        p.instant_precip = o.flag("instant_precip") ? 1 : 0;
        p.porosity = o.num("porosity");
        p.time_scalar = o.num("time_scalar");
        p.alpha_transfer = o.num("alpha_transfer");
        p.implicit_sat = o.flag("implicit_sat") ? 1 : 0;
        const std::string bc = o.str("bc"), sv = o.str("solver"), wm = o.str("warm");
        p.bc_mode = bc == "zfaces" ? LS_BC_ZFACES : LS_BC_REFERENCE;
        p.solver = sv == "plain" ? LS_SOLVER_PLAIN64 : sv == "learned" ? LS_SOLVER_LEARNED : LS_SOLVER_MIXED;
        p.warm_start = wm == "previous" ? LS_WARM_PREVIOUS : wm == "learned" ? LS_WARM_LEARNED : LS_WARM_PROJECTED;

        // (i, j, k) = (Z, Y, X) (src/main.cpp:93-94)
        lshost::Grid g;
        g.ni = (int)o.integer("zext");
        g.nj = (int)o.integer("yext");
        g.nk = (int)o.integer("xext");
        const int lb = (int)o.integer("label_bytes");
        const std::string out_dir = o.str("out_dir");

        lshost::Simulation sim(g, lb, p);
        if (p.solver == LS_SOLVER_LEARNED || p.warm_start == LS_WARM_LEARNED) {
            if (o.str("weights").empty())
                throw std::runtime_error("--solver/--warm learned needs --weights");
            sim.setNetwork(read_weights(o.str("weights")));
        }
        sim.readRAW(o.str("raw-file"), (size_t)o.integer("header_size"));

        // time step stability check (src/main.cpp:104-148)
        const double dt_user = o.num("dt");
        const double dt_stable_max = (p.dx * p.dx) / (2.0 * p.D + 1e-30);
        double dt_actual = dt_user;
        if (p.theta >= 0.5) {
            std::cout << "Using Theta=" << p.theta << " (Unconditionally Stable). Using user dt = "
                      << dt_actual << "s." << std::endl;
        } else if (dt_user > dt_stable_max) {
            dt_actual = 0.9 * dt_stable_max;
            std::cout << "[Warning] Explicit scheme requested. User dt (" << dt_user
                      << "s) exceeds stability limit (" << dt_stable_max
                      << "s). Using adjusted dt = " << dt_actual << "s." << std::endl;
        } else {
            std::cout << "Using user-specified dt = " << dt_actual << "s (stable)." << std::endl;
        }

        double t = 0.0;
        size_t n = 0;
        const int start_step = (int)o.integer("restart_step");
        if (start_step > 0) {
            std::cout << "Restarting simulation from step " << start_step << "..." << std::endl;
            sim.initialise();   // sulphide count and fresh scratch state; fields replaced below
            double tc = 0.0, dtc = 0.0;
            const bool own = sim.loadCheckpoint(out_dir, start_step, &tc, &dtc);
            n = (size_t)start_step;
            // the reference restarts at start_step x dt (src/main.cpp:178-180); the
            // own format records the time itself
            t = own ? tc : start_step * dt_actual;
        } else {
            std::cout << "Initialising new simulation..." << std::endl;
            sim.initialise();
        }

        const double tmax = o.num("tmax");
        const long long nout = o.integer("nout");
        const bool use_out_time = o.flag("use_out_time");
        const double out_time_interval = o.num("out_time");
        double next_out_time = std::floor(t / out_time_interval) * out_time_interval;
        if (t > 0.0 && t >= next_out_time) next_out_time += out_time_interval;

        std::cout << "\n======================================\n  SIMULATION PARAMETERS\n"
                  << "========================================\n"
                  << "  Input file                       : " << o.str("raw-file") << "\n"
                  << "  Output folder                    : " << out_dir << "\n"
                  << "  Grid (i, j, k)                   : " << g.ni << " x " << g.nj << " x " << g.nk << "\n"
                  << "  Voxel size (dx)                  : " << p.dx << " m\n"
                  << "  Porosity                         : " << p.porosity << "\n"
                  << "  Diffusion (D)                    : " << p.D << "\n"
                  << "  Reaction rate (kreac)            : " << p.kreac << "\n"
                  << "  Saturation limit (csat)          : " << p.c_sat << "\n"
                  << "  Top Pressure (Irrigation)        : " << p.top_pressure << " Pa\n"
                  << "  Evaporative flux                 : " << p.evap_flux << "\n"
                  << "  Time step (dt)                   : " << dt_user << " s\n"
                  << "  Time discretisation (Theta)      : " << p.theta << "\n"
                  << "  Precipitation mode               : " << (p.instant_precip ? "True (Instant)" : "False (Gradual)") << "\n"
                  << "  Saturation Solver                : " << (p.implicit_sat ? "Implicit (upwind, 3 Picard passes)" : "Explicit (Sub-stepping)") << "\n"
                  << "  Pressure solver / warm start     : " << sv << " / " << wm << " (boundary " << bc << ")\n"
                  << "  Output Mode                      : " << (use_out_time ? "Physical time based (seconds)" : "Time steps based (timesteps)") << "\n"
                  << "==================================================\n" << std::endl;

        make_dirs(out_dir);
        std::cout << "\n*** Starting Simulation ***" << std::endl;
        const double dt_target = dt_actual;
        double current_dt = dt_target;
        const bool use_ramping = true;
        if (use_ramping && start_step == 0) {
            current_dt = std::min(60.0, dt_target);
            std::cout << "Temporal Ramping Enabled: Starting at dt = " << current_dt << "s" << std::endl;
        }
        std::vector<double> history_t;
        std::vector<size_t> history_n;
        const long long chk_freq = o.integer("chk_freq");
        const long long max_steps = o.integer("max_steps");
        long long accepted = 0;
        const std::string root = out_dir + "/output";

        do {
            bool ok = false;
            while (!ok) {
                std::cout << "\nIteration: " << n << " | Time: " << t << "s | dt: " << current_dt << "s"
                          << std::endl;
                ok = sim.step(t, current_dt);   // snapshots the state first (saveState)
                const ls_step_report& r = sim.report();
                std::cout << lshost::step_log(r, current_dt, p)
                          << "    [GPU] pressure relative residual " << r.press_relres
                          << (r.press_fallback ? " (fallback)" : "") << ", "
                          << r.t_total_ms << " ms" << std::endl;
                if (ok) {
                    sim.total_leached_count += (unsigned long long)r.leached;
                    const double pct = sim.initial_sulphide_count > 0
                                           ? 100.0 * (double)sim.total_leached_count / (double)sim.initial_sulphide_count
                                           : 0.0;
                    std::cout << "Done iteration. [" << r.leached << " voxels fully leached this step.]\n"
                              << "    [Leaching Progress] Total Leached: " << sim.total_leached_count << " / "
                              << sim.initial_sulphide_count << " (" << std::fixed << std::setprecision(4) << pct
                              << "%)" << std::defaultfloat << std::setprecision(6) << std::endl;
                } else {
                    std::cout << "    [Retry Manager] Physics diverged! Rolling back memory and halving dt..."
                              << std::endl;
                    sim.restoreState();
                    current_dt *= 0.5;
                    if (current_dt < 1e-3)
                        throw std::runtime_error("Time step shrank below 1ms. Simulation is irreversibly unstable.");
                }
            }
            bool trigger = false;
            if (use_out_time) {
                if (t >= next_out_time - 1e-6) {
                    trigger = true;
                    while (next_out_time <= t) next_out_time += out_time_interval;
                }
            } else if (nout > 0 && n % (size_t)nout == 0) {
                trigger = true;
            }
            if (trigger) {
                std::cout << "Writing output for time step " << n << "..." << std::endl;
                sim.writeVTKFile(root, n, lshost::VTK_ALL);
                history_t.push_back(t);
                history_n.push_back(n);
                lshost::write_pvd(out_dir + "/time_mapping.pvd", history_t, history_n, "output");
            }
            if (n > 0 && chk_freq > 0 && n % (size_t)chk_freq == 0)
                sim.writeCheckpoint(out_dir, (int)n, t, current_dt);
            t += current_dt;
            n++;
            ++accepted;
            if (use_ramping && current_dt < dt_target) {
                current_dt *= 1.5;
                if (current_dt > dt_target) {
                    current_dt = dt_target;
                    std::cout << "    [Time Manager] Ramping complete. Reached target dt = " << dt_target
                              << "s." << std::endl;
                } else {
                    std::cout << "    [Time Manager] Accelerating next dt to " << current_dt << "s." << std::endl;
                }
            }
            if (max_steps > 0 && accepted >= max_steps) break;
        } while (t < tmax);
        std::cout << "\n*** Simulation Finished ***" << std::endl;
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "An unhandled std::exception occurred:\n\t" << e.what() << std::endl;
        return 1;
    }
}
