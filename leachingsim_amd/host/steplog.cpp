// steplog.cpp — the per-solve log lines the reference prints on rank 0, in
// its step order and formats, from one step's ls_step_report:
//   [Pressure]            src/Simulation.cpp:836-846
//   [Implicit Saturation] src/Simulation.cpp:2016-2027
//   [Saturation]          src/Simulation.cpp:1770-1779
//   [Concentration]       src/Simulation.cpp:1276-1286
//   [Acid Transport]      src/Simulation.cpp:2206-2215
//   [Leaching]            src/Simulation.cpp:455-464
// The benchmark parser of the reference reads "Sub-steps: N" and
// "[Implicit Saturation] Converged in N" from these lines
// (plot_benchmarks.py:22-24).  Numbers use the stream's default format, as
// the reference's std::cout does.  The reason code of a converged solve is
// PETSc's KSP_CONVERGED_RTOL (2): every solver here stops on the relative
// residual; a failed one prints KSP_DIVERGED_ITS (-3).
#include <sstream>

#include "ls_host.h"

namespace lshost {

std::string step_log(const ls_step_report& r, double dt, const ls_params& P) {
    std::ostringstream o;
    const bool press_ok = r.press_relres <= P.press_rtol;
    if (press_ok) {
        o << "    [Pressure] Converged in " << r.press_iters << " iterations. (Reason Code: 2)\n"
          << "    [Pressure] Min Value: " << r.p_min << " Pa | Max Value: " << r.p_max << " Pa\n";
    } else {
        o << "    [Pressure] DIVERGED/FAILED in " << r.press_iters
          << " iterations! (Error Code: -3)\n";
        return o.str();
    }
    if (P.implicit_sat) {
        if (r.sat_valid) {
            o << "    [Implicit Saturation] Converged in " << r.sat_iters
              << " iterations. (Reason Code: 2)\n"
              << "    [Implicit Saturation] Raw Min: " << r.sat_raw_min
              << " | Raw Max: " << r.sat_raw_max << "\n";
        } else {
            o << "    [Implicit Saturation] DIVERGED/FAILED in " << r.sat_iters
              << " iterations. (Error/Physics Code: -100)\n"
              << "    [Implicit Saturation] Unphysical Raw Min: " << r.sat_raw_min
              << " | Raw Max: " << r.sat_raw_max << "\n";
            return o.str();
        }
    } else {
        const int n = r.sat_substeps > 0 ? r.sat_substeps : 1;
        const double dt_sub = dt / n;
        o << "    [Saturation] Max div(q): " << r.sat_max_divq << " | Sub-steps: " << n << " ("
          << dt_sub << "s)\n";
        if (n > 1)
            o << "    [Saturation] Sub-stepping required: " << n << " steps of " << dt_sub
              << "s.\n";
    }
    const bool adr_ok = r.adr_relres <= P.adr_rtol;
    if (adr_ok) {
        o << "    [Concentration] Converged in " << r.adr_iters_conc
          << " iterations. (Reason Code: 2)\n"
          << "    [Concentration] Min Value: " << r.conc_min << " Pa | Max Value: " << r.conc_max
          << " Pa\n"
          << "    [Acid Transport] Converged in " << r.adr_iters_acid
          << " iterations. (Reason Code: 2)\n"
          << "    [Acid Transport] Min Value: " << r.acid_min << " | Max Value: " << r.acid_max
          << "\n";
    } else {
        o << "    [Concentration] DIVERGED/FAILED in " << r.adr_iters
          << " iterations! (Error Code: -3)\n";
        return o.str();
    }
    const int n = r.leach_substeps > 0 ? r.leach_substeps : 1;
    const double dt_sub = dt / n;
    o << "    [Leaching] Max rate: " << r.leach_max_rate << " | Sub-steps: " << n << " ("
      << dt_sub << "s)\n";
    if (n > 1)
        o << "    [Leaching] Sub-stepping required: " << n << " steps of " << dt_sub << "s.\n";
    return o.str();
}

}  // namespace lshost
