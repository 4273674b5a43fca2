// Host-side reader of the reference's input files: the .par keyword file,
// ESRI ASCII rasters, .bci boundary/point-source lines, .bdy time series and
// .stage gauge points. Produces the row-major initial modal state and the
// scalar parameters of one scenario.
#pragma once

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace mw {

struct InputError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <typename real>
struct HostSeries {
  std::vector<real> t, v;
};

template <typename real>
struct HostBoundary {                 // classes/Boundary.h:17-501
  int type = 1;                       // CLOSED
  int start = 0, end = 0;
  real inlet = 0;
  HostSeries<real> series;
  std::vector<int> codes;             // finest Morton codes of the listed cells
};

template <typename real>
struct HostSource {                   // classes/PointSources.h:15-473
  int code = 0, type = 0;
  real value = 0;
  HostSeries<real> series;
};

template <typename real>
struct HostCase {
  // built-in synthetic case (input/read_test_case.cpp:3-17); 0 = raster files
  int test_case = 0;
  // solver parameters (classes/SolverParams.cu:10-47)
  int L = 0, solver = 0;
  real epsilon = 0, wall_height = 0, initial_tstep = 0, CFL = 0, tol_Krivo = real(9999.0);
  bool limitslopes = false, refine_wall = false, startq2d = false;
  int ref_thickness = 0;
  // simulation parameters (classes/SimulationParams.cpp:17-33)
  real xmin = 0, ymin = 0, sim_time = 0, manning = 0, dx = 0, g = real(9.80665);
  int xsz = 0, ysz = 0;
  bool monai = false, oregon = false;
  // stage output every massint seconds (classes/SaveInterval.h); interval
  // 9999 and count 9999 when the .par names no massint, as the reference
  real massint = real(9999.0);
  int massint_count0 = 9999;
  std::string dirroot = "res", resroot = "res";   // classes/PlottingParams.h:17-45
  // raster maps every saveint seconds (classes/PlottingParams.h:25-33)
  bool raster_out = false, voutput = false, qoutput = false, elevoff = false, depthoff = false;
  bool voutput_stage = false;     // gauge velocities to .xvelocity / .yvelocity
  bool vtk = false;               // leaf mesh and solution as VTK every saveint
  bool c_prop = false;            // leaf discharges to <dirroot>q*-c-prop.csv every saveint
  bool cumulative = false;        // .cumu rows (time, runtimes, dt, reduction) every massint
  real saveint = real(9999.0);
  int saveint_count0 = 9999;
  HostBoundary<real> bdy[4];          // N, E, S, W
  std::vector<HostSource<real>> srcs;
  std::vector<int> stage_codes;
  // raster cells h, qx, qy, z (rows bottom-up, mesh_dim wide; qx/qy empty
  // without startq2d), or for a synthetic case its 1D depths
  std::vector<real> raster[4];
  real syn_hl = 0, syn_hr = 0;
  // row-major modal state given through mw_set_initial (else empty):
  // h0 h1x h1y qx0 qx1x qx1y qy0 qy1x qy1y z0 z1x z1y
  std::vector<real> modal[12];
};

// Reads everything the reference's run_simulation reads before its time loop
// (run_simulation.cu:11-136). argv follows input/read_command_line_params.cpp.
template <typename real>
void read_case(const char* par, int argc, const char* const* argv, HostCase<real>& c);

// Keyword helpers with the reference's first-token prefix match
// (input/read_keyword_*.cpp).
bool par_has(const char* file, const char* kw);
int par_int(const char* file, const char* kw);
std::string par_str(const char* file, const char* kw);
template <typename real> real par_real(const char* file, const char* kw);

}  // namespace mw
