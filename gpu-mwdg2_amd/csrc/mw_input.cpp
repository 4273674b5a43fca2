// Host reader of the drop-in input files. See mw_input.hpp.
#include "mw_input.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace mw {

namespace {

template <typename real> const char* num_fmt();
template <> const char* num_fmt<float>() { return "%f"; }
template <> const char* num_fmt<double>() { return "%lf"; }

int zmorton(int x, int y) {
  auto spread = [](int v) {
    v &= 0xffff;
    v = (v ^ (v << 8)) & 0x00ff00ff;
    v = (v ^ (v << 4)) & 0x0f0f0f0f;
    v = (v ^ (v << 2)) & 0x33333333;
    v = (v ^ (v << 1)) & 0x55555555;
    return v;
  };
  return spread(x) | (spread(y) << 1);
}

struct File {
  FILE* fp;
  explicit File(const std::string& name, bool required = true) : fp(std::fopen(name.c_str(), "r")) {
    if (!fp && required) throw InputError("cannot open " + name);
  }
  ~File() { if (fp) std::fclose(fp); }
  File(const File&) = delete;
  File& operator=(const File&) = delete;
};

// Scan lines until the first token starts with kw; returns the line.
bool find_line(const char* file, const char* kw, char (&line)[256]) {
  File f(file);
  const size_t n = std::strlen(kw);
  char tok[128] = {0};
  while (std::fgets(line, sizeof line, f.fp)) {
    char t[128] = {0};
    if (std::sscanf(line, "%127s", t) == 1) std::memcpy(tok, t, sizeof tok);
    if (std::strncmp(tok, kw, n) == 0) return true;
  }
  return false;
}

// The reference opens names relative to the working directory; a name that
// is not found there is looked up next to the .par file.
std::string beside(const char* par, const std::string& name) {
  if (name.empty()) return name;
  if (FILE* fp = std::fopen(name.c_str(), "r")) { std::fclose(fp); return name; }
  const std::string p(par);
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? name : p.substr(0, s + 1) + name;
}

template <typename real>
void read_raster(const std::string& name, std::vector<real>& out, int mesh_dim, real wall) {
  File f(name, false);                // input/read_raster_file.cpp:18-24: missing -> defaults
  if (!f.fp) return;
  char buf[256];
  int ncols = 0, nrows = 0;
  real dummy = 0;
  if (std::fscanf(f.fp, "%255s %d", buf, &ncols) != 2 || std::fscanf(f.fp, "%255s %d", buf, &nrows) != 2)
    throw InputError("bad raster header in " + name);
  if (ncols > mesh_dim || nrows > mesh_dim || ncols < 0 || nrows < 0)
    throw InputError("raster " + name + " is larger than the 2^max_ref_lvl mesh");
  char fmt[24];
  std::snprintf(fmt, sizeof fmt, "%%255s %s", num_fmt<real>());
  for (int k = 0; k < 4; ++k) std::fscanf(f.fp, fmt, buf, &dummy);
  for (int j = 0; j < nrows; ++j)
    for (int i = 0; i < ncols; ++i) {
      if (std::fscanf(f.fp, num_fmt<real>(), &dummy) != 1) throw InputError("short raster " + name);
      const bool nodata = std::fabs(dummy + real(9999.0)) < real(1e-10);
      out[(size_t)(nrows - 1 - j) * mesh_dim + i] = nodata ? wall : dummy;
    }
}

template <typename real>
void read_series(const char* par, const char* name, HostSeries<real>& s) {
  const std::string bdy = beside(par, par_str(par, "bdyfile"));
  if (bdy.empty()) throw InputError("time series '" + std::string(name) + "' needs a bdyfile");
  File f(bdy);
  char line[256], tok[128] = {0};
  const size_t n = std::strlen(name);
  bool found = false;
  while (std::fgets(line, sizeof line, f.fp)) {
    char t[128] = {0};
    if (std::sscanf(line, "%127s", t) == 1) std::memcpy(tok, t, sizeof tok);
    if (std::strncmp(tok, name, n) == 0) { found = true; break; }
  }
  if (!found) throw InputError("time series '" + std::string(name) + "' not in " + bdy);
  int rows = 0;
  char unit[64] = {0};
  if (!std::fgets(line, sizeof line, f.fp) || std::sscanf(line, "%d %63s", &rows, unit) < 1 || rows <= 0)
    throw InputError("zero-row time series '" + std::string(name) + "'");
  const int mult = !std::strncmp(unit, "seconds", 7) ? 1 : (!std::strncmp(unit, "minutes", 7) ? 60 : 3600);
  char fmt[16];
  std::snprintf(fmt, sizeof fmt, "%s %s", num_fmt<real>(), num_fmt<real>());
  for (int i = 0; i < rows; ++i) {
    real v = 0, t = 0;
    if (!std::fgets(line, sizeof line, f.fp)) throw InputError("short time series '" + std::string(name) + "'");
    std::sscanf(line, fmt, &v, &t);
    s.v.push_back(v);
    s.t.push_back(t * mult);
  }
}

int type_of(const char* t) {
  if (!std::strncmp(t, "CLOSED", 6)) return 1;
  if (!std::strncmp(t, "FREE", 4)) return 2;
  if (!std::strncmp(t, "HFIX", 4)) return 4;
  if (!std::strncmp(t, "HVAR", 4)) return 8;
  if (!std::strncmp(t, "QFIX", 4)) return 16;
  if (!std::strncmp(t, "QVAR", 4)) return 32;
  return 0;
}

}  // namespace

bool par_has(const char* file, const char* kw) {
  char line[256];
  return find_line(file, kw, line);
}

int par_int(const char* file, const char* kw) {
  File f(file);
  const size_t n = std::strlen(kw);
  char line[256], tok[128] = {0};
  int v = 0;
  while (std::fgets(line, sizeof line, f.fp)) {
    char t[128] = {0};
    int tv = v;
    const int got = std::sscanf(line, "%127s %d", t, &tv);
    if (got >= 1) std::memcpy(tok, t, sizeof tok);
    if (got == 2) v = tv;
    if (std::strncmp(tok, kw, n) == 0) return v;
  }
  return 0;
}

template <typename real>
real par_real(const char* file, const char* kw) {
  File f(file);
  const size_t n = std::strlen(kw);
  char line[256], tok[128] = {0}, fmt[24];
  std::snprintf(fmt, sizeof fmt, "%%127s %s", num_fmt<real>());
  real v = 0;
  while (std::fgets(line, sizeof line, f.fp)) {
    char t[128] = {0};
    real tv = v;
    const int got = std::sscanf(line, fmt, t, &tv);
    if (got >= 1) std::memcpy(tok, t, sizeof tok);
    if (got == 2) v = tv;
    if (std::strncmp(tok, kw, n) == 0) return v;
  }
  return real(0);
}

std::string par_str(const char* file, const char* kw) {
  File f(file);
  const size_t n = std::strlen(kw);
  char line[256], tok[128] = {0}, val[256] = {0};
  while (std::fgets(line, sizeof line, f.fp)) {
    char t[128] = {0}, tv[256] = {0};
    const int got = std::sscanf(line, "%127s %255s", t, tv);
    if (got >= 1) std::memcpy(tok, t, sizeof tok);
    if (got == 2) std::memcpy(val, tv, sizeof val);
    if (std::strncmp(tok, kw, n) == 0) return std::string(val);
  }
  return std::string();
}

// Built-in synthetic cases. Domain, friction and the 1D depths per case
// (classes/SimulationParams.cpp:15-147, classes/Depths1D.h:10-40).
template <typename real>
struct Synthetic {
  real xmin, xmax, ymin, ymax, manning, hl, hr;
};

template <typename real>
Synthetic<real> synthetic_params(int tc) {
  Synthetic<real> s{real(0.0), real(50.0), real(0.0), real(50.0), real(0.0), real(0.0), real(0.0)};
  if (tc == 9 || tc == 10 || (tc >= 11 && tc <= 14)) s.manning = real(0.02);
  if (tc == 15 || tc == 16) { s.xmax = s.ymax = real(38.0); s.manning = real(0.0125); }
  if (tc == 17 || tc == 18) { s.xmin = s.ymin = real(-50.0); s.xmax = s.ymax = real(50.0); }
  if (tc == 19) { s.xmin = real(10.0); s.xmax = real(70.0); s.ymin = real(-10.0); s.ymax = real(50.0); }
  if (tc == 21 || tc == 22) { s.xmax = s.ymax = real(75.0); }
  if (tc == 23) { s.xmin = s.ymin = real(-20.0); s.xmax = s.ymax = real(20.0); }
  if (tc == 1 || tc == 2) { s.hl = real(6.0); s.hr = real(6.0); }
  else if (tc == 3 || tc == 4) { s.hl = real(2.0); s.hr = real(2.0); }
  else if (tc == 5 || tc == 6 || tc == 11 || tc == 12) { s.hl = real(6.0); s.hr = real(2.0); }
  else if ((tc >= 7 && tc <= 10) || tc == 13 || tc == 14) { s.hl = real(6.0); s.hr = real(0.0); }
  else if (tc == 17) { s.hl = real(2.0); s.hr = real(0.0); }
  return s;
}

// Bed elevation at a node (synthetic/topo*.cuh).
template <typename real>
real synthetic_bed(real x, real y, int tc) {
  auto c_prop_1d = [](real a) {                               // topo_c_prop_1D.cuh
    real z = real(0.0);
    if (a >= real(22) && a < real(25)) z = real(0.05) * a - real(1.1);
    else if (a >= real(25) && a <= real(28)) z = real(-0.05) * a + real(1.4);
    else if (a > real(8) && a < real(12)) z = real(0.2) - real(0.05) * (a - real(10)) * (a - real(10));
    else if (a > real(39) && a < real(46.5)) z = real(0.3);
    return real(10.0) * z;
  };
  auto triangle = [](real a) {                                // topo_triangle.cuh
    const real slope = real(0.4) / real(3.0);
    if (a > real(25.5) && a <= real(28.5)) return (a - real(25.5)) * slope;
    if (a > real(28.5) && a <= real(31.5)) return real(0.4) - (a - real(28.5)) * slope;
    return real(0.0);
  };
  auto dist = [](real x, real y, real x0, real y0) { return std::sqrt((x - x0) * (x - x0) + (y - y0) * (y - y0)); };
  switch (tc) {
    case 1: case 3: case 11: case 13: return c_prop_1d(x);
    case 2: case 4: case 12: case 14: return c_prop_1d(y);
    case 15: return triangle(x);
    case 16: return triangle(y);
    case 17: return real(0.01) * x * x;                       // topo_parabolic_bowl.cuh
    case 18: return real(0.01) * y * y;
    case 19: case 20: {                                       // topo_three_cones.cuh
      const real z1 = (real(8) - dist(x, y, real(30), real(6))) / real(8);
      const real z2 = (real(8) - dist(x, y, real(30), real(24))) / real(8);
      const real z3 = real(0.3) * (real(10) - dist(x, y, real(47.5), real(15)));
      return std::fmax(std::fmax(std::fmax(z1, z2), z3), real(0.0));
    }
    case 21: {                                                // topo_diff.cuh
      real z = real(0);
      z = std::fmax(z, real(1.0) - real(0.2) * dist(x, y, real(20.0), real(15.0)));
      z = std::fmax(z, real(2.0) - real(0.5) * dist(x, y, real(40.0), real(15.0)));
      z = std::fmax(z, real(3.0) - real(0.3) * dist(x, y, real(60.0), real(15.0)));
      return z;
    }
    case 22: {                                                // topo_non_diff.cuh
      const bool yb = real(11.0) <= y && y <= real(19.0);
      if (real(16.0) <= x && x <= real(24.0) && yb) return real(0.86);
      if (real(36.0) <= x && x <= real(44.0) && yb) return real(1.78);
      if (real(56.0) <= x && x <= real(64.0) && yb) return real(2.30);
      return real(0);
    }
    default: return real(0.0);
  }
}

// Depth at a node (synthetic/h_init*.cuh). The c-property depth is the
// surface level minus the bed, negative above it, as the reference returns.
template <typename real>
real synthetic_depth(real x, real y, real z, const Synthetic<real>& s, int tc) {
  auto overtop = [&](real a) { return (a <= real(25)) ? std::fmax(real(0.0), s.hl - z) : std::fmax(real(0.0), s.hr - z); };
  switch (tc) {
    case 1: case 2: case 3: case 4: return s.hl - z;
    case 5: case 7: case 9: case 11: case 13: return overtop(x);
    case 6: case 8: case 10: case 12: case 14: return overtop(y);
    case 15: return (x < real(15.5)) ? real(0.75) : real(0.0);
    case 16: return (y < real(15.5)) ? real(0.75) : real(0.0);
    case 17: return std::fmax(real(0.0), (real(0.2) * x + real(10.0)) - z);
    case 18: return std::fmax(real(0.0), (real(0.2) * y + real(10.0)) - z);
    case 19: return real(1.0) - z;
    case 20: return (x < real(16.0)) ? std::fmax(real(1.875) - z, real(0.0)) : real(0.0);
    case 21: return real(1.95) - z;
    case 22: return real(1.78) - z;
    case 23: return (std::sqrt(x * x + y * y) < real(2.5)) ? real(2.5) : real(0.5);
    default: return real(0.0);
  }
}

// 4-node averages and (MWDG2) slopes of nodal values on the
// (md + 1) x (md + 1) node grid (mra/modal_projections.cu:28-88).
template <typename real>
void project_nodes(HostCase<real>& c, int md, const std::vector<real> (&node)[4]) {
  const size_t N = (size_t)md * md;
  const int nd = md + 1;
  for (auto& a : c.modal) a.assign(N, real(0));
  const real s3 = std::sqrt(real(3.0));
  for (size_t idx = 0; idx < N; ++idx) {
    const size_t x = idx % md, y = idx / md;
    const size_t SW = y * nd + x, SE = SW + 1, NW = (y + 1) * nd + x, NE = NW + 1;
    for (int v = 0; v < 4; ++v) {
      const std::vector<real>& a = node[v];
      c.modal[3 * v][idx] = real(0.25) * (a[NW] + a[NE] + a[SE] + a[SW]);
      if (c.solver != 1) {
        c.modal[3 * v + 1][idx] = (a[NE] - a[NW] + a[SE] - a[SW]) / (real(4.0) * s3);
        c.modal[3 * v + 2][idx] = (a[NE] - a[SE] + a[NW] - a[SW]) / (real(4.0) * s3);
      }
    }
  }
  if (c.solver == 1)
    for (int v = 0; v < 12; ++v)
      if (v % 3 != 0) c.modal[v].clear();
}

template <typename real>
void read_case(const char* par, int argc, const char* const* argv, HostCase<real>& c) {
  {
    File probe(par);
  }
  c.test_case = par_int(par, "test_case");                  // input/read_test_case.cpp:3-17
  if (c.test_case < 0 || c.test_case > 23) throw InputError("test_case must be in [0, 23]");
  if (c.test_case == 20)
    throw InputError("test_case 20 (three cones dam break, a 70 x 30 domain on a square mesh) needs non-square "
                     "cells, which are not supported");
  c.L = par_int(par, "max_ref_lvl");
  if (c.L < 2 || c.L > 14) throw InputError("max_ref_lvl must be in [2, 14]");
  c.initial_tstep = par_real<real>(par, "initial_tstep");
  c.epsilon = par_real<real>(par, "epsilon");
  c.wall_height = par_real<real>(par, "wall_height");
  if (par_has(par, "hwfv1")) { c.solver = 1; c.CFL = real(0.5); }
  else if (par_has(par, "mwdg2")) { c.solver = 2; c.CFL = real(0.3); }
  else throw InputError("invalid adaptive solver type: the .par needs a line 'hwfv1' or 'mwdg2'");
  c.limitslopes = par_has(par, "limitslopes");
  if (c.limitslopes) c.tol_Krivo = par_real<real>(par, "tol_Krivo");
  c.refine_wall = c.test_case == 0 && par_has(par, "refine_wall");   // mra/preflag_details.cu:77
  if (c.refine_wall) c.ref_thickness = par_int(par, "ref_thickness");
  c.startq2d = par_has(par, "startq2d");
  if (par_has(par, "cfl")) c.CFL = par_real<real>(par, "cfl");    // extension keyword

  const int md = 1 << c.L;
  const std::string dem = c.test_case == 0 ? beside(par, par_str(par, "DEMfile")) : std::string();
  const Synthetic<real> syn = synthetic_params<real>(c.test_case);
  if (c.test_case == 0) {
    if (dem.empty()) throw InputError("the .par names no DEMfile");
    c.xsz = par_int(dem.c_str(), "ncols");
    c.ysz = par_int(dem.c_str(), "nrows");
    c.xmin = par_real<real>(dem.c_str(), "xllcorner");
    c.ymin = par_real<real>(dem.c_str(), "yllcorner");
    c.dx = par_real<real>(dem.c_str(), "cellsize");
    c.manning = par_real<real>(par, "fpfric");
    c.monai = par_has(par, "monai");
    c.oregon = par_has(par, "oregon-seaside");
  } else {                                                   // run_simulation.cu:35-36
    c.xsz = c.ysz = md;
    c.xmin = syn.xmin;
    c.ymin = syn.ymin;
    c.dx = (syn.xmax - syn.xmin) / real(md);
    c.manning = syn.manning;
  }
  c.sim_time = par_real<real>(par, "sim_time");
  if (par_has(par, "massint")) {
    c.massint = par_real<real>(par, "massint");
    c.massint_count0 = 0;
  }
  if (par_has(par, "saveint")) {
    c.saveint = par_real<real>(par, "saveint");
    c.saveint_count0 = 0;
  }
  c.raster_out = par_has(par, "raster_out");
  c.voutput = par_has(par, "voutput");
  c.voutput_stage = par_has(par, "voutput_stage");
  c.vtk = par_has(par, "vtk");
  c.c_prop = par_has(par, "c_prop");
  c.cumulative = par_has(par, "cumulative");
  c.qoutput = par_has(par, "qoutput");
  c.elevoff = par_has(par, "elevoff");
  c.depthoff = par_has(par, "depthoff");
  if (par_has(par, "dirroot")) c.dirroot = par_str(par, "dirroot");
  if (par_has(par, "resroot")) c.resroot = par_str(par, "resroot");

  for (int i = 0; i + 1 < argc; ++i) {                       // input/read_command_line_params.cpp:12-47
    if (!std::strncmp(argv[i], "-epsilon", 8)) {
      real v = 0;
      std::sscanf(argv[i + 1], num_fmt<real>(), &v);
      c.epsilon = v;
    } else if (!std::strncmp(argv[i], "-solver", 7)) {
      if (!std::strncmp(argv[i + 1], "hwfv1", 5)) { c.solver = 1; c.CFL = real(0.5); }
      else if (!std::strncmp(argv[i + 1], "mwdg2", 5)) { c.solver = 2; c.CFL = real(0.3); }
      else throw InputError("invalid -solver value");
    }
  }
  if (c.xsz < 1 || c.ysz < 1 || c.xsz > md || c.ysz > md) throw InputError("DEM size does not fit the mesh");
  if (!(c.dx > real(0))) throw InputError("DEM cellsize must be positive");

  // boundaries (classes/Boundary.h:237-369)
  const std::string bci_name = c.test_case == 0 ? par_str(par, "bcifile") : std::string();   // Boundary.h:93
  const std::string bci = beside(par, bci_name);
  const char dchar[4] = {'N', 'E', 'S', 'W'};
  if (!bci_name.empty()) {
    for (int k = 0; k < 4; ++k) {
      File f(bci);
      char line[256] = {0}, t[16] = {0}, series[64] = {0}, dir = 0;
      real lower = 0, upper = 0, inlet = 0;
      char fmt4[40], fmt5[48];
      std::snprintf(fmt4, sizeof fmt4, "%%c %s %s %%15s", num_fmt<real>(), num_fmt<real>());
      bool found = false;
      while (std::fgets(line, sizeof line, f.fp)) {
        std::sscanf(line, fmt4, &dir, &lower, &upper, t);
        if (dir == dchar[k]) { found = true; break; }
      }
      if (!found) continue;
      HostBoundary<real>& b = c.bdy[k];
      b.type = type_of(t);
      if (b.type == 0) b.type = 1;
      if (b.type == 4 || b.type == 16) {
        std::snprintf(fmt5, sizeof fmt5, "%%c %s %s %%15s %s", num_fmt<real>(), num_fmt<real>(), num_fmt<real>());
        std::sscanf(line, fmt5, &dir, &lower, &upper, t, &inlet);
      } else if (b.type == 8 || b.type == 32) {
        std::snprintf(fmt5, sizeof fmt5, "%%c %s %s %%15s %%63s", num_fmt<real>(), num_fmt<real>());
        std::sscanf(line, fmt5, &dir, &lower, &upper, t, series);
        read_series<real>(par, series, b.series);
      }
      const real origin = (k == 0 || k == 2) ? c.xmin : c.ymin;
      b.start = (int)((lower - origin) / c.dx);
      b.end = (int)((upper - origin) / c.dx - 1);
      b.inlet = inlet;
      for (int i = 0; i < b.end - b.start + 1; ++i) {          // Boundary.h:182-235
        const int cc = b.start + i;
        const int x = (k == 2) ? cc : (k == 0) ? cc : (k == 1) ? c.xsz - 1 : 0;
        const int y = (k == 2) ? 0 : (k == 0) ? c.ysz - 1 : cc;
        if (x < 0 || y < 0 || x >= md || y >= md) throw InputError("boundary span leaves the mesh");
        b.codes.push_back(zmorton(x, y));
      }
    }
    // point sources (classes/PointSources.h:188-286)
    File f(bci);
    char line[256] = {0};
    while (std::fgets(line, sizeof line, f.fp)) {
      char t[16] = {0}, series[64] = {0}, pt = 0;
      real xs = 0, ys = 0, v = 0;
      char fmt4[40], fmt5[48];
      std::snprintf(fmt4, sizeof fmt4, "%%c %s %s %%15s", num_fmt<real>(), num_fmt<real>());
      std::sscanf(line, fmt4, &pt, &xs, &ys, t);
      if (pt != 'P') continue;
      HostSource<real> s;
      const int x = (int)((xs - c.xmin) / c.dx), y = (int)((ys - c.ymin) / c.dx);
      if (x < 0 || y < 0 || x >= md || y >= md) throw InputError("point source outside the mesh");
      s.code = zmorton(x, y);
      s.type = type_of(t);
      if (s.type == 4 || s.type == 16) {
        std::snprintf(fmt5, sizeof fmt5, "%%c %s %s %%15s %s", num_fmt<real>(), num_fmt<real>(), num_fmt<real>());
        std::sscanf(line, fmt5, &pt, &xs, &ys, t, &v);
        s.value = v;
      } else if (s.type == 8 || s.type == 32) {
        std::snprintf(fmt5, sizeof fmt5, "%%c %s %s %%15s %%63s", num_fmt<real>(), num_fmt<real>());
        std::sscanf(line, fmt5, &pt, &xs, &ys, t, series);
        read_series<real>(par, series, s.series);
      } else {
        throw InputError("point source with an unknown type");
      }
      c.srcs.push_back(s);
    }
  }
  // gauge points, used for preflagging (classes/StagePoints.h:40-140)
  if (par_has(par, "stagefile")) {
    const std::string sf = beside(par, par_str(par, "stagefile"));
    File f(sf);
    int n = 0;
    if (std::fscanf(f.fp, "%d", &n) != 1) n = 0;
    char fmt2[16];
    std::snprintf(fmt2, sizeof fmt2, "%s %s", num_fmt<real>(), num_fmt<real>());
    for (int i = 0; i < n; ++i) {
      real xs = 0, ys = 0;
      if (std::fscanf(f.fp, fmt2, &xs, &ys) != 2) throw InputError("short stage file");
      const int x = (int)((xs - c.xmin) / c.dx), y = (int)((ys - c.ymin) / c.dx);
      if (x < 0 || y < 0 || x >= md || y >= md) throw InputError("stage point outside the mesh");
      c.stage_codes.push_back(zmorton(x, y));
    }
  }

  // initial modal state (mra/get_modal_values.cu:14-52)
  const size_t N = (size_t)md * md;
  if (c.test_case != 0) {                                    // mra/init_nodal_values.cu:45-81
    const int nd = md + 1;
    std::vector<real> node[4];
    for (auto& a : node) a.assign((size_t)nd * nd, real(0));
    for (int yi = 0; yi < nd; ++yi)
      for (int xi = 0; xi < nd; ++xi) {
        const real x = c.xmin + c.dx * real(xi), y = c.ymin + c.dx * real(yi);
        const size_t i = (size_t)yi * nd + xi;
        node[3][i] = synthetic_bed<real>(x, y, c.test_case);
        node[0][i] = synthetic_depth<real>(x, y, node[3][i], syn, c.test_case);
      }
    project_nodes<real>(c, md, node);
    return;
  }
  std::vector<real> hr(N, real(0)), qxr(N, real(0)), qyr(N, real(0)), zr(N, c.wall_height);
  const std::string start = beside(par, par_str(par, "startfile"));
  if (!start.empty()) read_raster<real>(start, hr, md, c.wall_height);
  read_raster<real>(dem, zr, md, c.wall_height);
  if (c.startq2d && !start.empty()) {
    read_raster<real>(start + ".Qx", qxr, md, c.wall_height);
    read_raster<real>(start + ".Qy", qyr, md, c.wall_height);
  }
  for (auto& a : c.modal) a.clear();
  if (c.solver == 1) {                                          // input/read_and_project_modes_fv1.cu:39-53
    c.modal[0] = std::move(hr);
    c.modal[3] = std::move(qxr);
    c.modal[6] = std::move(qyr);
    c.modal[9] = std::move(zr);
    return;
  }
  // nodes from cells with the reference's flips and duplicated last row/column
  // (input/read_and_project_modes_dg2.cu:51-80), then 4-node averages and
  // slopes (mra/modal_projections.cu:71-86)
  const int nd = md + 1, iy = c.ysz + 1, ix = c.xsz + 1;
  std::vector<real> node[4];
  for (auto& a : node) a.assign((size_t)nd * nd, real(0));
  const std::vector<real>* src[4] = {&hr, &qxr, &qyr, &zr};
  for (int y = 0; y < nd; ++y)
    for (int x = 0; x < nd; ++x) {
      const int i = (x == ix - 1 || x == nd - 1) ? x - 1 : x;
      const int j = (y == iy - 1 || y == nd - 1) ? y - 1 : y;
      const size_t ii = (size_t)(nd - 1 - y) * nd + x, mi = (size_t)(md - 1 - j) * md + i;
      for (int v = 0; v < 4; ++v) node[v][ii] = (*src[v])[mi];
    }
  project_nodes<real>(c, md, node);
}

template void read_case<float>(const char*, int, const char* const*, HostCase<float>&);
template void read_case<double>(const char*, int, const char* const*, HostCase<double>&);
template float par_real<float>(const char*, const char*);
template double par_real<double>(const char*, const char*);

}  // namespace mw
