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
  for (auto& a : c.modal) a.clear();
  for (auto& a : c.raster) a.clear();
  if (c.test_case != 0) {                   // nodal values and projection on the device
    c.syn_hl = syn.hl;
    c.syn_hr = syn.hr;
    return;
  }
  // the raster cells, rows bottom-up; the device builds the modal values
  // (input/read_and_project_modes_{fv1,dg2}.cu, mra/modal_projections.cu)
  c.raster[0].assign(N, real(0));
  c.raster[3].assign(N, c.wall_height);
  const std::string start = beside(par, par_str(par, "startfile"));
  if (!start.empty()) read_raster<real>(start, c.raster[0], md, c.wall_height);
  read_raster<real>(dem, c.raster[3], md, c.wall_height);
  if (c.startq2d && !start.empty()) {
    c.raster[1].assign(N, real(0));
    c.raster[2].assign(N, real(0));
    read_raster<real>(start + ".Qx", c.raster[1], md, c.wall_height);
    read_raster<real>(start + ".Qy", c.raster[2], md, c.wall_height);
  }
}

template void read_case<float>(const char*, int, const char* const*, HostCase<float>&);
template void read_case<double>(const char*, int, const char* const*, HostCase<double>&);
template float par_real<float>(const char*, const char*);
template double par_real<double>(const char*, const char*);

}  // namespace mw
