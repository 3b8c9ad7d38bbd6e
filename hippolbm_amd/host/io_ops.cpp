// io_ops.cpp -- the output operators of the operator layer: write_paraview,
// plane_velocity_profile and plot_line_velocity, with the reference's file
// names, directory layout and text formats.  The statistics and line samples
// are computed on the device (hlbm_plane_profile / hlbm_line_velocity); the
// paraview writer copies the fields once and formats the text on host threads.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "operators.hpp"

namespace hlbm_host {

namespace {

void check(int rc, const char* what) {
  if (rc != HLBM_OK) throw OperatorError(std::string(what) + ": " + hlbm_last_error());
}

// onika::format_string(fmt, timestep) as the reference's file names use it:
// one integer conversion (%d, %i, %ld with flags / width), "%%" for '%'.
std::string format_step(const std::string& fmt, long step) {
  std::string out;
  bool used = false;
  for (size_t i = 0; i < fmt.size(); i++) {
    if (fmt[i] != '%') {
      out += fmt[i];
      continue;
    }
    if (i + 1 < fmt.size() && fmt[i + 1] == '%') {
      out += '%';
      i++;
      continue;
    }
    size_t j = i + 1;
    std::string spec = "%";
    while (j < fmt.size() && std::string("-+ 0#").find(fmt[j]) != std::string::npos) spec += fmt[j++];
    while (j < fmt.size() && std::isdigit((unsigned char)fmt[j])) spec += fmt[j++];
    while (j < fmt.size() && fmt[j] == 'l') j++;
    if (j >= fmt.size() || (fmt[j] != 'd' && fmt[j] != 'i') || used)
      throw OperatorError("file name format '" + fmt + "': only one integer conversion (%d) is supported");
    spec += "ld";
    char buf[64];
    std::snprintf(buf, sizeof buf, spec.c_str(), step);
    out += buf;
    used = true;
    i = j;
  }
  return out;
}

// std::ostream's default formatting of a double at precision p ("%.pg")
inline void put_g(std::string& s, double v, int prec = 6) {
  char b[40];
  const int n = std::snprintf(b, sizeof b, "%.*g", prec, v);
  s.append(b, (size_t)n);
}

void write_text(const std::string& path, const std::string& text) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw OperatorError("cannot create " + path);
  f.write(text.data(), (std::streamsize)text.size());
  if (!f) throw OperatorError("cannot write " + path);
}

// Text for items [0, n) produced by fn(i, out) on host threads, concatenated in order.
template <class Fn>
std::string parallel_text(size_t n, Fn fn) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::string> part(nt);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      const size_t a = n * t / nt, b = n * (t + 1) / nt;
      std::string& s = part[t];
      s.reserve((b - a) * 16);
      for (size_t i = a; i < b; i++) fn(i, s);
    });
  for (auto& x : th) x.join();
  size_t total = 0;
  for (auto& p : part) total += p.size();
  std::string out;
  out.reserve(total);
  for (auto& p : part) out += p;
  return out;
}

std::string slot_string(const msp::Node& a, const char* key, const std::string& dflt) {
  const msp::Node* n = a.find(key);
  return n && n->is_scalar() ? n->text : dflt;
}

bool slot_bool(const msp::Node& a, const char* key, bool dflt) {
  const msp::Node* n = a.find(key);
  return n && n->is_scalar() ? n->boolean() : dflt;
}

std::string output_dir(const Slots& s, const std::string& own) { return own.empty() ? s.output_directory : own; }

// ---- write_paraview (src/io/write_paraview.cpp:40-85,
//      src/io/include/hippoLBM/io/write_paraview.hpp:139-327) ---------------
// Writes <output_directory>/ParaviewOutput/<filename>.pvtr and
// <output_directory>/ParaviewOutput/<filename>/000000.vtr (ascii
// RectilinearGrid over the All box, two ghost layers per side): P =
// c^2*avg_rho/3*(rho-1) and U = dx/dtLB*u on fluid nodes (0 elsewhere), OBST
// the obstacle value, optionally Fi (the 19 populations).  The fields are the
// rho/u slots as held.  Ghost nodes carry their periodic image on periodic
// axes (update_ghost); on the other axes they hold the state the last
// set_distribution without bounds gave them (rho = value, u = 0, fluid).
struct WriteParaview : OperatorNode {
  std::string filename = "hippoLBM_%010d", outdir;
  bool distributions = false, display_filename = true;
  explicit WriteParaview(const msp::Node& a) {
    filename = slot_string(a, "filename", filename);
    outdir = slot_string(a, "output_directory", "");
    distributions = slot_bool(a, "distributions", false);
    display_filename = slot_bool(a, "display_filename", true);
  }
  void execute(Slots& s) override {
    if (!s.io) return;
    s.flush();
    const int n[3] = {s.nodes[0], s.nodes[1], s.nodes[2]};
    const int L[3] = {n[0] + 4, n[1] + 4, n[2] + 4};
    const size_t N = (size_t)n[0] * n[1] * n[2];
    const size_t Nall = (size_t)L[0] * L[1] * L[2];
    std::vector<double> rho(N), u(3 * N);
    std::vector<int32_t> obst(N);
    check(hlbm_get_macro(s.dom, rho.data(), u.data(), HLBM_LAYOUT_NODES), "write_paraview");
    check(hlbm_get_flags(s.dom, obst.data(), HLBM_LAYOUT_NODES), "write_paraview");
    std::vector<double> f;
    if (distributions) {
      f.resize((size_t)19 * N);
      check(hlbm_get_populations_f64(s.dom, f.data(), HLBM_LAYOUT_NODES), "write_paraview");
    }
    // the untouched ghost state: f_i = value * w_i, rho summed in direction order
    static const int WDEN[19] = {3, 18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36};
    double fg[19], rho_g = 0.0;
    for (int i = 0; i < 19; i++) {
      fg[i] = s.set_dist_value * (1.0 / WDEN[i]);
      rho_g += fg[i];
    }
    // All-box node -> Real node (or -1 for a ghost of a non-periodic axis)
    auto image = [&](size_t k) -> long {
      long c[3] = {(long)(k % L[0]), (long)((k / L[0]) % L[1]), (long)(k / ((size_t)L[0] * L[1]))};
      for (int d = 0; d < 3; d++) {
        long r = c[d] - 2;
        if (r < 0 || r >= n[d]) {
          if (!s.periodic[d]) return -1;
          r = (r + n[d]) % n[d];
        }
        c[d] = r;
      }
      return (c[2] * n[1] + c[1]) * n[0] + c[0];
    };
    const double dx = s.dx;
    const double pc = 1. / 3. * s.params.celerity * s.params.celerity * s.params.avg_rho;
    const double uc = dx / s.params.dtLB;

    const std::string basedir = output_dir(s, outdir) + "/ParaviewOutput/";
    const std::string file_name = format_step(filename, s.timestep);
    const std::string fullname = basedir + file_name;
    std::filesystem::create_directories(fullname);
    if (display_filename && !s.quiet) {
      std::printf("writing paraview file: %s\n", fullname.c_str());
      std::fflush(stdout);
    }
    // global All box of a single rank: [-2, n+1] on every axis
    const int gs[3] = {-2, -2, -2}, ge[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
    {
      std::string t;
      t += " <VTKFile type=\"PRectilinearGrid\"> \n";
      t += "   <PRectilinearGrid WholeExtent=\"0 " + std::to_string(n[0] - 1) + " 0 " + std::to_string(n[1] - 1) +
           " 0 " + std::to_string(n[2] - 1) + "\"\n";
      t += "                     GhostLevel=\"#\">\n";
      t += "     <Piece Extent=\" ";
      for (int d = 0; d < 3; d++) {
        t += std::to_string(gs[d]) + " " + std::to_string(ge[d]);
        t += d < 2 ? " " : "";
      }
      t += "\" Source=\"" + file_name + "/000000.vtr\"/>\n";
      t += "    <PCoordinates>\n";
      t += "      <PDataArray type=\"Float32\" Name=\"X\"/>\n";
      t += "      <PDataArray type=\"Float32\" Name=\"Y\"/>\n";
      t += "      <PDataArray type=\"Float32\" Name=\"Z\"/>\n";
      t += "    </PCoordinates>\n";
      t += "     <PPointData Scalars=\"P OBST\"  Vectors=\"U\" >\n";
      t += "       <PDataArray Name=\"P\" type=\"Float32\" NumberOfComponents=\"1\"/>\n";
      t += "       <PDataArray Name=\"OBST\" type=\"Float32\" NumberOfComponents=\"1\"/>\n";
      t += "       <PDataArray Name=\"U\" type=\"Float32\" NumberOfComponents=\"3\"/>\n";
      if (distributions) t += "       <PDataArray Name=\"Fi\" type=\"Float32\" NumberOfComponents=\"19\"/>\n";
      t += "     </PPointData> \n";
      t += "   </PRectilinearGrid>\n";
      t += " </VTKFile>\n";
      write_text(basedir + "/" + file_name + ".pvtr", t);
    }
    std::string t;
    t += "<VTKFile type=\"RectilinearGrid\">\n";
    t += " <RectilinearGrid WholeExtent=\" 0 " + std::to_string(n[0] - 1) + " 0 " + std::to_string(n[1] - 1) + " 0 " +
         std::to_string(n[2] - 1) + "\">\n";
    t += "      <Piece Extent=\"";
    for (int d = 0; d < 3; d++) t += std::to_string(gs[d]) + " " + std::to_string(ge[d]) + " ";
    t += "\">\n";
    t += "      <Coordinates>\n";
    const char* axis[3] = {"X", "Y", "Z"};
    for (int d = 0; d < 3; d++) {
      t += std::string("          <DataArray type=\"Float32\" Name=\"") + axis[d] + "\" format=\"ascii\">\n";
      for (int x = gs[d]; x <= ge[d]; x++) {
        put_g(t, (double)(x * dx));
        t += ' ';
      }
      t += "\n          </DataArray>\n";
    }
    t += "      </Coordinates>\n";
    t += "      <PointData>\n";
    t += "          <DataArray type=\"Float32\" Name=\"P\" format=\"ascii\">\n";
    t += parallel_text(Nall, [&](size_t k, std::string& o) {
      const long r = image(k);
      const bool fluid = r < 0 || obst[r] == -1;
      put_g(o, fluid ? pc * ((r < 0 ? rho_g : rho[r]) - 1) : 0.0);
      o += ' ';
    });
    t += "\n          </DataArray>\n";
    t += "          <DataArray type=\"Float32\" Name=\"U\" format=\"ascii\" NumberOfComponents=\"3\">\n";
    t += parallel_text(Nall, [&](size_t k, std::string& o) {
      const long r = image(k);
      for (int c = 0; c < 3; c++) {
        const double v = (r >= 0 && obst[r] == -1) ? uc * u[(size_t)c * N + r] : 0.0;
        put_g(o, (double)(float)v);
        o += ' ';
      }
    });
    t += "\n          </DataArray>\n";
    t += "          <DataArray type=\"Float32\" Name=\"OBST\" format=\"ascii\">\n";
    t += parallel_text(Nall, [&](size_t k, std::string& o) {
      const long r = image(k);
      o += r >= 0 && obst[r] != -1 ? "-2 " : "-1 ";
    });
    t += "\n          </DataArray>\n";
    if (distributions) {
      t += "          <DataArray type=\"Float32\" Name=\"Fi\" format=\"ascii\" NumberOfComponents=\"19\">\n";
      t += parallel_text(Nall, [&](size_t k, std::string& o) {
        const long r = image(k);
        for (int i = 0; i < 19; i++) {
          put_g(o, (double)(float)(r >= 0 ? f[(size_t)i * N + r] : fg[i]));
          o += ' ';
        }
      });
      t += "\n          </DataArray>\n";
    }
    t += "      </PointData>\n";
    t += "      </Piece>\n";
    t += " </RectilinearGrid>\n";
    t += "</VTKFile>\n";
    write_text(fullname + "/000000.vtr", t);
  }
};

// ---- plane_velocity_profile (src/check/plane_velocity_profile.cu:157-262) ----
struct PlaneVelocityProfile : OperatorNode {
  std::string dimension, dump_file = "profile_%010d.csv", outdir;
  bool display_filename = true;
  explicit PlaneVelocityProfile(const msp::Node& a) {
    if (!a.find("dimension")) throw OperatorError("plane_velocity_profile: the dimension slot is required");
    dimension = slot_string(a, "dimension", "");
    dump_file = slot_string(a, "dump_file", dump_file);
    outdir = slot_string(a, "output_directory", "");
    display_filename = slot_bool(a, "display_filename", true);
  }
  void execute(Slots& s) override {
    int dim;
    if (dimension == "X") dim = 0;
    else if (dimension == "Y") dim = 1;
    else if (dimension == "Z") dim = 2;
    else {
      if (!s.quiet) {
        std::printf("[plane_velocity_profile] Please, select a valid dimension \"X\", \"Y\", or \"Z\".\n");
        std::printf("[plane_velocity_profile] is skipped\n");
      }
      return;
    }
    if (!s.io) return;
    s.flush();
    const int n = s.nodes[dim];
    std::vector<double> sum(n), vmin(n), vmax(n);
    std::vector<int32_t> cnt(n);
    check(hlbm_plane_profile(s.dom, dim, s.dx / s.params.dtLB, sum.data(), cnt.data(), vmin.data(), vmax.data()),
          "plane_velocity_profile");
    const std::string dir = output_dir(s, outdir) + "/Profile/";
    std::filesystem::create_directories(dir);
    const std::string file = format_step(dir + dump_file, s.timestep);
    if (display_filename && !s.quiet) {
      std::printf("writing profile file: %s\n", file.c_str());
      std::fflush(stdout);
    }
    std::string t;
    char b[128];
    std::snprintf(b, sizeof b, "%16s%16s%16s%16s\n", "position", "avg", "min", "max");
    t += b;
    const double origin = s.bounds[0][dim];
    for (int i = 0; i < n; i++) {
      const double mean = cnt[i] > 0 ? sum[i] / cnt[i] : 0.0;
      std::snprintf(b, sizeof b, "%16.8g%16.8g%16.8g%16.8g\n", origin + i * s.dx, mean, vmin[i], vmax[i]);
      t += b;
    }
    write_text(file, t);
  }
};

// ---- plot_line_velocity (src/check/plot_over_line.cu:125-215) ---------------
struct PlotLineVelocity : OperatorNode {
  std::string dump_file = "line_%010d.csv", outdir;
  bool display_filename = true, has_line = false, has_lbm = false;
  double line[6] = {0, 0, 0, 0, 0, 0};
  long lbm[6] = {0, 0, 0, 0, 0, 0};
  explicit PlotLineVelocity(const msp::Node& a) {
    dump_file = slot_string(a, "dump_file", dump_file);
    outdir = slot_string(a, "output_directory", "");
    display_filename = slot_bool(a, "display_filename", true);
    if (const msp::Node* l = a.find("line")) {
      has_line = true;
      // onika AABB: [[x0,y0,z0],[x1,y1,z1]] or {min: [...], max: [...]}
      std::vector<double> lo, hi;
      if (l->is_map()) {
        lo = l->at("min").numbers();
        hi = l->at("max").numbers();
      } else if (l->is_seq() && l->items.size() == 2) {
        lo = l->items[0].numbers();
        hi = l->items[1].numbers();
      }
      if (lo.size() != 3 || hi.size() != 3) throw OperatorError("plot_line_velocity: line must be [[x,y,z],[x,y,z]]");
      for (int d = 0; d < 3; d++) line[d] = lo[d], line[3 + d] = hi[d];
    }
    if (const msp::Node* l = a.find("line_lbm")) {
      has_lbm = true;
      if (!l->is_seq() || l->items.size() != 2) throw OperatorError("plot_line_velocity: line_lbm must be [[i,j,k],[i,j,k]]");
      const std::vector<double> lo = l->items[0].numbers(), hi = l->items[1].numbers();
      if (lo.size() != 3 || hi.size() != 3) throw OperatorError("plot_line_velocity: line_lbm must be [[i,j,k],[i,j,k]]");
      for (int d = 0; d < 3; d++) lbm[d] = (long)lo[d], lbm[3 + d] = (long)hi[d];
    }
  }
  void execute(Slots& s) override {
    long g[6];
    if (has_line && has_lbm) {
      if (!s.quiet) std::printf("You can't define both slots: line and line_lbm\n");
      return;
    } else if (has_lbm) {
      for (int k = 0; k < 6; k++) g[k] = lbm[k];
    } else if (has_line) {
      // LBMGrid::project_to_grid<Global> (grid.hpp:332-341): floor((r - bmin) / dx)
      for (int k = 0; k < 6; k++) g[k] = (long)std::floor((line[k] - s.bounds[0][k % 3]) / s.dx);
    } else {
      if (!s.quiet) std::printf("Error [plot_line_velocity], you need to specify slot line or line_lbm\n");
      return;
    }
    // PlotLine::resize (plot_over_line.cu:53-72): one axis, end > start
    int count = 0, axis = 0;
    long size = 0;
    for (int d = 0; d < 3; d++) {
      if (g[d] != g[3 + d]) {
        count++;
        const long diff = g[3 + d] - g[d];
        if (diff <= 0) throw OperatorError("PlotLine::resize: end point must be greater than start point");
        size = diff + 1;
        axis = d;
      }
    }
    if (count != 1)
      throw OperatorError("PlotLine::resize: start and end points must be aligned along a single dimension");
    if (!s.io) return;
    s.flush();
    std::vector<double> u((size_t)size * 3);
    const int32_t start[3] = {(int32_t)g[0], (int32_t)g[1], (int32_t)g[2]};
    check(hlbm_line_velocity(s.dom, start, axis, (int32_t)size, s.dx / s.params.dtLB, u.data()),
          "plot_line_velocity");
    const std::string dir = output_dir(s, outdir) + "/Profile/";
    std::filesystem::create_directories(dir);
    const std::string file = format_step(dir + dump_file, s.timestep);
    if (display_filename && !s.quiet) {
      std::printf("writing profile file: %s\n", file.c_str());
      std::fflush(stdout);
    }
    std::string t;
    char b[160];
    std::snprintf(b, sizeof b, "%16s%16s%16s%16s%16s%16s\n", "rx", "ry", "rz", "ux", "uy", "uz");
    t += b;
    for (long i = 0; i < size; i++) {
      double r[3];
      for (int d = 0; d < 3; d++) r[d] = double(g[d] + (d == axis ? i : 0)) * s.dx;
      std::snprintf(b, sizeof b, "%16.8g%16.8g%16.8g%16.8g%16.8g%16.8g\n", r[0], r[1], r[2], u[3 * i], u[3 * i + 1],
                    u[3 * i + 2]);
      t += b;
    }
    write_text(file, t);
  }
};

template <class T>
OperatorFactoryFn with_args() {
  return [](const msp::Node& a) { return std::unique_ptr<OperatorNode>(new T(a)); };
}

struct IoRegistrar {
  IoRegistrar() {
    auto* f = OperatorNodeFactory::instance();
    f->register_factory("write_paraview", with_args<WriteParaview>());
    f->register_factory("plane_velocity_profile", with_args<PlaneVelocityProfile>());
    f->register_factory("plot_line_velocity", with_args<PlotLineVelocity>());
  }
};
const IoRegistrar io_registrar;

}  // namespace

}  // namespace hlbm_host
