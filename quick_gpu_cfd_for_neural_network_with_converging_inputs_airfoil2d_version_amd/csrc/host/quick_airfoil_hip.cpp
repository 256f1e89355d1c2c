// quick_airfoil_hip.cpp -- the solver executable of the airfoil pipeline on the MI355X backend.
//
// Same command line, settings file and outputs as the reference's quick_airfoil
// (src/quick_airfoil.cpp:14-123): --settings FILE --geometry FILE --patches FILE --out PREFIX;
// settings are "key value" lines plus "box x0 x1 y0 y1 level" lines (the keys written by
// pipeline/configuration.py:143-167).  The solve runs on the device through libamr_hip.so:
// mesh, viscous tables and the geometric audit on the host, then one device session driven by
// amr_hip_solve (steady check loop with the on-device monitor), so no state is downloaded per
// check.  Outputs: PREFIX.history.csv, PREFIX.result.json, PREFIX.gpu.json, the mesh arrays
// (PREFIX.cells/faces/walls.npy, PREFIX.mesh.json) and the initial / final snapshots
// (PREFIX.initial.state.npy + .json, PREFIX.final.state.npy + .json; PREFIX.failed_last.* after a
// failure) with their viscous parts (wall_stress.npy, viscous_loads.json, sa.npy), in the
// reference's column orders.
// Accepted backends: "hip" (this executable is the GPU backend; the reference's own CPU build
// remains the "cpu" backend).  Extra optional keys: model (laminar | euler; the reference
// executable is always laminar) and precision (64 | 32).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "amr_hip.h"
#include "mesher.h"
#include "views.h"

using namespace cfd;

namespace {

using Clock = std::chrono::steady_clock;

// NPY v1, little-endian <f8, C order, rows x columns; header padded to a multiple of 64 bytes.
bool write_npy(const std::string& path, const double* values, size_t rows, size_t columns) {
  std::string header = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + std::to_string(rows) + ", " +
                       std::to_string(columns) + "), }";
  const size_t preamble = 10;
  while ((preamble + header.size() + 1) % 64) header += ' ';
  header += '\n';
  std::ofstream out(path, std::ios::binary);
  if (!out) return false;
  const unsigned char magic[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, static_cast<unsigned char>(header.size() & 255),
                                   static_cast<unsigned char>(header.size() >> 8)};
  out.write(reinterpret_cast<const char*>(magic), 10);
  out.write(header.data(), static_cast<std::streamsize>(header.size()));
  if (rows && columns) out.write(reinterpret_cast<const char*>(values), static_cast<std::streamsize>(rows * columns * sizeof(double)));
  return static_cast<bool>(out);
}

bool write_mesh_files(const StaticAmrMesh2D& m, const Params2D& p, const std::string& prefix) {
  std::vector<double> cells, faces, walls;
  for (const AmrCell2D& c : m.cells)
    for (double v : {double(c.ix), double(c.iy), double(c.span), double(c.level), c.x0, c.x1, c.y0, c.y1, c.volume, c.cx,
                     c.cy, double(c.wall_begin), double(c.wall_count)})
      cells.push_back(v);
  for (const AmrFace2D& f : m.faces)
    for (double v : {double(f.left), double(f.right), double(f.axis), double(f.boundary), f.coordinate, f.lo, f.hi,
                     f.length, f.centroid})
      faces.push_back(v);
  for (const WallFragment2D& w : m.walls)
    for (double v : {double(w.owner_cell), double(w.source_edge_id), double(w.body_id), double(w.patch_id), w.x0, w.y0,
                     w.x1, w.y1, w.nx, w.ny, w.length})
      walls.push_back(v);
  if (!write_npy(prefix + ".cells.npy", cells.data(), m.cells.size(), 13) ||
      !write_npy(prefix + ".faces.npy", faces.data(), m.faces.size(), 9) ||
      !write_npy(prefix + ".walls.npy", walls.data(), m.walls.size(), 11))
    return false;
  // Agglomeration members (old id, new id, x0, x1, y0, y1, V, cx, cy) of an agglomerated mesh
  // (reference src/static_amr2d_binary.cpp:38-39).
  if (!m.agglomeration.empty()) {
    std::vector<double> members;
    for (const auto& row : m.agglomeration) members.insert(members.end(), row.begin(), row.end());
    if (!write_npy(prefix + ".agglomeration.npy", members.data(), m.agglomeration.size(), 9)) return false;
  }
  const StaticAmrOptions2D& o = m.options;
  const bool sa = m.transport.sa_model != 0;
  std::ofstream out(prefix + ".mesh.json");
  out << std::setprecision(17) << "{\n\"format\":\"NPY v1 <f8, C-order\",\n\"model\":\""
      << (m.transport.model == AMR_MODEL_EULER ? "2D Euler, static slip wall"
                                               : (sa ? "2D SA-neg mechanical viscous, kappa=0, stationary no-slip wall"
                                                     : "2D laminar mechanical viscous, kappa=0, stationary no-slip wall"))
      << (o.agglomerate_small ? ", agglomerated FV, no SRD, SSP-RK3\",\n" : ", zero-degree SRD, SSP-RK3\",\n")
      << "\"cells_columns\":[\"ix\",\"iy\",\"span\",\"level\",\"x0\",\"x1\",\"y0\",\"y1\",\"fluid_area\",\"cx\",\"cy\","
         "\"wall_begin\",\"wall_count\"],\n"
      << "\"faces_columns\":[\"left\",\"right\",\"axis\",\"boundary\",\"coordinate\",\"lo\",\"hi\",\"open_length\","
         "\"open_centroid\"],\n"
      << "\"walls_columns\":[\"owner\",\"source_edge\",\"body\",\"patch\",\"x0\",\"y0\",\"x1\",\"y1\",\"nx\",\"ny\","
         "\"length\"],\n"
      << "\"state_shape\":[" << (sa ? 5 : 4) << "," << m.cells.size() << "],\n\"state_rows\":[\"rho\",\"rho_u\","
      << "\"rho_v\",\"rho_E\"" << (sa ? ",\"rho_nutilde\"" : "") << "],\n"
      << "\"bounds\":[" << o.xmin << ',' << o.xmax << ',' << o.ymin << ',' << o.ymax << "],\n\"base_shape\":[" << o.nx
      << ',' << o.ny << "],\n\"levels\":" << o.levels << ",\n\"buffer\":" << o.buffer << ",\n\"gamma\":" << p.gamma
      << ",\n\"cfl\":" << p.cfl << ",\n\"fixed_dt\":" << p.dt_fixed << ",\n\"flux\":\""
      << (p.flux == FLUX_HLLC ? "hllc" : "ausm") << "\",\n\"bc\":[";
  for (int side = 0; side < 4; ++side)
    out << (side ? "," : "") << (m.boundaries[side].kind < 0 ? p.bc[side] : m.boundaries[side].kind);
  out << "],\n\"coarse_fine_subfaces\":" << m.coarse_fine_faces << ",\n\"small_cells\":" << m.small_cells
      << ",\n\"split_background_leaves\":" << m.split_leaves << ",\n\"redistribution_mode\":\""
      << (o.agglomerate_small ? "agglomerated-no-SRD" : "state-SRD") << "\"\n}\n";
  return static_cast<bool>(out);
}

bool write_viscous_snapshot(const StaticAmrMesh2D& m, const Params2D& p, const StaticAmrState2D& s,
                            const std::string& prefix);

// State snapshot: PREFIX.state.npy (nv x n) and PREFIX.json with bounds, pressure loads and
// conserved totals (the reference's snapshot summary fields), plus the viscous parts.
bool write_snapshot(const StaticAmrMesh2D& m, const Params2D& p, const StaticAmrState2D& s, const std::string& prefix,
                    std::string* error) {
  const size_t n = m.cells.size(), nv = m.transport.sa_model ? 5 : 4;
  if (s.u.size() != nv * n) {
    *error = "snapshot shape mismatch";
    return false;
  }
  double rhomin = INFINITY, rhomax = 0, pmin = INFINITY, pmax = 0, fx = 0, fy = 0, mz = 0;
  long double totals[4] = {0, 0, 0, 0};
  for (size_t k = 0; k < n; ++k) {
    const double rho = s.u[k], vx = s.u[n + k] / rho, vy = s.u[2 * n + k] / rho;
    const double pressure = (p.gamma - 1) * (s.u[3 * n + k] - .5 * rho * (vx * vx + vy * vy));
    if (!std::isfinite(rho) || !(rho > 0) || !std::isfinite(pressure) || !(pressure > 0)) {
      *error = "nonphysical snapshot; no file certified";
      return false;
    }
    rhomin = std::min(rhomin, rho);
    rhomax = std::max(rhomax, rho);
    pmin = std::min(pmin, pressure);
    pmax = std::max(pmax, pressure);
    for (int v = 0; v < 4; ++v) totals[v] += static_cast<long double>(m.cells[k].volume) * s.u[v * n + k];
    const AmrCell2D& c = m.cells[k];
    for (size_t q = c.wall_begin; q < c.wall_begin + c.wall_count; ++q) {
      const WallFragment2D& w = m.walls[q];
      const double x = pressure * w.nx * w.length, y = pressure * w.ny * w.length;
      fx += x;
      fy += y;
      mz += .5 * (w.x0 + w.x1) * y - .5 * (w.y0 + w.y1) * x;
    }
  }
  if (!write_npy(prefix + ".state.npy", s.u.data(), nv, n)) {
    *error = "NPY write failed: " + prefix + ".state.npy";
    return false;
  }
  if (!write_viscous_snapshot(m, p, s, prefix)) {
    *error = "viscous snapshot write failed: " + prefix;
    return false;
  }
  std::ofstream out(prefix + ".json");
  out << std::setprecision(17) << "{\"backend\":\"hip\",\"time\":" << s.time << ",\"steps\":" << s.steps
      << ",\"rejected\":" << s.rejected << ",\"rho_min\":" << rhomin << ",\"rho_max\":" << rhomax << ",\"p_min\":" << pmin
      << ",\"p_max\":" << pmax << ",\"fx_pressure\":" << fx << ",\"fy_pressure\":" << fy
      << ",\"mz_pressure_about_origin\":" << mz << ",\"integrated_conserved\":[" << static_cast<double>(totals[0]) << ','
      << static_cast<double>(totals[1]) << ',' << static_cast<double>(totals[2]) << ','
      << static_cast<double>(totals[3]) << "],\"state_bytes\":" << nv * n * sizeof(double)
      << ",\"dimensional_units\":false}\n";
  if (!out) {
    *error = "snapshot JSON write failed";
    return false;
  }
  return true;
}

// Molecular viscosity at a temperature (constant or Sutherland), as the device path evaluates it.
double viscosity_of(const AmrTransport2D& t, double temperature) {
  if (t.law == AMR_MU_CONSTANT) return t.mu_ref;
  const double ratio = temperature / t.temperature_ref;
  return t.mu_ref * ratio * std::sqrt(ratio) * (t.temperature_ref + t.sutherland) / (temperature + t.sutherland);
}

// Wall stress and body loads of a viscous snapshot, in FP64 from the snapshot state:
// PREFIX.wall_stress.npy (14 columns per wall fragment), PREFIX.viscous_loads.json (pressure and
// viscous force and moment about the origin, in total and per patch) and, with SA-neg,
// PREFIX.sa.npy (nutilde, eddy viscosity, wall distance per cell) - the files the pipeline's
// surface and load exports read.  The no-slip shear of a fragment comes from its donor cell:
// tau n with the velocity gradient (0 - u_donor) n / d (the shear the time step applies).
bool write_viscous_snapshot(const StaticAmrMesh2D& m, const Params2D& p, const StaticAmrState2D& s,
                            const std::string& prefix) {
  if (m.transport.model == AMR_MODEL_EULER) return true;
  const size_t n = m.cells.size(), nw = m.walls.size();
  const AmrTransport2D& t = m.transport;
  std::vector<double> rho(n), vx(n), vy(n), pr(n), mu(n);
  for (size_t k = 0; k < n; ++k) {
    rho[k] = s.u[k];
    vx[k] = s.u[n + k] / rho[k];
    vy[k] = s.u[2 * n + k] / rho[k];
    pr[k] = (p.gamma - 1) * (s.u[3 * n + k] - .5 * rho[k] * (vx[k] * vx[k] + vy[k] * vy[k]));
    mu[k] = viscosity_of(t, pr[k] / (rho[k] * t.gas_constant));
  }
  std::vector<double> rows;
  rows.reserve(14 * nw);
  std::map<int, std::array<double, 6>> patches;
  std::array<double, 6> total{};
  for (size_t q = 0; q < nw; ++q) {
    const WallFragment2D& w = m.walls[q];
    const AmrWallSample2D& sample = m.viscous.wall_samples[q];
    const size_t k = w.owner_cell, dn = static_cast<size_t>(sample.donor);
    const double dd = sample.distance;
    const double du = -vx[dn] / dd, dv = -vy[dn] / dd;
    const double g0 = w.nx * du, g1 = w.ny * du, g2 = w.nx * dv, g3 = w.ny * dv;
    const double dil = 2.0 / 3.0 * (g0 + g3);
    const double txx = mu[dn] * (2 * g0 - dil), txy = mu[dn] * (g1 + g2), tyy = mu[dn] * (2 * g3 - dil);
    const double flux_x = (txx * w.nx + txy * w.ny) * w.length, flux_y = (txy * w.nx + tyy * w.ny) * w.length;
    const double x = .5 * (w.x0 + w.x1), y = .5 * (w.y0 + w.y1);
    const double px = pr[k] * w.nx * w.length, py = pr[k] * w.ny * w.length, fvx = -flux_x, fvy = -flux_y;
    const std::array<double, 6> loads = {px, py, x * py - y * px, fvx, fvy, x * fvy - y * fvx};
    std::array<double, 6>& patch = patches[w.patch_id];
    for (int v = 0; v < 6; ++v) {
      patch[v] += loads[v];
      total[v] += loads[v];
    }
    for (double value : {double(q), double(k), double(dn), dd, mu[dn], flux_x / w.length, flux_y / w.length,
                         (-w.ny * flux_x + w.nx * flux_y) / w.length, pr[k], px, py, fvx, fvy, loads[5]})
      rows.push_back(value);
  }
  if (!write_npy(prefix + ".wall_stress.npy", rows.data(), nw, 14)) return false;
  std::ofstream f(prefix + ".viscous_loads.json");
  f << std::setprecision(17) << "{\"time\":" << s.time << ",\"model\":\""
    << (t.sa_model ? "sa-neg-mechanical" : "laminar")
    << "\",\"conductivity\":0,\"load_columns\":[\"fx_pressure\",\"fy_pressure\",\"mz_pressure\",\"fx_viscous\","
       "\"fy_viscous\",\"mz_viscous\"],\"total\":[";
  for (int v = 0; v < 6; ++v) f << (v ? "," : "") << total[v];
  f << "],\"patches\":[";
  bool first = true;
  for (const auto& [id, value] : patches) {
    f << (first ? "" : ",") << "{\"patch\":" << id << ",\"loads\":[";
    for (int v = 0; v < 6; ++v) f << (v ? "," : "") << value[v];
    f << "]}";
    first = false;
  }
  f << "],\"wall_stress_columns\":[\"wall_id\",\"owner\",\"donor\",\"normal_distance\",\"mu\",\"tau_n_x\","
       "\"tau_n_y\",\"tau_t_signed\",\"p\",\"fx_pressure\",\"fy_pressure\",\"fx_viscous\",\"fy_viscous\","
       "\"mz_viscous\"],\"moment_reference\":[0,0],\"wall_normal\":\"fluid to solid\",\"wall_energy_flux\":0}\n";
  if (!f) return false;
  if (t.sa_model) {
    std::vector<double> sa;
    sa.reserve(3 * n);
    for (size_t k = 0; k < n; ++k) {
      const double phi = s.u[4 * n + k] / rho[k], nu = mu[k] / rho[k];
      const double chi = phi / nu, cube = chi * chi * chi;
      const double eddy = phi > 0 ? rho[k] * phi * cube / (cube + 7.1 * 7.1 * 7.1) : 0.0;
      sa.push_back(phi);
      sa.push_back(eddy);
      sa.push_back(m.viscous.wall_distance[k]);
    }
    if (!write_npy(prefix + ".sa.npy", sa.data(), n, 3)) return false;
    std::ofstream d(prefix + ".sa.json");
    d << std::setprecision(17) << "{\"model\":\"SA-neg-ft2-Note1c-mechanical\",\"columns\":[\"nutilde\",\"mu_t\","
      << "\"wall_distance\"],\"nu_inf\":" << t.sa_nu_inf << ",\"wall_nutilde\":0,\"thermal_conductivity\":0}\n";
    if (!d) return false;
  }
  return true;
}

bool write_stats(const amr_gpu_stats& s, const std::string& path, double io) {
  std::ofstream out(path);
  out << std::setprecision(17) << "{\n\"allocated_device_bytes\":" << s.allocated_bytes << ",\n\"pinned_host_bytes\":"
      << s.pinned_bytes << ",\n\"geometry_h2d_bytes\":" << s.geometry_h2d_bytes << ",\n\"state_h2d_bytes\":"
      << s.state_h2d_bytes << ",\n\"control_h2d_bytes\":" << s.control_h2d_bytes << ",\n\"control_d2h_bytes\":"
      << s.control_d2h_bytes << ",\n\"state_d2h_bytes\":" << s.state_d2h_bytes << ",\n\"status_polls\":" << s.status_polls
      << ",\n\"full_downloads\":" << s.full_downloads << ",\n\"setup_seconds\":" << s.setup_seconds
      << ",\n\"advance_wall_seconds\":" << s.advance_seconds << ",\n\"download_wall_seconds\":" << s.download_seconds
      << ",\n\"binary_io_seconds\":" << io << ",\n\"advance_hip_event_seconds\":" << s.advance_event_seconds
      << ",\n\"event_batches\":" << s.event_batches << ",\n\"graph_nodes_per_trial\":" << s.graph_nodes_per_trial
      << ",\n\"event_scope\":\"sum of timed trial-batch stream intervals; excludes control/state transfers\""
      << ",\n\"transfer_counter_scope\":\"explicit hipMemcpy only; kernel arguments excluded\",\n\"overlap_claim\":false\n}\n";
  return static_cast<bool>(out);
}

const char* reason_name(int r) {
  switch (r) {
    case AMR_STOP_CONVERGED: return "raw-residual-and-three-coefficients";
    case AMR_STOP_BUDGET: return "time-budget";
    case AMR_STOP_TIME_CAP: return "physical-time-cap";
    case AMR_STOP_FAILURE: return "admissibility-or-advance-failure";
    default: return "short-validation";
  }
}

int run(int argc, char** argv) {
  if (argc != 9) {
    std::cerr << "usage: quick_airfoil_hip --settings FILE --geometry FILE --patches FILE --out PREFIX\n";
    return 2;
  }
  std::map<std::string, std::string> args;
  for (int i = 1; i < argc; i += 2) args[argv[i]] = argv[i + 1];
  for (const char* key : {"--settings", "--geometry", "--patches", "--out"})
    if (!args.count(key)) return 2;
  std::map<std::string, std::string> cfg;
  std::vector<AmrBox2D> boxes;
  std::ifstream settings(args["--settings"]);
  if (!settings) return 2;
  for (std::string line; std::getline(settings, line);) {
    std::istringstream row(line);
    std::string key, value;
    if (!(row >> key)) continue;
    if (key == "box") {
      AmrBox2D b;
      if (!(row >> b.x0 >> b.x1 >> b.y0 >> b.y1 >> b.level)) return 2;
      boxes.push_back(b);
    } else {
      if (!(row >> value)) return 2;
      cfg[key] = value;
    }
  }
  const auto number = [&](const char* key) { return std::stod(cfg.at(key)); };
  const auto option = [&](const char* key, const std::string& fallback) {
    return cfg.count(key) ? cfg.at(key) : fallback;
  };
  const std::string name = cfg.at("name"), prefix = args["--out"], mode = cfg.at("mode"), backend = cfg.at("backend");
  const std::string mesh_preset = cfg.at("mesh_preset"), viscosity = cfg.at("viscosity");
  const std::string turbulence = cfg.at("turbulence"), sa_layout = cfg.at("sa_layout");
  const std::string model = option("model", "laminar");
  const int precision = std::stoi(option("precision", "64"));
  const double stop_time = number("stop_time"), budget = number("budget"), mu = number("mu_star");
  const double mach = number("mach"), aoa = number("aoa"), gamma = number("gamma");
  const bool events = number("events") != 0;
  if (backend != "hip") {
    std::cerr << "backend '" << backend << "' is not served by this executable (it is the hip backend)\n";
    return 2;
  }
  if ((mode != "steady" && mode != "time" && mode != "mesh" && mode != "smoke" && mode != "rest") ||
      (viscosity != "constant" && viscosity != "sutherland") || (turbulence != "none" && turbulence != "sa-neg") ||
      (model != "laminar" && model != "euler") || (precision != 64 && precision != 32) || !std::isfinite(stop_time) ||
      stop_time <= 0 || budget < 2 || !std::isfinite(mu) || mu <= 0 || !(mach > 0) || !(gamma > 1))
    return 2;
  const std::filesystem::path parent = std::filesystem::path(prefix).parent_path();
  if (!parent.empty()) std::filesystem::create_directories(parent);
  const auto begin = Clock::now();
  const auto elapsed = [&]() { return std::chrono::duration<double>(Clock::now() - begin).count(); };

  std::vector<double> x, y;
  std::vector<int> patches;
  std::ifstream geo(args["--geometry"]);
  for (std::string line; std::getline(geo, line);) {
    std::replace(line.begin(), line.end(), ',', ' ');
    std::istringstream row(line);
    double a, b;
    if (row >> a >> b) {
      x.push_back(a);
      y.push_back(b);
    }
  }
  std::ifstream pat(args["--patches"]);
  for (std::string line; std::getline(pat, line);) {
    std::replace(line.begin(), line.end(), ',', ' ');
    std::istringstream row(line);
    int a, b;
    if (row >> a >> b) {
      if (a != static_cast<int>(patches.size())) return 2;
      patches.push_back(b);
    }
  }
  if (x.size() < 3 || x.size() != patches.size()) return 2;

  StaticAmrOptions2D options;
  options.nx = static_cast<int>(number("nx"));
  options.ny = static_cast<int>(number("ny"));
  options.levels = static_cast<int>(number("levels"));
  options.buffer = static_cast<int>(number("buffer"));
  options.xmin = number("xmin");
  options.xmax = number("xmax");
  options.ymin = number("ymin");
  options.ymax = number("ymax");
  options.max_cells = static_cast<size_t>(number("max_cells"));
  options.agglomerate_small = true;
  options.boxes = boxes;
  const double wall_dx = (options.xmax - options.xmin) / options.nx / (1 << options.levels);
  const double wall_dy = (options.ymax - options.ymin) / options.ny / (1 << options.levels);
  Params2D p = hipamr::params_from_c(amr_params{});
  p.gamma = gamma;
  p.cfl = number("cfl");
  p.dt_fixed = number("fixed_dt");
  p.limiter = LIM_FIRST_ORDER;
  p.time_integrator = TIME_SSP_RK3;
  p.flux = cfg.at("flux") == "hllc" ? FLUX_HLLC : FLUX_AUSMPUP;
  p.ausm_mco = 1.0;
  for (int side = 0; side < 4; ++side) p.bc[side] = BC_TRANSMISSIVE;
  const double alpha = aoa * std::acos(-1.) / 180, cs = std::cos(alpha), sn = std::sin(alpha);
  const std::array<double, 4> reference = {1, mode == "rest" ? 0 : mach * cs, mode == "rest" ? 0 : mach * sn, 1 / gamma};

  StaticAmrMesh2D mesh;
  std::string error;
  if (hipamr::build_mesh(options, x, y, patches, &mesh, &error)) {
    std::cerr << "geometry: " << error << '\n';
    return 1;
  }
  for (int side = 0; side < 4; ++side) {
    AmrBoundary2D& b = mesh.boundaries[side];
    std::copy(reference.begin(), reference.end(), b.reference);
    const std::string kind = cfg.at("bc" + std::to_string(side));
    b.kind = kind == "pressure-outlet" ? AMR_BC_PRESSURE_OUTLET : kind == "transmissive" ? AMR_BC_TRANSMISSIVE : AMR_BC_FARFIELD;
    b.outlet_pressure = number("outlet_p_star");
  }
  AmrTransport2D& t = mesh.transport;
  const bool viscous = model == "laminar";
  t.model = viscous ? AMR_MODEL_LAMINAR : AMR_MODEL_EULER;
  t.sa_model = viscous && turbulence == "sa-neg";
  t.sa_nu_inf = number("sa_nu_inf");
  t.mu_ref = mu;
  if (viscosity == "sutherland") {
    t.law = AMR_MU_SUTHERLAND;
    t.temperature_ref = 1 / gamma;
    t.sutherland = number("sutherland_star");
  }
  t.viscous_cfl = number("viscous_cfl");
  t.sa_cfl = number("sa_cfl");
  t.wall_probe_min_h = number("wall_probe_min_h");
  t.stabilization = AMR_VISCOUS_VOLUME;
  if (!t.sa_model) t.sa_nu_inf = 0;
  if (hipamr::prepare_viscous(&mesh, p, &error)) {
    std::cerr << error;
    return 1;
  }
  // Geometric audit before any flow work: fluid area, wall length, per-cell closure
  // (src/quick_airfoil.cpp:51-58).
  long double solid = 0, perimeter = 0, area = 0, wall_length = 0;
  const size_t n = mesh.cells.size();
  std::vector<std::array<double, 2>> closure(n, {0.0, 0.0});
  std::vector<double> surface(n, 0.0);
  for (size_t i = 0; i < x.size(); ++i) {
    const size_t j = (i + 1) % x.size();
    solid += (static_cast<long double>(x[i]) * y[j] - static_cast<long double>(y[i]) * x[j]) * .5;
    perimeter += std::hypot(x[i] - x[j], y[i] - y[j]);
  }
  for (const AmrCell2D& c : mesh.cells) {
    if (!(c.volume > 0)) return 1;
    area += c.volume;
  }
  for (const AmrFace2D& f : mesh.faces) {
    if (f.left >= 0) {
      closure[f.left][f.axis] += f.length;
      surface[f.left] += f.length;
    }
    if (f.right >= 0) {
      closure[f.right][f.axis] -= f.length;
      surface[f.right] += f.length;
    }
  }
  for (const WallFragment2D& w : mesh.walls) {
    wall_length += w.length;
    closure[w.owner_cell][0] += w.nx * w.length;
    closure[w.owner_cell][1] += w.ny * w.length;
    surface[w.owner_cell] += w.length;
  }
  double max_closure = 0;
  for (size_t k = 0; k < n; ++k) max_closure = std::max(max_closure, std::hypot(closure[k][0], closure[k][1]) / surface[k]);
  const double domain = (options.xmax - options.xmin) * (options.ymax - options.ymin);
  const double area_error = std::fabs(static_cast<double>(area - (domain - std::fabs(solid))));
  const double wall_error = std::fabs(static_cast<double>(wall_length - perimeter));
  if (area_error > 1e-10 || wall_error > 1e-11 || max_closure > 1e-9) {
    std::cerr << "geometry audit failed " << area_error << ' ' << wall_error << ' ' << max_closure;
    return 1;
  }
  StaticAmrState2D state;
  if (hipamr::initialize_state(mesh, p, reference, &state)) return 1;
  const std::vector<double> initial = state.u;
  const double mesh_seconds = elapsed();

  double io = 0;
  const auto timed = [&](auto&& fn) {
    const auto t0 = Clock::now();
    const bool ok = fn();
    io += std::chrono::duration<double>(Clock::now() - t0).count();
    return ok;
  };
  if (!timed([&] { return write_mesh_files(mesh, p, prefix); }) ||
      !timed([&] { return write_snapshot(mesh, p, state, prefix + ".initial", &error); })) {
    std::cerr << (error.empty() ? "mesh file write failed" : error) << '\n';
    return 1;
  }

  amr_mesh_view view;
  hipamr::fill_view(mesh, &view);
  const amr_params params = hipamr::params_to_c(p);
  amr_session* session = nullptr;
  amr_gpu_stats stats{};
  amr_steady_result r{};
  r.reason = mode == "mesh" ? -1 : AMR_STOP_BUDGET;
  std::vector<amr_steady_check> history;
  char message[512] = {0};
  if (mode != "mesh") {
    amr_session_options so{};
    so.batch_trials = 32;
    so.graph = 1;
    so.event_timing = events ? 1 : 0;
    so.precision = precision;
    const amr_state_meta meta{state.time, state.steps, state.rejected};
    if (amr_hip_create(&view, &params, state.u.data(), state.u.size(), &meta, &so, &session, message, sizeof(message))) {
      std::cerr << message << '\n';
      return 1;
    }
    amr_steady_options o{};
    o.mach = mach;
    o.aoa_degrees = aoa;
    o.moment[0] = number("moment_x");
    o.moment[1] = number("moment_y");
    o.region[0] = number("res_xmin");
    o.region[1] = number("res_xmax");
    o.region[2] = number("res_ymin");
    o.region[3] = number("res_ymax");
    o.raw_tol = number("raw_tol");
    o.load_tol = number("load_tol");
    o.min_time = number("tmin");
    o.period = number("period");
    o.hold = static_cast<int>(number("hold"));
    o.mode = mode == "steady" ? 0 : mode == "time" ? 1 : 2;
    o.stop_time = stop_time;
    o.budget_seconds = std::max(2.0, budget - elapsed());
    history.resize(static_cast<size_t>(std::ceil(stop_time / o.period)) + 4);
    const int status = amr_hip_solve(session, &o, &r, history.data(), static_cast<int32_t>(history.size()), message,
                                     sizeof(message));
    if (status) std::cerr << "flow: " << message << " rejected=" << r.rejected << '\n';
    amr_state_meta after{};
    if (amr_hip_download(session, state.u.data(), state.u.size(), &after, message, sizeof(message))) {
      std::cerr << message << '\n';
      amr_hip_destroy(session);
      return 1;
    }
    state.time = after.time;
    state.steps = after.steps;
    state.rejected = after.rejected;
    amr_hip_stats(session, &stats);
    amr_hip_destroy(session);
  }
  const bool failed = r.failed != 0;
  std::ofstream hist(prefix + ".history.csv");
  hist << std::setprecision(17)
       << "time,steps,rejected,elapsed,res_rho,res_mx,res_my,res_E,res_sa,Cl,Cd,Cm,delta_Cl,delta_Cd,delta_Cm,hold\n";
  for (int i = 0; i < std::min<int>(r.checks, static_cast<int>(history.size())); ++i) {
    const amr_steady_check& h = history[i];
    hist << h.time << ',' << h.steps << ',' << h.rejected << ',' << h.elapsed + mesh_seconds;
    for (double v : h.residual) hist << ',' << v;
    hist << ',' << h.cl << ',' << h.cd << ',' << h.cm << ',' << h.dcl << ',' << h.dcd << ',' << h.dcm << ',' << h.hold
         << '\n';
    std::cout << "check " << name << " t=" << h.time << " steps=" << h.steps << " raw="
              << *std::max_element(h.residual, h.residual + 5) << " Cl=" << h.cl << " Cd=" << h.cd << " Cm=" << h.cm
              << " hold=" << h.hold << '\n';
  }
  hist.close();
  if (!timed([&] { return write_snapshot(mesh, p, state, prefix + (failed ? ".failed_last" : ".final"), &error); })) {
    std::cerr << error << '\n';
    return 1;
  }
  if (!write_stats(stats, prefix + ".gpu.json", io)) return 1;
  double rest_error = 0;
  if (mode == "rest")
    for (size_t i = 0; i < 4 * n; ++i) rest_error = std::max(rest_error, std::fabs(initial[i] - state.u[i]));
  const int nv = t.sa_model ? 5 : 4;
  std::ofstream out(prefix + ".result.json");
  out << std::setprecision(17) << "{\"case\":" << std::quoted(name) << ",\"mode\":" << std::quoted(mode)
      << ",\"backend\":\"hip\",\"precision\":" << precision << ",\"converged\":" << (r.converged ? "true" : "false")
      << ",\"failed\":" << (failed ? "true" : "false") << ",\"stop_reason\":"
      << std::quoted(mode == "mesh" ? "mesh-only" : reason_name(r.reason)) << ",\"elapsed_seconds\":" << elapsed()
      << ",\"mesh_setup_seconds\":" << mesh_seconds << ",\"cpu_advance_seconds\":0,\"physical_time\":" << state.time
      << ",\"steps\":" << state.steps << ",\"rejected\":" << state.rejected << ",\"fluid_control_volumes\":" << n
      << ",\"background_leaves\":" << mesh.background_leaves << ",\"split_leaves\":" << mesh.split_leaves
      << ",\"sa_layout\":" << std::quoted(sa_layout) << ",\"stop_time_cap\":" << stop_time
      << ",\"turbulence_model\":" << std::quoted(turbulence) << ",\"mesh_preset\":" << std::quoted(mesh_preset)
      << ",\"viscosity_law\":" << std::quoted(viscosity) << ",\"wall_dx\":" << wall_dx << ",\"wall_dy\":" << wall_dy
      << ",\"Mach\":" << mach << ",\"AOA_degrees\":" << aoa << ",\"mu_star\":" << mu << ",\"kappa\":0,\"Cl\":" << r.cl
      << ",\"Cd\":" << r.cd << ",\"Cm\":" << r.cm << ",\"moment_reference\":[" << number("moment_x") << ","
      << number("moment_y") << "],\"moment_sign\":\"CCW-positive\",\"raw_residual\":[";
  for (int v = 0; v < nv; ++v) out << (v ? "," : "") << r.residual[v];
  out << "],\"checks\":" << r.checks << ",\"hold\":" << r.hold << ",\"rest_linf\":" << rest_error
      << ",\"area_error\":" << area_error << ",\"wall_length_error\":" << wall_error
      << ",\"closure_relative_linf\":" << max_closure << "}\n";
  return failed || (mode == "rest" && rest_error > 2e-11) ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::cerr << "configuration/runtime error: " << e.what() << "\n";
    return 2;
  }
}
