// sim_host.cpp -- Config2D/3D and Simulator2D/3D over the C ABI.
//
// Config loading follows include/FSI_Simulator/utils/Config.hpp:29-225 and
// src/utils/Config.cpp:13-89 (keys, defaults, boundary list semantics);
// meshes follow src/io/MeshLoader.cpp (Medit .mesh sections),
// src/fem/DeformableMesh.cpp:159-187, 403-504 (initial transform, boundary
// extraction) and the scene merge of src/fem/FemScene.cpp (surface vertices
// in increasing volume index, meshes concatenated); the step and getters
// follow src/core/Simulator2D.cpp / Simulator3D.cpp and
// src/fem/StaticBodySolver.cpp:122-180 (kinematic control).

#include "sim_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <sys/stat.h>

#include "json_lite.h"

namespace ffsim {

int flag_from_name(const std::string& s) {
  static const std::map<std::string, int> names = {
      {"Wall", 5},         {"InletLeft", 13},   {"InletRight", 14}, {"InletUp", 15},    {"InletDown", 16},
      {"InletFront", 17},  {"InletBack", 18},   {"OutletLeft", 19}, {"OutletRight", 20}, {"OutletUp", 21},
      {"OutletDown", 22},  {"OutletFront", 23}, {"OutletBack", 24}};
  auto it = names.find(s);
  return it == names.end() ? 0 : it->second;
}

// ---------------------------------------------------------------- config
static void read_options(const ffjson::Value& j, FemSolverOptions& o) {
  o.optimizer_type = j.string_or("optimizer_type", o.optimizer_type);
  o.iterations = j.int_or("iterations", o.iterations);
  o.line_search_method = j.string_or("line_search_method", o.line_search_method);
  o.ls_max_iter = j.int_or("ls_max_iter", o.ls_max_iter);
  o.ls_beta = j.number_or("ls_beta", o.ls_beta);
  o.ls_alpha = j.number_or("ls_alpha", o.ls_alpha);
  o.linear_solver_type = j.string_or("linear_solver_type", o.linear_solver_type);
  o.substeps = j.int_or("substeps", o.substeps);
  o.omega = j.number_or("omega", o.omega);
  o.vbd_iterations = j.int_or("vbd_iterations", o.vbd_iterations);
  o.force_density_abs_tol = j.number_or("force_density_abs_tol", o.force_density_abs_tol);
  o.grad_check = j.bool_or("grad_check", o.grad_check);
  o.thread_ct = j.int_or("thread_ct", o.thread_ct);
}

template <size_t N>
static void read_array(const ffjson::Value& j, const char* key, std::array<double, N>& a) {
  if (!j.contains(key)) return;
  std::vector<double> v = j.numbers_or(key, {});
  if (v.size() != N) throw std::runtime_error(std::string("'") + key + "' must have " + std::to_string(N) + " entries");
  for (size_t i = 0; i < N; i++) a[i] = v[i];
}

template <int D>
static void read_solid(const ffjson::Value& j, SolidBodyConfig<D>& c) {
  c.mesh_path = j.string_or("mesh_path", c.mesh_path);
  c.density = j.number_or("density", c.density);
  c.youngs_modulus = j.number_or("youngs_modulus", c.youngs_modulus);
  c.poisson_ratio = j.number_or("poisson_ratio", c.poisson_ratio);
  read_array(j, "translate", c.translate);
  if constexpr (D == 2) {
    c.rotate[0] = j.number_or("rotate", c.rotate[0]);
  } else {
    read_array(j, "rotate", c.rotate);
  }
  read_array(j, "scale", c.scale);
  read_array(j, "initial_velocity", c.initial_velocity);
  if (j.contains("lbs_control_config")) {
    const ffjson::Value& l = j.at("lbs_control_config");
    LBSControlConfig& o = c.lbs_control_config;
    o.cnum = l.int_or("cnum", o.cnum);
    o.lbs_distance_type = l.string_or("lbs_distance_type", o.lbs_distance_type);
    o.random_first = l.bool_or("random_first", o.random_first);
    o.omega = l.number_or("omega", o.omega);
    o.stiffness = l.number_or("stiffness", o.stiffness);
  }
}

template <int D>
bool Config<D>::load(const std::string& filepath) {
  std::ifstream f(filepath);
  if (!f.is_open()) {
    std::cerr << "[Config FATAL] Configuration file not found or unreadable: '" << filepath << "'" << std::endl;
    return false;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  try {
    const ffjson::Value j = ffjson::parse(ss.str());
    SimulationParameters<D> p;
    p.dimension = j.int_or("dimension", D);
    p.fluid_viscosity = j.number_or("fluid_viscosity", p.fluid_viscosity);
    p.fluid_density = j.number_or("fluid_density", p.fluid_density);
    p.fluid_nx = j.int_or("fluid_nx", p.fluid_nx);
    p.fluid_ny = j.int_or("fluid_ny", p.fluid_ny);
    if (D == 3) p.fluid_nz = j.int_or("fluid_nz", p.fluid_nz);
    p.fluid_dx = j.number_or("fluid_dx", 1.0);
    p.solid_solver_type = j.string_or("solid_solver_type", p.solid_solver_type);
    p.total_time = j.number_or("total_time", p.total_time);
    p.dt = j.number_or("dt", p.dt);
    if (D == 2) p.output_frequency = j.int_or("output_frequency", p.output_frequency);  // the 3D reader keeps 100
    p.output_path = j.string_or("output_path", p.output_path);
    p.log_level = j.string_or("log_level", p.log_level);
    p.log_file = j.string_or("log_file", p.log_file);
    p.use_newton = j.bool_or("use_newton", p.use_newton);
    if (D == 2) p.boundary_velocities.assign(4, 0.0f);
    if (j.contains("boundary_velocities")) {
      std::vector<double> bv = j.numbers_or("boundary_velocities", {});
      p.boundary_velocities.assign(bv.begin(), bv.end());
    }
    if (j.contains("global_fem_options") && j.at("global_fem_options").is_object())
      read_options(j.at("global_fem_options"), p.global_fem_options);
    if (j.contains("boundaries") && j.at("boundaries").is_array()) {
      for (const ffjson::Value& b : j.at("boundaries").arr) {
        if (!b.is_object()) {
          std::cerr << "[JSON PARSE WARN] Item in 'boundaries' array is not a JSON object. Skipping." << std::endl;
          continue;
        }
        BoundaryCondition bc;
        for (double x : b.numbers_or("pos", {})) bc.position.push_back((int)x);
        if (!b.contains("pos") || !b.contains("type")) throw std::runtime_error("boundary entry needs 'pos' and 'type'");
        if (!b.at("type").is_string()) throw std::runtime_error("boundary 'type' must be a string");
        bc.flag = flag_from_name(b.at("type").str);
        p.boundaries.push_back(bc);
      }
    }
    if (j.contains("solids") && j.at("solids").is_array()) {
      for (const ffjson::Value& s : j.at("solids").arr) {
        if (!s.is_object()) continue;
        SolidBodyConfig<D> c;
        read_solid<D>(s, c);
        p.solids.push_back(c);
      }
    }
    if (j.contains("freeflow_amd") && j.at("freeflow_amd").is_object()) {
      const ffjson::Value& o = j.at("freeflow_amd");
      p.device_options.storage = o.string_or("storage", p.device_options.storage);
      p.device_options.stage_fp16 = o.bool_or("stage_fp16", p.device_options.stage_fp16);
      p.device_options.periodic = o.bool_or("periodic", p.device_options.periodic);
      p.device_options.fast_math = o.bool_or("fast_math", p.device_options.fast_math);
      p.device_options.device = o.int_or("device", p.device_options.device);
    }
    // postProcess (Config.cpp:72-89): warnings only
    for (const BoundaryCondition& bc : p.boundaries) {
      if ((int)bc.position.size() != D)
        std::cerr << "[Config ERROR] boundary position has " << bc.position.size() << " entries, expected " << D << std::endl;
      if (bc.flag == 0) std::cerr << "[Config WARN] boundary condition with an unrecognized type" << std::endl;
    }
    const size_t nbv = D == 2 ? 4 : 6;
    if (!p.boundary_velocities.empty() && p.boundary_velocities.size() != nbv)
      std::cerr << "[Config WARN] `boundary_velocities` should have " << nbv << " elements" << std::endl;
    m_params = p;
  } catch (const std::exception& e) {
    std::cerr << "[Config FATAL] " << filepath << ": " << e.what() << std::endl;
    return false;
  }
  return true;
}

template class Config<2>;
template class Config<3>;

// ---------------------------------------------------------------- meshes
static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

struct RawMesh {
  int dim = 0;
  std::vector<double> v;
  std::vector<unsigned> edges, tris, tets;
};

// Medit .mesh reader (MeshLoader.cpp:197-257 for 2D, the 3D twin for
// Tetrahedra); indices are 1-based in the file.
static RawMesh load_mesh(const std::string& path, int dim) {
  std::ifstream f(path);
  if (!f.is_open()) throw std::runtime_error("cannot open mesh file '" + path + "'");
  RawMesh m;
  std::string line;
  auto read_block = [&](int per, std::vector<unsigned>* idx, std::vector<double>* pts) {
    if (!std::getline(f, line)) throw std::runtime_error("truncated mesh file '" + path + "'");
    const int count = std::stoi(trim(line));
    for (int i = 0; i < count; i++) {
      if (!std::getline(f, line)) throw std::runtime_error("truncated mesh file '" + path + "'");
      std::stringstream ss(line);
      if (pts) {
        for (int c = 0; c < dim; c++) {
          double x;
          ss >> x;
          pts->push_back(x);
        }
      } else {
        for (int c = 0; c < per; c++) {
          unsigned k;
          ss >> k;
          if (k == 0) throw std::runtime_error("mesh index 0 in '" + path + "' (indices are 1-based)");
          idx->push_back(k - 1);
        }
      }
    }
  };
  while (std::getline(f, line)) {
    const std::string t = trim(line);
    if (t.empty()) continue;
    if (t.rfind("MeshVersionFormatted", 0) == 0) continue;
    if (t.rfind("Dimension", 0) == 0) {
      std::stringstream ss(t);
      std::string kw;
      ss >> kw >> m.dim;
      if (!ss) {
        std::getline(f, line);
        m.dim = std::stoi(trim(line));
      }
      if (m.dim != dim)
        throw std::runtime_error("mesh '" + path + "' has dimension " + std::to_string(m.dim) + ", expected " +
                                 std::to_string(dim));
    } else if (t == "Vertices") {
      read_block(0, nullptr, &m.v);
    } else if (t == "Edges") {
      read_block(2, &m.edges, nullptr);
    } else if (t == "Triangles") {
      read_block(3, &m.tris, nullptr);
    } else if (t == "Tetrahedra") {
      read_block(4, &m.tets, nullptr);
    } else if (t == "End") {
      break;
    } else {
      std::cerr << "[Mesh WARN] unknown keyword '" << t << "' in " << path << std::endl;
    }
  }
  const unsigned nv = (unsigned)(m.v.size() / dim);
  for (const auto* a : {&m.edges, &m.tris, &m.tets})
    for (unsigned k : *a)
      if (k >= nv) throw std::runtime_error("mesh '" + path + "' references a missing vertex");
  return m;
}

// glm::radians<float>: degrees times the float value of pi/180.
static float radians_f(float deg) { return deg * 0.01745329251994329576923690768489f; }

// Mesh2D::applyInitialTransform (DeformableMesh.cpp:403-425), in the
// reference's double arithmetic: scale and translation rounded to float, the
// angle converted to radians in float, and glm's column-major
// mat2(cos, -sin, sin, cos), which maps (x, y) -> (cos x + sin y, -sin x + cos y).
static void transform2(std::vector<double>& v, const SolidBodyConfig<2>& c) {
  const double sx = (float)c.scale[0], sy = (float)c.scale[1];
  const double ang = (double)radians_f((float)c.rotate[0]);
  const double cs = std::cos(ang), sn = std::sin(ang);
  const double tx = (float)c.translate[0], ty = (float)c.translate[1];
  for (size_t i = 0; i + 1 < v.size(); i += 2) {
    const double x = sx * v[i], y = sy * v[i + 1];
    v[i] = cs * x + sn * y + tx;
    v[i + 1] = -sn * x + cs * y + ty;
  }
}

// Mesh3D::applyInitialTransform (DeformableMesh.cpp:159-187), in float as
// the reference's glm::mat4 is: M = T * R * S with
// R = eulerAngleYXZ(yaw = rotate[1], pitch = rotate[0], roll = rotate[2]),
// so column c of M is R's column c times scale[c] and column 3 is the
// translation; M * (x, y, z, 1) sums glm's column products as
// (c0 x + c1 y) + (c2 z + t).
static void transform3(std::vector<double>& v, const SolidBodyConfig<3>& c) {
  const float yaw = radians_f((float)c.rotate[1]), pitch = radians_f((float)c.rotate[0]);
  const float roll = radians_f((float)c.rotate[2]);
  const float ch = std::cos(yaw), sh = std::sin(yaw), cp = std::cos(pitch), sp = std::sin(pitch);
  const float cb = std::cos(roll), sb = std::sin(roll);
  // R[col][row] of Ry(yaw) * Rx(pitch) * Rz(roll)
  const float R[3][3] = {{ch * cb + sh * sp * sb, sb * cp, -sh * cb + ch * sp * sb},
                         {-ch * sb + sh * sp * cb, cb * cp, sb * sh + ch * sp * cb},
                         {sh * cp, -sp, ch * cp}};
  const float s[3] = {(float)c.scale[0], (float)c.scale[1], (float)c.scale[2]};
  const float t[3] = {(float)c.translate[0], (float)c.translate[1], (float)c.translate[2]};
  float M[3][3];
  for (int col = 0; col < 3; col++)
    for (int r = 0; r < 3; r++) M[col][r] = R[col][r] * s[col];
  for (size_t i = 0; i + 2 < v.size(); i += 3) {
    const float x = (float)v[i], y = (float)v[i + 1], z = (float)v[i + 2];
    for (int r = 0; r < 3; r++) {
      const float a = M[0][r] * x + M[1][r] * y;
      const float b = M[2][r] * z + t[r];
      v[i + r] = (double)(a + b);
    }
  }
}

// Mesh2D::extractEdges (DeformableMesh.cpp:427-504): explicit edges plus
// triangle edges used once, as a sorted set of sorted vertex pairs.
static std::vector<unsigned> boundary_edges(const RawMesh& m) {
  std::set<std::array<unsigned, 2>> es;
  for (size_t i = 0; i + 1 < m.edges.size(); i += 2) {
    std::array<unsigned, 2> e = {m.edges[i], m.edges[i + 1]};
    std::sort(e.begin(), e.end());
    es.insert(e);
  }
  std::map<std::array<unsigned, 2>, int> cnt;
  for (size_t i = 0; i + 2 < m.tris.size(); i += 3) {
    const unsigned* t = &m.tris[i];
    const std::array<unsigned, 2> ed[3] = {{t[0], t[1]}, {t[1], t[2]}, {t[2], t[0]}};
    for (auto e : ed) {
      std::sort(e.begin(), e.end());
      cnt[e]++;
    }
  }
  for (const auto& kv : cnt)
    if (kv.second == 1) es.insert(kv.first);
  std::vector<unsigned> out;
  for (const auto& e : es) {
    out.push_back(e[0]);
    out.push_back(e[1]);
  }
  return out;
}

// Mesh3D::extractSurface (DeformableMesh.cpp:188-260): file triangles first,
// then tetrahedron faces used once, in sorted-face order with the face's
// original orientation.
static std::vector<unsigned> boundary_tris(const RawMesh& m) {
  std::vector<unsigned> out(m.tris.begin(), m.tris.end());
  std::map<std::array<unsigned, 3>, std::pair<int, std::array<unsigned, 3>>> faces;
  for (size_t i = 0; i + 3 < m.tets.size(); i += 4) {
    const unsigned* t = &m.tets[i];
    const std::array<unsigned, 3> fs[4] = {{t[0], t[2], t[1]}, {t[0], t[3], t[2]}, {t[0], t[1], t[3]}, {t[1], t[2], t[3]}};
    for (const auto& f : fs) {
      std::array<unsigned, 3> k = f;
      std::sort(k.begin(), k.end());
      faces[k].first++;
      faces[k].second = f;
    }
  }
  for (const auto& kv : faces)
    if (kv.second.first == 1) out.insert(out.end(), kv.second.second.begin(), kv.second.second.end());
  return out;
}

// ---------------------------------------------------------------- simulator
template <int D>
Simulator<D>::Simulator(const SimulationParameters<D>& params) : m_params(params) {}

template <int D>
Simulator<D>::~Simulator() {
  if (m_dom) ff_domain_destroy(m_dom);
}

template <int D>
void Simulator<D>::check(int rc, const char* what) const {
  if (rc != FF_OK) throw std::runtime_error(std::string(what) + ": " + ff_last_error());
}

template <int D>
void Simulator<D>::upload_flags_and_state() {
  const int nx = m_params.fluid_nx, ny = m_params.fluid_ny, nz = D == 3 ? m_params.fluid_nz : 1;
  std::vector<uint8_t> flags((size_t)nx * ny * nz, FF_FLAG_FLUID);
  // LbmInitializer2D.cu:106-122: list order, later entries win, out-of-grid ignored
  for (const BoundaryCondition& bc : m_params.boundaries) {
    if ((int)bc.position.size() < D) continue;
    const int x = bc.position[0], y = bc.position[1], z = D == 3 ? bc.position[2] : 0;
    if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
      flags[((size_t)z * ny + y) * nx + x] = (uint8_t)bc.flag;
    else
      std::cerr << "[WARN] boundary condition outside the grid is ignored" << std::endl;
  }
  check(ff_set_flags(m_dom, flags.data()), "ff_set_flags");
  const std::vector<float>& bv = m_params.boundary_velocities;
  check(ff_init_state(m_dom, bv.empty() ? nullptr : bv.data(), (int)bv.size()), "ff_init_state");
}

template <int D>
void Simulator<D>::load_bodies() {
  m_bodies.clear();
  m_vol_offset.clear();
  m_pos.clear();
  m_vel.clear();
  m_surf_elems.clear();
  m_surf_to_vol.clear();
  m_elem_body.clear();
  int surf_offset = 0;
  for (size_t b = 0; b < m_params.solids.size(); b++) {
    const SolidBodyConfig<D>& c = m_params.solids[b];
    if (c.mesh_path.empty()) throw std::runtime_error("solid has an empty mesh_path");
    RawMesh m = load_mesh(c.mesh_path, D);
    Body<D> body;
    body.rest = m.v;
    if constexpr (D == 2) {
      transform2(body.rest, c);
      body.cells = m.tris;
      body.surf_elems = boundary_edges(m);
    } else {
      transform3(body.rest, c);
      body.cells = m.tets;
      body.tris = m.tris;
      body.surf_elems = boundary_tris(m);
    }
    for (int k = 0; k < D; k++) body.v0[k] = c.initial_velocity[k];
    const int nv = (int)(body.rest.size() / D);
    std::vector<int> v2s(nv, -1);
    std::vector<bool> on(nv, false);
    for (unsigned k : body.surf_elems) on[k] = true;
    for (int i = 0; i < nv; i++)
      if (on[i]) {
        v2s[i] = (int)body.surf_to_vol.size();
        body.surf_to_vol.push_back(i);
      }
    const int vol_offset = (int)(m_pos.size() / D);
    m_vol_offset.push_back(vol_offset);
    m_pos.insert(m_pos.end(), body.rest.begin(), body.rest.end());
    for (int i = 0; i < nv; i++)
      for (int k = 0; k < D; k++) m_vel.push_back(body.v0[k]);
    for (unsigned k : body.surf_elems) m_surf_elems.push_back((unsigned)(v2s[k] + surf_offset));
    for (int s : body.surf_to_vol) m_surf_to_vol.push_back(s + vol_offset);
    for (size_t e = 0; e < body.surf_elems.size() / D; e++) m_elem_body.push_back((int)b);
    surf_offset += (int)body.surf_to_vol.size();
    m_bodies.push_back(std::move(body));
  }
}

// Surface proxy from the volume state (update_surface_pos_kernel2d: FP64 ->
// FP32 gather, src/fsi/FsiCouplingKernels.cu:38-58).
template <int D>
void Simulator<D>::push_boundary(bool set_topology) {
  const int ns = (int)m_surf_to_vol.size();
  std::vector<float> sp((size_t)ns * D), sv((size_t)ns * D);
  for (int s = 0; s < ns; s++)
    for (int k = 0; k < D; k++) {
      sp[(size_t)s * D + k] = (float)m_pos[(size_t)m_surf_to_vol[s] * D + k];
      sv[(size_t)s * D + k] = (float)m_vel[(size_t)m_surf_to_vol[s] * D + k];
    }
  if (set_topology) {
    const int ne = (int)(m_surf_elems.size() / D);
    check(ff_set_boundary(m_dom, ns, sp.data(), sv.data(), ne, m_surf_elems.data(), m_elem_body.data(),
                          std::max<int>(1, (int)m_bodies.size())),
          "ff_set_boundary");
  } else {
    check(ff_update_boundary(m_dom, sp.data(), sv.data()), "ff_update_boundary");
  }
}

template <int D>
void Simulator<D>::initialize() {
  if (m_params.solid_solver_type != "static" && m_params.solid_solver_type != "vbd")
    throw std::runtime_error("Invalid solid_solver_type '" + m_params.solid_solver_type +
                             "'. Supported types are 'static' and 'vbd'.");
  if (m_params.solid_solver_type == "vbd")
    std::cerr << "[WARN] solid_solver_type 'vbd': solids are moved kinematically only in this build" << std::endl;
  if (m_dom) {
    ff_domain_destroy(m_dom);
    m_dom = nullptr;
  }
  ff_domain_desc desc{};
  desc.dim = D;
  desc.nx = m_params.fluid_nx;
  desc.ny = m_params.fluid_ny;
  desc.nz = D == 3 ? m_params.fluid_nz : 1;
  desc.viscosity = m_params.fluid_viscosity;
  desc.dx = m_params.fluid_dx;
  desc.dt = m_params.dt;
  desc.fluid_density = m_params.fluid_density;
  const std::string& st = m_params.device_options.storage;
  desc.storage = st == "f16" ? FF_STORAGE_F16
                : st == "bf16" ? FF_STORAGE_BF16
                : st == "f16u" ? FF_STORAGE_F16U
                               : FF_STORAGE_F32;
  desc.options = (m_params.device_options.stage_fp16 && D == 2 ? FF_OPT_STAGE_FP16 : 0u) |
                 (m_params.device_options.periodic ? FF_OPT_PERIODIC : 0u) |
                 (m_params.device_options.fast_math ? FF_OPT_FAST_MATH : 0u);
  desc.device = m_params.device_options.device;
  check(ff_domain_create(&desc, &m_dom), "ff_domain_create");
  upload_flags_and_state();
  load_bodies();
  if (!m_surf_to_vol.empty()) push_boundary(true);
  if (m_params.output_frequency > 0 && !m_params.output_path.empty()) {
    std::string p;
    for (const char ch : m_params.output_path + "/") {
      if (ch == '/' && !p.empty()) mkdir(p.c_str(), 0755);
      p += ch;
    }
  }
  m_targets.clear();
  m_target_vel = 0.0;
  m_target_idx = -1;
  m_wave = false;
}

template <int D>
void Simulator<D>::reset() {
  if (!m_dom) return;
  upload_flags_and_state();
  load_bodies();
  if (!m_surf_to_vol.empty()) push_boundary(true);
  m_targets.clear();
  m_target_vel = 0.0;
  m_target_idx = -1;
  m_wave = false;
}

template <int D>
void Simulator<D>::step() {
  if (!m_dom) throw std::runtime_error("Simulator not initialized");
  if (m_fluid_enabled) check(ff_step(m_dom, 1), "ff_step");
  // StaticBodySolver2D::advanceTimeStep (StaticBodySolver.cpp:122-151) with
  // glm's double arithmetic: length = sqrt(dot), normalize = d * (1 / sqrt(dot)),
  // every volume vertex gets velocity vel * dir and moves by (vel * dir) * dt.
  if (!m_targets.empty() && m_target_vel > 0.0 && !m_pos.empty()) {
    std::array<double, D> marker;
    for (int k = 0; k < D; k++) marker[k] = m_pos[k];
    if (m_target_idx < 0) m_target_idx = 0;
    auto diff = [&](const std::array<double, D>& t) {
      std::array<double, D> d;
      for (int k = 0; k < D; k++) d[k] = t[k] - marker[k];
      return d;
    };
    auto dot = [](const std::array<double, D>& d) {
      double s2 = d[0] * d[0] + d[1] * d[1];
      if constexpr (D == 3) s2 = s2 + d[2] * d[2];
      return s2;
    };
    if (std::sqrt(dot(diff(m_targets[m_target_idx]))) < m_params.dt * m_target_vel)
      m_target_idx = (m_target_idx + 1) % (int)m_targets.size();
    const std::array<double, D> d = diff(m_targets[m_target_idx]);
    const double inv = 1.0 / std::sqrt(dot(d));
    std::array<double, D> v, shift;
    for (int k = 0; k < D; k++) {
      v[k] = m_target_vel * (d[k] * inv);
      shift[k] = v[k] * m_params.dt;
    }
    const size_t nvol = m_pos.size() / D;
    for (size_t i = 0; i < nvol; i++)
      for (int k = 0; k < D; k++) {
        m_vel[i * D + k] = v[k];
        m_pos[i * D + k] += shift[k];
      }
    if (!m_surf_to_vol.empty()) push_boundary(false);
  }
}

template <int D>
std::vector<std::array<int, 2>> Simulator<D>::fillSolid() {
  if (!m_dom) throw std::runtime_error("Simulator not initialized");
  std::vector<std::array<int, 2>> out;
  if constexpr (D != 2) {
    throw NotImplemented("fillSolid is a 2D operation");
  } else {
    if (m_surf_to_vol.empty()) return out;
    check(ff_fill_solid(m_dom), "ff_fill_solid");
    const int nx = m_params.fluid_nx, ny = m_params.fluid_ny;
    std::vector<uint8_t> flags((size_t)nx * ny);
    check(ff_get_flags(m_dom, flags.data()), "ff_get_flags");
    for (size_t i = 0; i < flags.size(); i++)
      if (flags[i] == FF_FLAG_SOLID_DYNAMIC) out.push_back({(int)(i % nx), (int)(i / nx)});
  }
  return out;
}

template <int D>
void Simulator<D>::applyKinematicControl(int mesh_id, const std::vector<std::array<double, D>>& target_pos, double vel) {
  // StaticBodySolver2D::applyKinematicControl (StaticBodySolver.cpp:158-180)
  if (mesh_id < 0 || mesh_id >= (int)m_bodies.size()) {
    std::cerr << "[WARN] Invalid mesh_id " << mesh_id << " for kinematic control." << std::endl;
    return;
  }
  if (target_pos.size() < 2) {
    std::cerr << "[WARN] At least two target positions are required for kinematic control." << std::endl;
    return;
  }
  if (vel <= 0.0) {
    std::cerr << "[WARN] Velocity must be positive for kinematic control." << std::endl;
    return;
  }
  m_targets = target_pos;
  m_target_vel = vel;
  m_target_idx = -1;
}

template <int D>
std::vector<float> Simulator<D>::getFluidMoment(int which) const {
  if (!m_dom) throw std::runtime_error("Simulator not initialized");
  const size_t n = (size_t)m_params.fluid_nx * m_params.fluid_ny * (D == 3 ? m_params.fluid_nz : 1);
  std::vector<float> out(n * (D == 2 ? 6 : 10));
  check(ff_get_moments(m_dom, out.data(), which), "ff_get_moments");
  return out;
}

template <int D>
std::vector<std::array<double, D>> Simulator<D>::getVertices() const {
  std::vector<std::array<double, D>> out(m_pos.size() / D);
  for (size_t i = 0; i < out.size(); i++)
    for (int k = 0; k < D; k++) out[i][k] = m_pos[i * D + k];
  return out;
}

template <int D>
std::vector<std::array<double, D>> Simulator<D>::getVelocity() const {
  std::vector<std::array<double, D>> out(m_vel.size() / D);
  for (size_t i = 0; i < out.size(); i++)
    for (int k = 0; k < D; k++) out[i][k] = m_vel[i * D + k];
  return out;
}

template <int D>
std::vector<std::array<double, D>> Simulator<D>::getForce() const {
  std::vector<std::array<double, D>> out(m_pos.size() / D);
  for (auto& a : out) a.fill(0.0);
  if (!m_dom || m_surf_to_vol.empty()) return out;
  std::vector<float> f(m_surf_to_vol.size() * D);
  check(ff_get_step_vertex_forces(m_dom, f.data()), "ff_get_step_vertex_forces");
  for (size_t s = 0; s < m_surf_to_vol.size(); s++)
    for (int k = 0; k < D; k++) out[m_surf_to_vol[s]][k] = f[s * D + k];
  return out;
}

template <int D>
std::vector<unsigned> Simulator<D>::getCells() const {
  std::vector<unsigned> out;
  for (size_t b = 0; b < m_bodies.size(); b++)
    for (unsigned k : m_bodies[b].cells) out.push_back(k + (unsigned)m_vol_offset[b]);
  return out;
}

template <int D>
std::vector<std::array<double, D>> Simulator<D>::getBoundaryVertices() const {
  std::vector<std::array<double, D>> out(m_surf_to_vol.size());
  if (out.empty()) return out;
  if (!m_dom) {  // before initialize(): the proxy's FP32 gather of the volume state
    for (size_t i = 0; i < out.size(); i++)
      for (int k = 0; k < D; k++) out[i][k] = (float)m_pos[(size_t)m_surf_to_vol[i] * D + k];
    return out;
  }
  std::vector<float> v(out.size() * D);
  check(ff_get_boundary(m_dom, v.data(), nullptr), "ff_get_boundary");
  for (size_t i = 0; i < out.size(); i++)
    for (int k = 0; k < D; k++) out[i][k] = v[i * D + k];
  return out;
}

template <int D>
std::vector<unsigned> Simulator<D>::getBoundaryElements() const {
  return m_surf_elems;
}

template <int D>
std::vector<unsigned> Simulator<D>::getBoundaryPointIdx() const {
  return std::vector<unsigned>(m_surf_to_vol.begin(), m_surf_to_vol.end());
}

template <int D>
std::vector<double> Simulator<D>::getBodyWrench() const {
  const int nb = std::max<int>(1, (int)m_bodies.size());
  std::vector<double> out((size_t)nb * (D == 2 ? 3 : 6), 0.0);
  if (m_dom && !m_surf_to_vol.empty()) check(ff_get_body_wrench(m_dom, out.data()), "ff_get_body_wrench");
  return out;
}

// Per-body power of the most recent step, sum over the body's boundary
// vertices of -F . u: what the reference's RL loop computes on the host from
// every vertex force and velocity (rl/env.py:175-178).
template <int D>
std::vector<double> Simulator<D>::getBodyPower() const {
  const int nb = std::max<int>(1, (int)m_bodies.size());
  std::vector<double> out((size_t)nb, 0.0);
  if (m_dom && !m_surf_to_vol.empty()) check(ff_get_body_power(m_dom, out.data()), "ff_get_body_power");
  return out;
}

template <int D>
void Simulator<D>::setWave(const std::vector<float>& s_coord, const std::vector<double>& params) {
  if (m_surf_to_vol.empty()) throw std::runtime_error("no boundary to drive");
  if (s_coord.size() != m_surf_to_vol.size()) throw std::runtime_error("s_coord needs one value per boundary vertex");
  if (params.size() != 4 * std::max<size_t>(1, m_bodies.size())) throw std::runtime_error("params needs 4 values per body");
  std::vector<float> rest(m_surf_to_vol.size() * D);
  for (size_t s = 0; s < m_surf_to_vol.size(); s++)
    for (int k = 0; k < D; k++) rest[s * D + k] = (float)m_pos[(size_t)m_surf_to_vol[s] * D + k];
  check(ff_set_wave(m_dom, rest.data(), s_coord.data(), params.data()), "ff_set_wave");
  m_wave = true;
}

template <int D>
double Simulator<D>::fluidMass() const {
  double m = 0.0;
  check(ff_mass(m_dom, &m), "ff_mass");
  return m;
}

// Legacy-VTK output (the role of VtkWriter: velocity image data for the fluid,
// volume cells for the solid).
template <int D>
void Simulator<D>::saveFrameData(int frame_index, bool save_fluid, bool save_solid) const {
  if (m_params.output_frequency <= 0 || m_params.output_path.empty()) return;
  const int nx = m_params.fluid_nx, ny = m_params.fluid_ny, nz = D == 3 ? m_params.fluid_nz : 1;
  if (save_fluid && m_dom) {
    const std::vector<float> m = getFluidMoment(0);
    const int nm = D == 2 ? 6 : 10;
    const std::string path = m_params.output_path + "/fluid_frame_" + std::to_string(frame_index) + ".vtk";
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << "# vtk DataFile Version 3.0\nfluid velocity\nASCII\nDATASET STRUCTURED_POINTS\n";
    f << "DIMENSIONS " << nx << " " << ny << " " << nz << "\nORIGIN 0 0 0\nSPACING " << m_params.fluid_dx << " "
      << m_params.fluid_dx << " " << m_params.fluid_dx << "\n";
    f << "POINT_DATA " << (size_t)nx * ny * nz << "\nVECTORS velocity float\n";
    for (size_t c = 0; c < (size_t)nx * ny * nz; c++)
      f << m[c * nm + 1] << " " << m[c * nm + 2] << " " << (D == 3 ? m[c * nm + 3] : 0.0f) << "\n";
  }
  if (save_solid && !m_pos.empty()) {
    const std::string path = m_params.output_path + "/solid_frame_" + std::to_string(frame_index) + ".vtk";
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    const size_t nv = m_pos.size() / D;
    f << "# vtk DataFile Version 3.0\nsolid\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS " << nv << " double\n";
    for (size_t i = 0; i < nv; i++) f << m_pos[i * D] << " " << m_pos[i * D + 1] << " " << (D == 3 ? m_pos[i * D + 2] : 0.0) << "\n";
    const std::vector<unsigned> cells = getCells();
    const int per = D == 2 ? 3 : 4;
    const size_t nc = cells.size() / per;
    f << "CELLS " << nc << " " << nc * (per + 1) << "\n";
    for (size_t c = 0; c < nc; c++) {
      f << per;
      for (int k = 0; k < per; k++) f << " " << cells[c * per + k];
      f << "\n";
    }
    f << "CELL_TYPES " << nc << "\n";
    for (size_t c = 0; c < nc; c++) f << (D == 2 ? 5 : 10) << "\n";
  }
}

template class Simulator<2>;
template class Simulator<3>;

}  // namespace ffsim
