// sim_host.h -- host-side drop-in for FreeFlow's Config/Simulator layer, driving
// the fluid step through the C ABI (include/freeflow_lbm.h).
//
// Mirrors include/FSI_Simulator/utils/Config.hpp (schema and defaults),
// include/FSI_Simulator/core/Simulator2D.hpp / Simulator3D.hpp (methods) and
// the static/kinematic body of src/fem/StaticBodySolver.cpp.  Solids are rigid
// here: their boundary (segments in 2D, triangles in 3D) is extracted from
// the mesh as src/fem/DeformableMesh.cpp does, placed by the initial
// transform and moved only by kinematic control or a prescribed wave.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/freeflow_lbm.h"

namespace ffsim {

struct NotImplemented : std::logic_error {
  using std::logic_error::logic_error;
};

// LbmNodeFlag names accepted by the JSON loader (Config.hpp:29-62); anything
// else maps to Invalid.
int flag_from_name(const std::string& s);

struct BoundaryCondition {
  std::vector<int> position;
  int flag = 0;
};

struct LBSControlConfig {  // FemConfig.hpp: kept for schema compatibility
  int cnum = 2;
  std::string lbs_distance_type = "geodesic";
  bool random_first = false;
  double omega = 0.5;
  double stiffness = 10.0;
};

struct FemSolverOptions {  // FemConfig.hpp: kept for schema compatibility
  std::string optimizer_type = "newton";
  int iterations = 10;
  std::string line_search_method = "backtracking";
  int ls_max_iter = 10;
  double ls_beta = 0.5;
  double ls_alpha = 1e-4;
  std::string linear_solver_type = "eigen_ldlt";
  double force_density_abs_tol = 1e-3;
  bool grad_check = false;
  int thread_ct = 8;
  int substeps = 3;
  double omega = 0.8;
  int vbd_iterations = 30;
};

template <int D>
struct SolidBodyConfig {
  std::string mesh_path;
  double density = 1000.0;
  double youngs_modulus = 1.0e6;
  double poisson_ratio = 0.45;
  std::array<double, D> translate{};
  std::array<double, D == 2 ? 1 : 3> rotate{};
  std::array<double, D> scale{};
  std::array<double, D> initial_velocity{};
  LBSControlConfig lbs_control_config;
  SolidBodyConfig() { scale.fill(1.0); }
};
using SolidBodyConfig2D = SolidBodyConfig<2>;
using SolidBodyConfig3D = SolidBodyConfig<3>;

// Fluid options beyond the reference schema ("freeflow_amd" object in the
// JSON); defaults reproduce the reference's behaviour.
struct DeviceOptions {
  std::string storage = "f32";  // f32 | f16 | bf16 | f16u (FP16, velocity as a hi + lo pair)
  bool stage_fp16 = true;       // 2D: the reference's FP16 population tile
  bool periodic = false;        // periodic pulls instead of the clamped halo
  bool fast_math = false;       // FMA-reassociated update (2D without stage_fp16; 3D)
  int device = 0;
};

template <int D>
struct SimulationParameters {
  int dimension = D;
  double fluid_viscosity = 0.001;
  double fluid_density = 1000.0;
  int fluid_nx = D == 2 ? 400 : 128;
  int fluid_ny = D == 2 ? 400 : 128;
  int fluid_nz = 128;
  double fluid_dx = 1.0;
  std::string solid_solver_type = "static";
  double total_time = 10.0;
  double dt = 0.001;
  int output_frequency = 100;
  std::string output_path = "./output";
  std::string log_level = "info";
  std::string log_file = "simulation.log";
  bool use_newton = true;
  std::vector<BoundaryCondition> boundaries;
  std::vector<float> boundary_velocities;
  FemSolverOptions global_fem_options;
  std::vector<SolidBodyConfig<D>> solids;
  DeviceOptions device_options;
  // SimulationParameters2D/3D::getCf (Config.hpp:96-100, 154-158)
  float getCf() const {
    float cm = D == 2 ? (float)(fluid_density * fluid_dx * fluid_dx) : (float)(fluid_density * fluid_dx * fluid_dx * fluid_dx);
    return (float)((double)cm * fluid_dx / (dt * dt));
  }
};
using SimulationParameters2D = SimulationParameters<2>;
using SimulationParameters3D = SimulationParameters<3>;

template <int D>
class Config {
 public:
  bool load(const std::string& filepath);
  const SimulationParameters<D>& getParams() const { return m_params; }
  SimulationParameters<D>& params() { return m_params; }

 private:
  SimulationParameters<D> m_params;
};
using Config2D = Config<2>;
using Config3D = Config<3>;

// One loaded solid: volume vertices (double, like the reference's real),
// cells (triangles in 2D, tetrahedra in 3D) and the extracted boundary.
template <int D>
struct Body {
  std::vector<double> rest;            // nv * D after the initial transform
  std::vector<unsigned> cells;         // triangles (2D) / tetrahedra (3D)
  std::vector<unsigned> tris;          // 3D: triangles given in the mesh file
  std::vector<unsigned> surf_elems;    // boundary elements, local vertex ids
  std::vector<int> surf_to_vol;        // surface vertex -> local volume vertex
  std::array<double, D> v0{};          // initial velocity
};

template <int D>
class Simulator {
 public:
  explicit Simulator(const SimulationParameters<D>& params);
  ~Simulator();
  Simulator(const Simulator&) = delete;
  Simulator& operator=(const Simulator&) = delete;

  void initialize();
  void reset();
  void step();
  void enableFluidSolver(bool enable) { m_fluid_enabled = enable; }
  std::vector<std::array<int, 2>> fillSolid();
  void applyKinematicControl(int mesh_id, const std::vector<std::array<double, D>>& target_pos, double vel);
  void saveFrameData(int frame_index, bool save_fluid, bool save_solid) const;

  std::vector<float> getFluidMoment(int which = 0) const;
  std::vector<std::array<double, D>> getVertices() const;
  std::vector<std::array<double, D>> getVelocity() const;
  std::vector<std::array<double, D>> getForce() const;
  std::vector<unsigned> getCells() const;
  std::vector<std::array<double, D>> getBoundaryVertices() const;
  std::vector<unsigned> getBoundaryElements() const;
  std::vector<unsigned> getBoundaryPointIdx() const;

  // extensions
  // Mesh loading, initial transform and boundary extraction alone (no device
  // work): what initialize() does for the solids.
  void loadSolids() { load_bodies(); }
  std::vector<double> getBodyWrench() const;
  std::vector<double> getBodyPower() const;
  void setWave(const std::vector<float>& s_coord, const std::vector<double>& params);
  double fluidMass() const;
  ff_domain* domain() const { return m_dom; }

  const SimulationParameters<D>& m_params;

 private:
  void check(int rc, const char* what) const;
  void load_bodies();
  void upload_flags_and_state();
  void push_boundary(bool set_topology);
  std::vector<Body<D>> m_bodies;
  std::vector<int> m_vol_offset;       // first volume vertex of each body
  std::vector<double> m_pos, m_vel;    // scene volume state (double)
  std::vector<unsigned> m_surf_elems;  // global surface element indices
  std::vector<int> m_surf_to_vol;      // global surface vertex -> volume vertex
  std::vector<int> m_elem_body;
  ff_domain* m_dom = nullptr;
  bool m_fluid_enabled = true;
  bool m_wave = false;
  // static-body kinematic control (StaticBodySolver2D::advanceTimeStep)
  std::vector<std::array<double, D>> m_targets;
  double m_target_vel = 0.0;
  int m_target_idx = -1;
};
using Simulator2D = Simulator<2>;
using Simulator3D = Simulator<3>;

}  // namespace ffsim
