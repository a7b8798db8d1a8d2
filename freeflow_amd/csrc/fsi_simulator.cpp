// fsi_simulator.cpp -- pybind11 module with the reference's Python surface.
//
// Class, method and argument names follow src/python/bindings.cpp:133-328
// (module `fsi_simulator`, submodule `fem`, Config2D/3D, SimulationParameters2D/3D,
// Simulator2D/3D).  The fluid step underneath is the gfx950 library behind
// include/freeflow_lbm.h.  Vectors the reference returns as lists of 2/3-element
// numpy arrays (its vec2_t/vec3_t casters, bindings.cpp:24-90) come back as one
// (n, dim) float64 array, which indexes and converts the same way.  Methods of
// the deformable/LBS solid solvers raise NotImplementedError: solids in this
// library are kinematic (static body + prescribed motion).

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "sim_host.h"

namespace py = pybind11;
using namespace ffsim;

namespace {

template <int D>
py::array_t<double> rows(const std::vector<std::array<double, D>>& v) {
  py::array_t<double> a({(py::ssize_t)v.size(), (py::ssize_t)D});
  auto m = a.template mutable_unchecked<2>();
  for (size_t i = 0; i < v.size(); i++)
    for (int k = 0; k < D; k++) m(i, k) = v[i][k];
  return a;
}

template <int D>
std::vector<std::array<double, D>> points_from(py::handle obj) {
  py::array_t<double, py::array::c_style | py::array::forcecast> a =
      py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(obj);
  if (!a) throw py::type_error("target_pos must be convertible to a float array");
  if (a.ndim() != 2 || a.shape(1) != D)
    throw py::value_error("target_pos must have shape (n, " + std::to_string(D) + ")");
  std::vector<std::array<double, D>> out((size_t)a.shape(0));
  auto r = a.template unchecked<2>();
  for (size_t i = 0; i < out.size(); i++)
    for (int k = 0; k < D; k++) out[i][k] = r(i, k);
  return out;
}

template <int D>
void bind_params(py::module_& m, const char* name) {
  using P = SimulationParameters<D>;
  auto c = py::class_<P>(m, name)
               .def(py::init<>())
               .def_readwrite("solid_solver_type", &P::solid_solver_type)
               .def_readwrite("total_time", &P::total_time)
               .def_readwrite("dt", &P::dt)
               .def_readwrite("solids", &P::solids)
               .def_readwrite("global_fem_options", &P::global_fem_options)
               .def_readwrite("use_newton", &P::use_newton)
               .def_readwrite("output_path", &P::output_path)
               .def_readwrite("output_frequency", &P::output_frequency)
               .def_readwrite("log_level", &P::log_level)
               .def_readwrite("log_file", &P::log_file)
               .def_readwrite("boundary_velocities", &P::boundary_velocities)
               .def_readwrite("boundaries", &P::boundaries)
               .def_readwrite("fluid_viscosity", &P::fluid_viscosity)
               .def_readwrite("fluid_density", &P::fluid_density)
               .def_readwrite("fluid_nx", &P::fluid_nx)
               .def_readwrite("fluid_ny", &P::fluid_ny)
               .def_readwrite("fluid_dx", &P::fluid_dx)
               .def_readwrite("device_options", &P::device_options)
               .def("getCf", &P::getCf);
  if constexpr (D == 3) c.def_readwrite("fluid_nz", &P::fluid_nz);
}

template <int D>
void bind_solid(py::module_& fem, const char* name) {
  using S = SolidBodyConfig<D>;
  auto c = py::class_<S>(fem, name)
               .def(py::init<>())
               .def_readwrite("mesh_path", &S::mesh_path)
               .def_readwrite("density", &S::density)
               .def_readwrite("youngs_modulus", &S::youngs_modulus)
               .def_readwrite("poisson_ratio", &S::poisson_ratio)
               .def_readwrite("translate", &S::translate)
               .def_readwrite("scale", &S::scale)
               .def_readwrite("initial_velocity", &S::initial_velocity)
               .def_readwrite("lbs_control_config", &S::lbs_control_config);
  if constexpr (D == 2) {
    // SolidBodyConfig2D::rotate is one angle in degrees (FemConfig.hpp:98)
    c.def_property(
        "rotate", [](const S& s) { return s.rotate[0]; }, [](S& s, double v) { s.rotate[0] = v; });
  } else {
    c.def_readwrite("rotate", &S::rotate);
  }
}

template <int D>
void bind_config(py::module_& m, const char* name) {
  py::class_<Config<D>>(m, name)
      .def(py::init<>(), "Default constructor")
      .def("load", &Config<D>::load, py::arg("filepath"), "Load configuration from a JSON file.")
      .def("get_params", &Config<D>::getParams, py::return_value_policy::reference_internal,
           "Get a reference to the loaded simulation parameters.");
}

[[noreturn]] void not_here(const char* what) {
  throw NotImplemented(std::string(what) +
                       " needs the deformable (LBS/active-strain) solid solver, which this library does not provide");
}

template <int D>
void bind_simulator(py::module_& m, const char* name) {
  using S = Simulator<D>;
  auto c = py::class_<S>(m, name);
  c.def(py::init<const SimulationParameters<D>&>(), py::arg("params"), py::keep_alive<1, 2>(),
        "Constructor that takes simulation parameters.")
      .def("initialize", &S::initialize, "Initialize the simulation environment.")
      .def("reset", &S::reset, "Reset the simulation environment.")
      .def("step", &S::step, "Advance one time step.")
      .def("enableFluidSolver", &S::enableFluidSolver, py::arg("enable"), "enable/disable fluid solver.")
      .def(
          "apply_kinematic_control",
          [](S& s, int mesh_id, py::handle target_pos, double vel) {
            s.applyKinematicControl(mesh_id, points_from<D>(target_pos), vel);
          },
          py::arg("mesh_id"), py::arg("target_pos"), py::arg("vel"), "Apply kinematic control to a specific mesh.")
      .def(
          "apply_lbs_control", [](S&, int, py::object, py::object) { not_here("apply_lbs_control"); },
          py::arg("mesh_id"), py::arg("lbs_shift"), py::arg("lbs_rotation"))
      .def(
          "get_fluid_moments",
          [](const S& s) {
            const std::vector<float> mom = s.getFluidMoment(0);
            const auto& p = s.m_params;
            std::vector<py::ssize_t> shape;
            if constexpr (D == 2)
              shape = {p.fluid_ny, p.fluid_nx, 6};
            else
              shape = {p.fluid_nz, p.fluid_ny, p.fluid_nx, 10};
            return py::array_t<float>(shape, mom.data());
          },
          "Get the current fluid moment fields.")
      .def("getVertices", [](const S& s) { return rows<D>(s.getVertices()); }, "Get the position of solid vertices")
      .def("getVelocity", [](const S& s) { return rows<D>(s.getVelocity()); }, "Get the velocity of solid vertices")
      .def(
          "getForce", [](const S& s) { return rows<D>(s.getForce()); },
          "Hydrodynamic force of the most recent step on each solid vertex")
      .def(
          "getBoundaryVertices", [](const S& s) { return rows<D>(s.getBoundaryVertices()); },
          "Get the position of boundary vertices")
      .def("getBoundaryElements", &S::getBoundaryElements, "Get the indices of surface elements of solid meshes")
      .def("getBoundaryPointIdx", &S::getBoundaryPointIdx, "Get the indices of boundary points")
      .def("getControlPointIdx", [](const S&) { return std::vector<int>(); },
           "Indices of LBS control points (none: solids are kinematic)")
      .def("getLBSVertices", [](const S&) -> py::object { not_here("getLBSVertices"); })
      .def("begin_profiler", [](S&, const std::string&) {}, py::arg("name"),
           "Accepted for compatibility; use rocprofv3 for kernel timing.")
      .def("end_profiler", [](S&) {})
      // extensions of this library
      .def(
          "get_body_wrench",
          [](const S& s) {
            std::vector<double> w = s.getBodyWrench();
            const py::ssize_t per = D == 2 ? 3 : 6;
            return py::array_t<double>({(py::ssize_t)w.size() / per, per}, w.data());
          },
          "Per-body FP64 wrench of the most recent step: 2D (Fx, Fy, Tz), 3D (Fx, Fy, Fz, Tx, Ty, Tz).")
      .def(
          "get_body_power",
          [](const S& s) {
            std::vector<double> p = s.getBodyPower();
            return py::array_t<double>((py::ssize_t)p.size(), p.data());
          },
          "Per-body FP64 power of the most recent step: the sum over the body's boundary vertices of -F . u.")
      .def(
          "get_fluid_moments_buffer",
          [](const S& s, int which) {
            const std::vector<float> mom = s.getFluidMoment(which);
            return py::array_t<float>((py::ssize_t)mom.size(), mom.data());
          },
          py::arg("which"), "Flat AoS moments of ping-pong buffer `which` (0 current, 1 other).")
      .def(
          "set_wave",
          [](S& s, const std::vector<float>& s_coord, const std::vector<double>& params) { s.setWave(s_coord, params); },
          py::arg("s_coord"), py::arg("params"),
          "Travelling-wave kinematics of the boundary evaluated on the GPU each step.")
      .def("fluid_mass", &S::fluidMass, "FP64 sum of rho over fluid cells.")
      .def("load_solids", &S::loadSolids,
           "Load the solid meshes, apply their initial transforms and extract their boundaries without GPU work.");
  if constexpr (D == 2) {
    c.def("fillSolid", &S::fillSolid, "Set the flags of fluid nodes that lie in the interior of solid meshes.")
        .def("getTriangles", &S::getCells, "Get the indices of triangles of solid meshes")
        .def(
            "apply_active_strain", [](S&, int, int, py::object, double) { not_here("apply_active_strain"); },
            py::arg("mesh_id"), py::arg("tri_id"), py::arg("dir"), py::arg("magnitude"))
        .def(
            "save_frame_data", [](const S& s, int frame) { s.saveFrameData(frame, true, true); },
            py::arg("frame_index"), "Save simulation data for a given frame.");
  } else {
    c.def("getTetrahedra", &S::getCells, "Get the indices of tetrahedra of solid meshes")
        .def(
            "apply_active_strain", [](S&, int, int, py::object, double) { not_here("apply_active_strain"); },
            py::arg("mesh_id"), py::arg("tet_id"), py::arg("dir"), py::arg("magnitude"))
        .def("save_frame_data", &S::saveFrameData, py::arg("frame_index"), py::arg("save_fluid") = true,
             py::arg("save_solid") = true, "Save simulation data for a given frame.");
  }
}

}  // namespace

PYBIND11_MODULE(fsi_simulator, m) {
  m.doc() = "Fluid-structure simulator: HOME-LBM fluid step on MI355X (gfx950) with kinematic solids";
  py::register_exception<NotImplemented>(m, "NotImplementedInThisBuild", PyExc_NotImplementedError);

  py::module_ fem = m.def_submodule("fem", "FEM related classes");
  py::class_<FemSolverOptions>(fem, "FemSolverOptions")
      .def(py::init<>())
      .def_readwrite("optimizer_type", &FemSolverOptions::optimizer_type)
      .def_readwrite("iterations", &FemSolverOptions::iterations)
      .def_readwrite("force_density_abs_tol", &FemSolverOptions::force_density_abs_tol)
      .def_readwrite("grad_check", &FemSolverOptions::grad_check)
      .def_readwrite("thread_ct", &FemSolverOptions::thread_ct)
      .def_readwrite("substeps", &FemSolverOptions::substeps)
      .def_readwrite("omega", &FemSolverOptions::omega)
      .def_readwrite("vbd_iterations", &FemSolverOptions::vbd_iterations)
      .def_readwrite("ls_max_iter", &FemSolverOptions::ls_max_iter)
      .def_readwrite("ls_beta", &FemSolverOptions::ls_beta)
      .def_readwrite("ls_alpha", &FemSolverOptions::ls_alpha)
      .def_readwrite("linear_solver_type", &FemSolverOptions::linear_solver_type);
  py::class_<LBSControlConfig>(fem, "LBSControlConfig")
      .def(py::init<>())
      .def_readwrite("cnum", &LBSControlConfig::cnum)
      .def_readwrite("lbs_distance_type", &LBSControlConfig::lbs_distance_type)
      .def_readwrite("random_first", &LBSControlConfig::random_first)
      .def_readwrite("omega", &LBSControlConfig::omega)
      .def_readwrite("stiffness", &LBSControlConfig::stiffness);
  bind_solid<2>(fem, "SolidBodyConfig2D");
  bind_solid<3>(fem, "SolidBodyConfig3D");

  py::class_<BoundaryCondition>(m, "BoundaryCondition")
      .def(py::init<>())
      .def_readwrite("position", &BoundaryCondition::position)
      .def_readwrite("flag", &BoundaryCondition::flag);
  py::class_<DeviceOptions>(m, "DeviceOptions")
      .def(py::init<>())
      .def_readwrite("storage", &DeviceOptions::storage)
      .def_readwrite("stage_fp16", &DeviceOptions::stage_fp16)
      .def_readwrite("periodic", &DeviceOptions::periodic)
      .def_readwrite("fast_math", &DeviceOptions::fast_math)
      .def_readwrite("device", &DeviceOptions::device);
  m.def("flag_from_name", &flag_from_name, py::arg("name"), "LbmNodeFlag ordinal of a boundary type name.");

  bind_params<3>(m, "SimulationParameters3D");
  bind_params<2>(m, "SimulationParameters2D");
  bind_config<2>(m, "Config2D");
  bind_config<3>(m, "Config3D");
  bind_simulator<3>(m, "Simulator3D");
  bind_simulator<2>(m, "Simulator2D");
}
