// getSimulation and the GPU-backed SimulationFrictionEstimation
// (src/simulation_building.cpp:80-98, src/physics/SimulationFrictionEstimation.cpp).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <iterator>
#include <stdexcept>

#include "ProjectiveFriction.hpp"

namespace {
void fail_stop(const char* where) {
    std::cerr << "pdfc: " << where << " failed: " << pdfc_last_error() << std::endl;
    std::abort();
}
}  // namespace

SimulationBase::SimulationBase(PhysicScene& scene, double time_step, std::size_t iteration_number,
                               double air_damping, bool handle_self_collision, double self_collision_tolerance)
    : m_scene(scene),
      m_time_step(time_step),
      m_iteration_number(iteration_number),
      m_damping_coefficient(air_damping),
      m_handle_self_collision(handle_self_collision),
      m_self_collision_tolerance(self_collision_tolerance),
      m_current_step_error(iteration_number, 0.),
      m_current_step_duration(iteration_number, IterationDuration::zero()),
      m_current_step_number_of_collisions(iteration_number, 0) {}

// SimulationBase::step (src/physics/SimulationBase.cpp:103-173), hook by hook.
void SimulationBase::step() {
    initializeStep();
    updateCollision();
    const auto t0 = std::chrono::high_resolution_clock::now();
    m_current_step_error.assign(m_iteration_number, 0.);
    m_current_step_duration.assign(m_iteration_number, IterationDuration::zero());
    for (std::size_t i = 0; i < m_iteration_number; ++i) {
        buildIteration();
        m_current_step_error[i] = solve();
        m_current_step_duration[i] = std::chrono::high_resolution_clock::now() - t0;
        updateIterationData();
    }
    updateScene();
    m_current_time += m_time_step;
}

SimulationFrictionEstimationHIP::SimulationFrictionEstimationHIP(PhysicScene& scene, const SimulationParameters& p)
    : SimulationBase(scene, p.time_step, p.local_global_iterations, p.air_damping, p.self_collision,
                     p.self_collision_tolerance) {
    auto mesh = scene.getMesh(0);
    const auto& rest = mesh->getRestPositions();
    const auto& tris = mesh->getTriangles();
    std::vector<double> rest_flat, cur_flat = scene.getGeneralizedPositions();
    for (const auto& x : rest) rest_flat.insert(rest_flat.end(), x.begin(), x.end());
    std::vector<int32_t> tri_flat;
    for (const auto& t : tris) tri_flat.insert(tri_flat.end(), t.begin(), t.end());
    const auto& table = scene.getFrictionCoefficientTable();
    std::vector<double> fr;
    for (const auto& row : table) fr.insert(fr.end(), row.begin(), row.end());
    const auto& obs = scene.getObstacles();
    pdfc_scene sc{};
    sc.n_vertices = rest.size();
    sc.rest_positions = rest_flat.data();
    sc.positions = cur_flat.data();
    sc.n_triangles = tris.size();
    sc.triangles = tri_flat.data();
    sc.area_density = mesh->getPhysicParameters().area_density;
    sc.bending = mesh->getPhysicParameters().bend;
    sc.stretching = mesh->getPhysicParameters().stretch;
    sc.material = (int32_t)mesh->getMaterialIdentifier();
    sc.n_obstacles = obs.size();
    sc.obstacles = obs.data();
    sc.n_materials = table.size();
    sc.friction = fr.data();
    std::vector<int32_t> att_v;
    std::vector<double> att_p, att_w;
    for (const auto& c : scene.getUserConstraints()) {
        att_v.push_back((int32_t)c->index);
        att_p.insert(att_p.end(), c->position.begin(), c->position.end());
        att_w.push_back(c->weight);
    }
    sc.n_attachments = att_v.size();
    sc.attachment_vertices = att_v.data();
    sc.attachment_positions = att_p.data();
    sc.attachment_weights = att_w.data();
    pdfc_params prm{};
    prm.time_step = p.time_step;
    prm.local_global_iterations = (int32_t)p.local_global_iterations;
    prm.self_collision = p.self_collision ? 1 : 0;
    prm.self_collision_tolerance = p.self_collision_tolerance;
    for (int k = 0; k < 3; ++k) prm.gravity[k] = p.gravity[k];
    prm.air_damping = p.air_damping;
    prm.precision = p.precision;
    prm.refinement_sweeps = p.refinement_sweeps;
    prm.warm_start = p.warm_start;
    prm.record_iterations = 1;
    int rc = pdfc_create(&sc, &prm, &m_sim);
    if (rc != PDFC_OK) {
        if (rc == PDFC_ERR_INVALID) throw std::invalid_argument(pdfc_last_error());
        throw std::runtime_error(pdfc_last_error());
    }
    std::vector<double> v0 = scene.getGeneralizedSpeeds();
    if (pdfc_set_state(m_sim, nullptr, v0.data()) != PDFC_OK) throw std::runtime_error(pdfc_last_error());
    pdfc_info info{};
    if (pdfc_get_info(m_sim, &info) != PDFC_OK) throw std::runtime_error(pdfc_last_error());
    m_lhs_times.push_back(1e6 * info.lhs_seconds);
    m_cholesky_times.push_back(1e6 * info.factor_seconds);
    m_phase_timing = p.phase_timing;
    if (m_phase_timing && pdfc_set_kernel_timing(m_sim, 1) != PDFC_OK) throw std::runtime_error(pdfc_last_error());
}

SimulationFrictionEstimationHIP::~SimulationFrictionEstimationHIP() { pdfc_destroy(m_sim); }

// initializeStep + the device's detection (SimulationSpeed.cpp:62-76, SimulationBase.cpp:333-413,
// SimulationFrictionEstimation.cpp:107-159).
void SimulationFrictionEstimationHIP::initializeStep() {
    if (pdfc_step_begin(m_sim) != PDFC_OK) fail_stop("step_begin");
}

// One local-global iteration on the device (buildIteration + solve); SimulationSpeed::solve
// returns 0 (SimulationSpeed.cpp:172-187), so the per-iteration "errors" are zeros.
double SimulationFrictionEstimationHIP::solve() {
    if (pdfc_step_iterate(m_sim, 1) != PDFC_OK) fail_stop("step_iterate");
    return 0.;
}

// updateScene (SimulationFrictionEstimation.cpp:64-87) and the stats bookkeeping of
// SimulationBase::step / baseUpdateColisions (SimulationBase.cpp:103-173, 333-413), in device
// time: the step, its prologue (detection), and per iteration its mean. With phase timing the
// local kernel (base right-hand side and obstacle friction, fused), the self-contact levels and
// the substitution are timed apart as well.
void SimulationFrictionEstimationHIP::updateScene() {
    if (pdfc_step_end(m_sim) != PDFC_OK) fail_stop("step_end");
    pdfc_phase_stats ps{};
    pdfc_info info{};
    if (pdfc_get_phase_stats(m_sim, &ps) != PDFC_OK || pdfc_get_info(m_sim, &info) != PDFC_OK) fail_stop("stats");
    m_step_times.push_back(ps.step_us);
    const std::size_t K = ps.iterations > 0 ? (std::size_t)ps.iterations : m_iteration_number;
    if (ps.timed_steps > 0) {
        m_collision_detection_times.push_back(ps.collision_us);
        for (std::size_t k = 0; k < K; ++k) m_iteration_times.push_back(ps.iterations_us / (double)std::max<std::size_t>(1, K));
    }
    if (info.n_contacts > 0) m_collision_numbers.push_back((double)info.n_contacts);
    if (m_handle_self_collision) {
        m_self_collision_detection_times.push_back(ps.detection_us);
        if (info.n_self_contacts > 0) {
            m_self_collision_numbers.push_back((double)info.n_self_contacts);
            m_self_collision_ordering_times.push_back(ps.order_us);
        }
    }
    if (m_phase_timing) {
        pdfc_kernel_times kt{};
        if (pdfc_get_kernel_times(m_sim, &kt) != PDFC_OK) fail_stop("kernel_times");
        // device time of all the step's substitution launches (per-launch mean times launches)
        const double solver = (kt.forward_us * (double)kt.forward_launches +
                               kt.backward_us * (double)kt.backward_launches +
                               kt.top_us * (double)kt.top_launches) / (double)K;
        for (std::size_t k = 0; k < K; ++k) {
            m_global_times.push_back(solver);
            m_rhs_base_times.push_back(kt.local_us);
            m_rhs_times.push_back(kt.local_us + (kt.self_iterations > 0 ? kt.self_us : 0.));
            if (kt.self_iterations > 0) m_self_friction_times.push_back(kt.self_us);
        }
        if (pdfc_set_kernel_timing(m_sim, 1) != PDFC_OK) fail_stop("kernel_timing");  // next step's sums
    }
    m_current_step_error.assign(K, 0.);
    m_current_step_number_of_collisions.assign(K, (std::size_t)info.n_contacts);
    m_current_step_duration.assign(K, IterationDuration::zero());
    for (std::size_t k = 0; k < K; ++k) {  // time since the step began, after iteration k
        const double us = ps.collision_us + (double)(k + 1) * ps.iterations_us / (double)std::max<std::size_t>(1, K);
        m_current_step_duration[k] =
            std::chrono::duration_cast<IterationDuration>(std::chrono::duration<double, std::micro>(us));
    }
    if (m_mirror) mirrorState();
}

void SimulationFrictionEstimationHIP::mirrorState() {
    const std::size_t n3 = 3 * m_scene.getNumberOfVertices();
    m_q.resize(n3);
    m_v.resize(n3);
    if (pdfc_get_state(m_sim, m_q.data(), m_v.data()) != PDFC_OK) fail_stop("get_state");
    m_scene.setGeneralizedPositions(m_q);
    m_scene.setGeneralizedSpeeds(m_v);
}

void SimulationFrictionEstimationHIP::saveCheckpoint(const std::string& path) const {
    std::size_t n = 0;
    int rc = pdfc_save_checkpoint(m_sim, nullptr, 0, &n);
    if (rc != PDFC_ERR_CAPACITY && rc != PDFC_OK) throw std::runtime_error(pdfc_last_error());
    std::vector<unsigned char> buf(n);
    if (pdfc_save_checkpoint(m_sim, buf.data(), buf.size(), &n) != PDFC_OK) throw std::runtime_error(pdfc_last_error());
    std::ofstream of(path, std::ios::binary);
    of.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)n);
    if (!of) throw std::runtime_error("cannot write checkpoint " + path);
}

void SimulationFrictionEstimationHIP::loadCheckpoint(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::invalid_argument("cannot open checkpoint " + path);
    std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    int rc = pdfc_load_checkpoint(m_sim, buf.data(), buf.size());
    if (rc == PDFC_ERR_INVALID) throw std::invalid_argument(pdfc_last_error());
    if (rc != PDFC_OK) throw std::runtime_error(pdfc_last_error());
    mirrorState();
}

std::size_t SimulationFrictionEstimationHIP::getStepsDone() const {
    pdfc_info info{};
    if (pdfc_get_info(m_sim, &info) != PDFC_OK) throw std::runtime_error(pdfc_last_error());
    return (std::size_t)info.steps_done;
}

// The sections of SimulationFrictionEstimation::exportStats in its order
// (src/physics/SimulationFrictionEstimation.cpp:755-771, with SimulationBase::exportStats,
// src/physics/SimulationBase.cpp:473-494, in the middle). LHS: host build of A^T A / K;
// Cholesky: host factorisation; iteration and collision-detection times: device time per
// step (HIP events); self-collision detection: host round trip and pairing; ordering: host
// level ordering. With phase_timing, solver (the substitution), full RHS (local kernel plus
// self-contact levels), RHS base (the local kernel, which fuses the obstacle friction) and
// self collision times come from device events around those launches; the obstacle friction
// alone is not timed apart from the local kernel, so "RHS friction overhead" is "(No data)".
void SimulationFrictionEstimationHIP::exportStats(std::ofstream& of) const noexcept {
    computeAndWriteStats("LHS computation time (µs)", of, m_lhs_times);
    computeAndWriteStats("Cholesky computation time (µs)", of, m_cholesky_times);
    computeAndWriteStats("Iteration time (µs)", of, m_iteration_times);
    computeAndWriteStats("Collision detection time (µs)", of, m_collision_detection_times);
    if (m_handle_self_collision) {
        computeAndWriteStats("Number of collisions", of, m_collision_numbers);
        computeAndWriteStats("Self collision detection time (µs)", of, m_self_collision_detection_times);
    }
    computeAndWriteStats("Number of self collisions", of, m_self_collision_numbers);
    computeAndWriteStats("Solver time (µs)", of, m_global_times);
    computeAndWriteStats("Full RHS computation time (µs)", of, m_rhs_times);
    computeAndWriteStats("Step time (µs)", of, m_step_times);
    computeAndWriteStats("RHS base computation time (µs)", of, m_rhs_base_times);
    computeAndWriteStats("RHS friction overhead (µs)", of, m_friction_times);
    if (m_handle_self_collision) {
        computeAndWriteStats("Self collision time (µs)", of, m_self_friction_times);
        computeAndWriteStats("Self collision ordering time (µs)", of, m_self_collision_ordering_times);
    }
}

std::unique_ptr<SimulationBase> getSimulation(PhysicScene& scene, const SimulationParameters& params) {
    return std::make_unique<SimulationFrictionEstimationHIP>(scene, params);
}
