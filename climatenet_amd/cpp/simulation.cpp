// getSimulation and the GPU-backed SimulationFrictionEstimation
// (src/simulation_building.cpp:80-98, src/physics/SimulationFrictionEstimation.cpp).
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

SimulationFrictionEstimationHIP::SimulationFrictionEstimationHIP(PhysicScene& scene, const SimulationParameters& p)
    : m_scene(scene), m_iterations(p.local_global_iterations) {
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
    m_self = p.self_collision;
    // SimulationSpeed::solve returns 0 (SimulationSpeed.cpp:172-187): the per-iteration
    // "errors" of the reference are zeros
    m_current_step_error.assign(m_iterations, 0.);
    m_current_step_duration.assign(m_iterations, IterationDuration::zero());
    m_current_step_number_of_collisions.assign(m_iterations, 0);
}

SimulationFrictionEstimationHIP::~SimulationFrictionEstimationHIP() { pdfc_destroy(m_sim); }

void SimulationFrictionEstimationHIP::step() noexcept {
    if (pdfc_step(m_sim, 1) != PDFC_OK) fail_stop("step");
    pdfc_phase_stats ps{};
    pdfc_info info{};
    if (pdfc_get_phase_stats(m_sim, &ps) != PDFC_OK || pdfc_get_info(m_sim, &info) != PDFC_OK) fail_stop("stats");
    // SimulationBase::step / baseUpdateColisions bookkeeping (SimulationBase.cpp:103-173,
    // 333-413): device times per step; each iteration gets the step's mean iteration time
    m_step_times.push_back(ps.step_us);
    const std::size_t K = ps.steps > 0 ? (std::size_t)(ps.iterations / ps.steps) : m_iterations;
    if (ps.timed_steps > 0) {
        m_collision_detection_times.push_back(ps.collision_us);
        for (std::size_t k = 0; k < K; ++k) m_iteration_times.push_back(ps.iterations_us / (double)std::max<std::size_t>(1, K));
    }
    if (info.n_contacts > 0) m_collision_numbers.push_back((double)info.n_contacts);
    if (m_self) {
        m_self_collision_detection_times.push_back(ps.detection_us);
        if (info.n_self_contacts > 0) m_self_collision_numbers.push_back((double)info.n_self_contacts);
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
// step (HIP events); self-collision detection: host round trip and pairing. The solver,
// RHS and friction phases run fused on the device and are not timed apart: "(No data)".
void SimulationFrictionEstimationHIP::exportStats(std::ofstream& of) const noexcept {
    computeAndWriteStats("LHS computation time (µs)", of, m_lhs_times);
    computeAndWriteStats("Cholesky computation time (µs)", of, m_cholesky_times);
    computeAndWriteStats("Iteration time (µs)", of, m_iteration_times);
    computeAndWriteStats("Collision detection time (µs)", of, m_collision_detection_times);
    if (m_self) {
        computeAndWriteStats("Number of collisions", of, m_collision_numbers);
        computeAndWriteStats("Self collision detection time (µs)", of, m_self_collision_detection_times);
    }
    computeAndWriteStats("Number of self collisions", of, m_self_collision_numbers);
    computeAndWriteStats("Solver time (µs)", of, m_global_times);
    computeAndWriteStats("Full RHS computation time (µs)", of, m_rhs_times);
    computeAndWriteStats("Step time (µs)", of, m_step_times);
    computeAndWriteStats("RHS base computation time (µs)", of, m_rhs_base_times);
    computeAndWriteStats("RHS friction overhead (µs)", of, m_friction_times);
    if (m_self) {
        computeAndWriteStats("Self collision time (µs)", of, m_self_friction_times);
        computeAndWriteStats("Self collision ordering time (µs)", of, m_self_collision_ordering_times);
    }
}

std::unique_ptr<SimulationBase> getSimulation(PhysicScene& scene, const SimulationParameters& params) {
    return std::make_unique<SimulationFrictionEstimationHIP>(scene, params);
}
