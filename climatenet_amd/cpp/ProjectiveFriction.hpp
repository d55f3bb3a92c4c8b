// C++ host interface of the simulation, mirroring the reference's public API
// (guibquatre/ClimateNet) on top of the pdfc C ABI:
//
//   PhysicScene, PhysicMesh          include/physics/PhysicScene.hpp, PhysicMesh.hpp
//   SimulationParameters             include/simulation_building.hpp:34-45
//   SimulationBase::step/exportStats include/physics/SimulationBase.hpp:70,117
//   getLastStepErrors/Durations/NumberCollisions  SimulationBase.hpp:100-111
//   getSimulation                    include/simulation_building.hpp:70-71
//   getSceneFromJSONDocument         include/scene_building.hpp:38-39
//   getSimulationParametersFromJSONDocument  include/simulation_building.hpp:55-56
//   StepCallBacks                    include/StepCallBack.hpp:9-27
//   getLogCallBacksFromJSONDocument  include/log_building.hpp:21-22
//   read_obj                         include/geometry/io.hpp / src/geometry/io.cpp:65-121
//   computeAndWriteStats             template/physics/io.tpp:23-63
//
// Construction errors throw std::invalid_argument (as the reference's JSON
// builders do); the step's hooks abort with the library's message on a device
// failure, the reference's fail-stop behaviour (its step methods are noexcept).
//
// One deliberate departure: the JSON document type is this build's own parser's
// (json::Value, json.hpp) where the reference takes rapidjson::Document; rapidjson is a
// third-party header library vendored inside the reference (extern/rapidjson) and is not
// part of this build. Key names, defaults and error messages are the reference's.
#pragma once

#include <array>
#include <chrono>
#include <cstddef>
#include <fstream>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "json.hpp"
#include "pdfc.h"

// ---------------------------------------------------------------- geometry IO
// Reads vertices ("v x y z", stored as float like tinyobjloader's real_t = float)
// and faces ("f ...", fan-triangulated, 1-based or negative indices, v/vt/vn forms).
bool read_obj(const std::string& filename, std::vector<std::array<double, 3>>& positions,
              std::vector<std::array<int, 3>>& triangles);

// ------------------------------------------------------------------- scene
struct PhysicParameters {
    double area_density = 0.;
    double bend = 0.;
    double stretch = 0.;
};

class PhysicMesh {
public:
    PhysicMesh(std::vector<std::array<double, 3>> rest, std::vector<std::array<int, 3>> triangles,
               const PhysicParameters& params, std::size_t material_identifier);
    std::size_t getNumberOfVertices() const { return m_rest.size(); }
    std::size_t getNumberOfTriangles() const { return m_triangles.size(); }
    std::size_t getMaterialIdentifier() const { return m_material; }
    const PhysicParameters& getPhysicParameters() const { return m_params; }
    std::array<double, 3> getVertexPosition(std::size_t i) const { return m_q[i]; }
    std::array<double, 3> getVertexSpeed(std::size_t i) const { return m_v[i]; }
    const std::vector<std::array<double, 3>>& getRestPositions() const { return m_rest; }
    const std::vector<std::array<int, 3>>& getTriangles() const { return m_triangles; }
    void setPositions(const std::vector<std::array<double, 3>>& q);
    void setSpeeds(const std::vector<std::array<double, 3>>& v);
    // OBJ frame ("v x y z" with the stream's default precision, "f i j k" 1-based),
    // as Mesh::writeObj (src/geometry/Mesh.cpp:95-109).
    void writeObj(std::ofstream& of) const;

private:
    std::vector<std::array<double, 3>> m_rest, m_q, m_v;
    std::vector<std::array<int, 3>> m_triangles;
    PhysicParameters m_params;
    std::size_t m_material;
};

// AttachmentConstraint (include/physics/AttachmentConstraint.hpp:33-70): one row of
// weight `weight` on vertex `index`, projected on weight * `position`. Added with
// PhysicScene::addConstraint after the scene is read; the JSON format has no key for it.
struct AttachmentConstraint {
    AttachmentConstraint(double weight, std::size_t index, const std::array<double, 3>& position)
        : weight(weight), index(index), position(position) {}
    double weight;
    std::size_t index;
    std::array<double, 3> position;
};

class PhysicScene {
public:
    using FrictionCoefficientTable = std::vector<std::vector<double>>;
    explicit PhysicScene(FrictionCoefficientTable friction);
    void addMesh(std::shared_ptr<PhysicMesh> mesh);
    void addObstacle(const pdfc_obstacle& obstacle);
    // Triangle-mesh obstacle (StaticMesh: one frame; MovingMesh: key frames and their
    // times). The scene owns the arrays; the stored pdfc_obstacle points into them.
    struct MeshObstacleData {
        std::vector<int32_t> triangles;  // 3T, frame 0's faces
        std::vector<double> frames;      // n_frames x 3 n_mesh_vertices
        std::vector<double> times;       // n_frames (empty for a static mesh)
    };
    void addMeshObstacle(pdfc_obstacle obstacle, std::shared_ptr<const MeshObstacleData> data);
    // PhysicScene::addConstraint (src/physics/PhysicScene.cpp:82-86); call before getSimulation.
    void addConstraint(const std::shared_ptr<AttachmentConstraint>& constraint);
    const std::vector<std::shared_ptr<AttachmentConstraint>>& getUserConstraints() const { return m_user_constraints; }
    void finalize();
    std::shared_ptr<PhysicMesh> getMesh(std::size_t index) { return m_meshes.at(index); }
    std::shared_ptr<const PhysicMesh> getMesh(std::size_t index) const { return m_meshes.at(index); }
    std::size_t getNumberOfVertices() const;
    const FrictionCoefficientTable& getFrictionCoefficientTable() const { return m_friction; }
    const std::vector<pdfc_obstacle>& getObstacles() const { return m_obstacles; }
    // Generalized vectors (3n, x y z per vertex), as PhysicScene::getGeneralizedPositions/Speeds.
    std::vector<double> getGeneralizedPositions() const;
    std::vector<double> getGeneralizedSpeeds() const;
    void setGeneralizedPositions(const std::vector<double>& q);
    void setGeneralizedSpeeds(const std::vector<double>& v);

private:
    FrictionCoefficientTable m_friction;
    std::vector<std::shared_ptr<PhysicMesh>> m_meshes;
    std::vector<pdfc_obstacle> m_obstacles;
    std::vector<std::shared_ptr<const MeshObstacleData>> m_mesh_data;  // kept alive for m_obstacles
    std::vector<std::shared_ptr<AttachmentConstraint>> m_user_constraints;
};

// -------------------------------------------------------------- simulation
struct SimulationParameters {
    std::size_t max_output_frame = 0;
    std::size_t local_global_iterations = 0;
    double time_step = 0.;
    std::string simulation_type;
    std::size_t sim_steps_per_frame = 0;
    bool self_collision = false;
    double self_collision_tolerance = 0.;
    std::array<double, 3> gravity{{0., 0., 0.}};
    double air_damping = 0.;
    // Build-specific options (not in the reference's JSON): precision, warm start, and device
    // timing of the step's phases for exportStats (events around every launch; measurement only).
    int precision = PDFC_FP64;
    int refinement_sweeps = 2;
    int warm_start = PDFC_WARM_PLAIN;
    bool phase_timing = false;
};

// SimulationBase (include/physics/SimulationBase.hpp:39-290). step() is not virtual: it runs
// the protected hooks in the reference's order (src/physics/SimulationBase.cpp:103-173):
// initializeStep, collision handling (updateCollision), then for each local-global iteration
// buildIteration, solve and updateIterationData, then updateScene. The reference detects the
// collisions on the host between initializeStep and updateCollision (baseUpdateColisions,
// SimulationBase.cpp:333-413); here detection is part of the device prologue.
class SimulationBase {
public:
    using IterationDuration = std::chrono::high_resolution_clock::duration;
    SimulationBase(PhysicScene& scene, double time_step, std::size_t iteration_number, double air_damping,
                   bool handle_self_collision, double self_collision_tolerance);
    virtual ~SimulationBase() = default;
    void step();
    virtual void exportStats(std::ofstream& of) const noexcept = 0;
    const std::vector<double>& getLastStepErrors() const noexcept { return m_current_step_error; }
    const std::vector<IterationDuration>& getLastStepDurations() const noexcept { return m_current_step_duration; }
    const std::vector<std::size_t>& getLastStepNumberCollisions() const noexcept {
        return m_current_step_number_of_collisions;
    }

protected:
    virtual void initializeStep() = 0;
    virtual void buildIteration() = 0;
    virtual void updateCollision() = 0;
    virtual double solve() = 0;
    virtual void updateScene() = 0;
    virtual void updateIterationData() noexcept {}

    PhysicScene& m_scene;
    const double m_time_step;
    double m_current_time = 0.;
    const std::size_t m_iteration_number;
    const double m_damping_coefficient;
    const bool m_handle_self_collision;
    const double m_self_collision_tolerance;
    std::vector<double> m_current_step_error;
    std::vector<IterationDuration> m_current_step_duration;
    std::vector<std::size_t> m_current_step_number_of_collisions;
};

// SimulationFrictionEstimation on the GPU. Its hooks drive the device step through the
// phase entry points: initializeStep -> pdfc_step_begin (per-step constant, obstacle and
// self-collision detection, pairing, levels); buildIteration -> nothing (the local step is
// fused into the device iteration); solve -> pdfc_step_iterate(1); updateScene ->
// pdfc_step_end, and the scene's q and v are mirrored back (callbacks and OBJ output read
// them right after step(), src/main.cpp:91-105).
class SimulationFrictionEstimationHIP : public SimulationBase {
public:
    SimulationFrictionEstimationHIP(PhysicScene& scene, const SimulationParameters& params);
    ~SimulationFrictionEstimationHIP() override;
    void exportStats(std::ofstream& of) const noexcept override;
    pdfc_sim* handle() const { return m_sim; }
    // Skip the per-step state copy when no caller reads the scene between steps.
    void setMirrorState(bool on) { m_mirror = on; }
    // Binary checkpoint (pdfc_save_checkpoint / pdfc_load_checkpoint): positions, velocities,
    // obstacle clocks, warm-start reactions, step count. No counterpart in the reference.
    // Loading copies q and v into the scene. Errors throw std::runtime_error /
    // std::invalid_argument (a file that is not a checkpoint of this scene).
    void saveCheckpoint(const std::string& path) const;
    void loadCheckpoint(const std::string& path);
    std::size_t getStepsDone() const;

protected:
    void initializeStep() override;
    void buildIteration() override {}
    void updateCollision() override {}
    double solve() override;
    void updateScene() override;

private:
    pdfc_sim* m_sim = nullptr;
    bool m_mirror = true;
    bool m_phase_timing = false;
    // the vectors of SimulationSpeed/SimulationBase/SimulationFrictionEstimation that
    // exportStats writes; the ones this build does not measure stay empty ("(No data)")
    std::vector<double> m_lhs_times, m_cholesky_times, m_iteration_times, m_collision_detection_times;
    std::vector<double> m_collision_numbers, m_self_collision_detection_times, m_self_collision_numbers;
    std::vector<double> m_global_times, m_rhs_times, m_step_times, m_rhs_base_times, m_friction_times;
    std::vector<double> m_self_friction_times, m_self_collision_ordering_times;
    std::vector<double> m_q, m_v;
    void mirrorState();
};

std::unique_ptr<SimulationBase> getSimulation(PhysicScene& scene, const SimulationParameters& params);

// --------------------------------------------------------------- callbacks
class StepCallBacks {
public:
    using Callback = std::function<void(std::size_t, const PhysicScene&)>;
    void addCallback(const Callback& callback) { m_callbacks.push_back(callback); }
    void addCallback(Callback&& callback) { m_callbacks.push_back(std::move(callback)); }
    template <typename It>
    void addCallbacks(It begin, It end) {
        for (; begin != end; ++begin) addCallback(*begin);
    }
    void operator()(std::size_t step, const PhysicScene& scene) const {
        for (const auto& c : m_callbacks) c(step, scene);
    }

private:
    std::vector<Callback> m_callbacks;
};

// ------------------------------------------------------------------ JSON
json::Value getJSONDocumentFromFilename(const std::string& filename);
PhysicScene getSceneFromJSONDocument(const json::Value& document);
SimulationParameters getSimulationParametersFromJSONDocument(const json::Value& document);
StepCallBacks getLogCallBacksFromJSONDocument(const json::Value& document);

// ------------------------------------------------------------------ stats
void computeAndWriteStats(const std::string& name, std::ofstream& of, const std::vector<double>& values);
