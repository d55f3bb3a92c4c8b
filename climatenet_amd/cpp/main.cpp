// pdfc_main <scene.json> [--checkpoint-every K] [--resume FILE] [--phase-timing]: the reference's driver
// (src/main.cpp:40-124) on the GPU simulation: output directory dataN/ with a copy of
// the scene, one OBJ frame every "Steps per frame" steps, log callbacks after each step,
// stats.txt. The options have no counterpart in the reference: --checkpoint-every writes
// dataN/checkpoint.bin every K steps and after the last one; --resume loads a checkpoint
// of the same scene and continues from its step (frames keep their numbers); --phase-timing
// times the device phases of every step (events around each launch, a synchronisation per
// step) so stats.txt also gets the solver, RHS and self-collision sections.
#include <sys/stat.h>

#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <list>
#include <sstream>
#include <string>

#include "ProjectiveFriction.hpp"

namespace {
bool is_dir(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}
std::string basename_of(const std::string& p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}
}  // namespace

int main(int argc, char* argv[]) {
    if (argc <= 1) {
        std::cerr << "Not enough arguments - 1 argument expected" << std::endl;
        std::cerr << " (1) Path of the json configuration file " << std::endl;
        return EXIT_FAILURE;
    }
    const std::string scene_path(argv[1]);
    std::size_t checkpoint_every = 0;
    std::string resume;
    bool phase_timing = false;
    for (int a = 2; a < argc; ++a) {
        const std::string opt = argv[a];
        if (opt == "--checkpoint-every" && a + 1 < argc) {
            checkpoint_every = (std::size_t)std::strtoull(argv[++a], nullptr, 10);
        } else if (opt == "--resume" && a + 1 < argc) {
            resume = argv[++a];
        } else if (opt == "--phase-timing") {
            phase_timing = true;
        } else {
            std::cerr << "Unknown option " << opt << std::endl;
            return EXIT_FAILURE;
        }
    }
    std::cout << "# Initializing." << std::endl;
    std::string out_dir;
    for (unsigned k = 0;; ++k) {
        out_dir = "data" + std::to_string(k);
        if (!is_dir(out_dir)) break;
    }
    mkdir(out_dir.c_str(), 0755);
    mkdir((out_dir + "/output").c_str(), 0755);
    {
        std::ifstream src(scene_path, std::ios::binary);
        std::ofstream dst(out_dir + "/" + basename_of(scene_path), std::ios::binary);
        dst << src.rdbuf();
    }
    try {
        json::Value doc = getJSONDocumentFromFilename(scene_path);
        PhysicScene scene = getSceneFromJSONDocument(doc);
        auto mesh = scene.getMesh(0);
        SimulationParameters params = getSimulationParametersFromJSONDocument(doc);
        params.phase_timing = phase_timing;
        auto simulation = getSimulation(scene, params);
        StepCallBacks log_callback = getLogCallBacksFromJSONDocument(doc);
        const size_t max_steps = params.max_output_frame * params.sim_steps_per_frame;
        auto* hip_sim = dynamic_cast<SimulationFrictionEstimationHIP*>(simulation.get());
        size_t first = 0;
        if (!resume.empty()) {
            hip_sim->loadCheckpoint(resume);
            first = hip_sim->getStepsDone();
            std::cout << "# Resumed from " << resume << " at step " << first << std::endl;
        }
        std::cout << "# Initialization done. Starting simulation, " << max_steps << " timesteps" << std::endl;
        std::list<int> milestones;
        for (int i = 5; i <= 100; i += 5) milestones.push_back(i);
        for (size_t i = first; i < max_steps; ++i) {
            std::cout << "Time Step: " << i << std::endl;
            simulation->step();
            log_callback(i, scene);
            if (i % params.sim_steps_per_frame == 0) {
                std::ostringstream name;
                name << out_dir << "/output/out_" << std::setw(6) << std::setfill('0')
                     << i / params.sim_steps_per_frame << ".obj";
                std::ofstream of(name.str());
                mesh->writeObj(of);
            }
            if (checkpoint_every > 0 && ((i + 1) % checkpoint_every == 0 || i + 1 == max_steps))
                hip_sim->saveCheckpoint(out_dir + "/checkpoint.bin");
            if (!milestones.empty() && i >= milestones.front() * (max_steps - 1) / 100.) {
                std::cerr << milestones.front() << "% done" << std::endl;
                milestones.pop_front();
            }
        }
        std::cout << "# Simulation done. Writing stats." << std::endl;
        std::ofstream of(out_dir + "/stats.txt");
        simulation->exportStats(of);
        std::cout << "# Writing done. " << std::endl;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
