// Test program for the C++ host interface: reads a scene JSON and prints what the
// builders produced; with a second argument "sim" it also builds the simulation and
// runs the given number of steps, printing the final position of vertex 0 (and, with a
// fourth argument, writing every final coordinate to that file; a fifth argument
// "i,j,...:w" adds attachment constraints first).
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "ProjectiveFriction.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    if (std::string(argv[1]) == "--stats" && argc >= 3) {  // computeAndWriteStats on fixed vectors
        std::ofstream of(argv[2]);
        computeAndWriteStats("Empty", of, {});
        computeAndWriteStats("One", of, {2.5});
        computeAndWriteStats("Three", of, {3., 1., 2.});
        computeAndWriteStats("Negative", of, {-1., -2.});
        return 0;
    }
    try {
        json::Value doc = getJSONDocumentFromFilename(argv[1]);
        PhysicScene scene = getSceneFromJSONDocument(doc);
        SimulationParameters p = getSimulationParametersFromJSONDocument(doc);
        StepCallBacks cb = getLogCallBacksFromJSONDocument(doc);
        auto m = scene.getMesh(0);
        std::cout << "vertices " << m->getNumberOfVertices() << "\n";
        std::cout << "triangles " << m->getNumberOfTriangles() << "\n";
        std::cout << "obstacles " << scene.getObstacles().size() << "\n";
        std::cout << "materials " << scene.getFrictionCoefficientTable().size() << "\n";
        std::cout.precision(17);
        std::cout << "timestep " << p.time_step << "\n";
        std::cout << "iterations " << p.local_global_iterations << "\n";
        std::cout << "self " << p.self_collision << " " << p.self_collision_tolerance << "\n";
        std::cout << "damping " << p.air_damping << "\n";
        for (std::size_t i = 0; i < scene.getObstacles().size(); ++i) {
            const pdfc_obstacle& o = scene.getObstacles()[i];
            std::cout << "obstacle" << i << " " << o.type << " " << o.material << " " << o.tolerance << " "
                      << o.n_mesh_vertices << " " << o.n_mesh_triangles << " " << o.n_frames;
            if (o.type == PDFC_OBSTACLE_BOX)
                std::cout << " " << o.point[0] << " " << o.point[1] << " " << o.point[2] << " " << o.size[0] << " "
                          << o.size[1] << " " << o.size[2];
            if (o.type == PDFC_OBSTACLE_ROTATING_SPHERE)
                std::cout << " " << o.radius << " " << o.rotation_axis[0] << " " << o.rotation_axis[1] << " "
                          << o.rotation_axis[2] << " " << o.angular_speed << " " << o.rotation_start;
            if (o.type == PDFC_OBSTACLE_MESH) {
                const double* last = o.mesh_frames + (o.n_frames - 1) * 3 * o.n_mesh_vertices;
                std::cout << " " << o.mesh_triangles[3 * o.n_mesh_triangles - 1] << " "
                          << (o.frame_times ? o.frame_times[o.n_frames - 1] : -1.0) << " " << last[0] << " "
                          << last[1] << " " << last[2];
            }
            std::cout << "\n";
        }
        auto q = m->getVertexPosition(1);
        std::cout << "v1 " << q[0] << " " << q[1] << " " << q[2] << "\n";
        if (argc >= 4 && std::string(argv[2]) == "sim") {
            if (argc >= 6) {  // "i,j,...:w": attach those vertices at their rest positions, weight w
                std::string spec = argv[5];
                const double w = std::atof(spec.substr(spec.find(':') + 1).c_str());
                std::string list = spec.substr(0, spec.find(':'));
                std::size_t pos = 0;
                while (pos < list.size()) {
                    std::size_t e = list.find(',', pos);
                    if (e == std::string::npos) e = list.size();
                    std::size_t v = (std::size_t)std::atol(list.substr(pos, e - pos).c_str());
                    scene.addConstraint(std::make_shared<AttachmentConstraint>(w, v, m->getRestPositions().at(v)));
                    pos = e + 1;
                }
            }
            auto sim = getSimulation(scene, p);
            int steps = std::atoi(argv[3]);
            for (int s = 0; s < steps; ++s) {
                sim->step();
                cb((size_t)s, scene);
            }
            auto x = scene.getMesh(0)->getVertexPosition(0);
            std::cout << "final0 " << x[0] << " " << x[1] << " " << x[2] << "\n";
            if (argc >= 5) {  // all final positions, one coordinate per line
                std::ofstream of(argv[4]);
                of.precision(17);
                for (double c : scene.getGeneralizedPositions()) of << c << "\n";
            }
        }
    } catch (const std::invalid_argument& e) {
        std::cout << "invalid_argument: " << e.what() << "\n";
        return 3;
    } catch (const std::exception& e) {
        std::cout << "error: " << e.what() << "\n";
        return 4;
    }
    return 0;
}
