// JSON scene and parameter readers, OBJ reader, scene containers, log callbacks
// and the stats.txt writer of the C++ host interface.
//   src/scene_building.cpp, src/simulation_building.cpp:37-78, src/log_building.cpp,
//   src/geometry/io.cpp:65-121, template/physics/io.tpp:23-63
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>

#include "ProjectiveFriction.hpp"

namespace {

const json::Value& require(const json::Value& obj, const char* key) {
    const json::Value* v = obj.FindMember(key);
    if (!v) throw std::invalid_argument(std::string("JSON reader - Error : cannot found key ") + key);
    return *v;
}
[[noreturn]] void type_error(const char* key) {
    throw std::invalid_argument(std::string("JSON reader - Error : type error with key ") + key);
}
double require_double(const json::Value& obj, const char* key) {
    const json::Value& v = require(obj, key);
    if (!v.IsDouble()) type_error(key);
    return v.GetDouble();
}
unsigned require_uint(const json::Value& obj, const char* key) {
    const json::Value& v = require(obj, key);
    if (!v.IsUint()) type_error(key);
    return v.GetUint();
}
bool require_bool(const json::Value& obj, const char* key) {
    const json::Value& v = require(obj, key);
    if (!v.IsBool()) type_error(key);
    return v.GetBool();
}
std::string require_string(const json::Value& obj, const char* key) {
    const json::Value& v = require(obj, key);
    if (!v.IsString()) type_error(key);
    return v.GetString();
}
const json::Value& require_array(const json::Value& obj, const char* key) {
    const json::Value& v = require(obj, key);
    if (!v.IsArray()) type_error(key);
    return v;
}
std::array<double, 3> vector3(const json::Value& a) {
    if (a.Size() != 3u) throw std::invalid_argument("JSON reader - Error : Array representing vector must be of size 3");
    std::array<double, 3> r{};
    for (int i = 0; i < 3; ++i) {
        if (!a[i].IsDouble())
            throw std::invalid_argument("JSON reader - Error : Array representing vector must contain doubles");
        r[i] = a[i].GetDouble();
    }
    return r;
}

}  // namespace

// ------------------------------------------------------------------ OBJ
bool read_obj(const std::string& filename, std::vector<std::array<double, 3>>& positions,
              std::vector<std::array<int, 3>>& triangles) {
    std::ifstream f(filename);
    if (!f) {
        std::cerr << "Cannot open file [" << filename << "]" << std::endl;
        return false;
    }
    positions.clear();
    triangles.clear();
    std::string line;
    std::vector<int> poly;
    while (std::getline(f, line)) {
        size_t i = line.find_first_not_of(" \t");
        if (i == std::string::npos) continue;
        if (line.compare(i, 2, "v ") == 0 || line.compare(i, 2, "v\t") == 0) {
            std::istringstream ss(line.substr(i + 2));
            std::array<double, 3> p{};
            for (int k = 0; k < 3; ++k) {
                std::string tok;
                if (!(ss >> tok)) return false;
                // stored as float, then widened (tinyobjloader real_t = float)
                p[k] = (double)(float)std::strtod(tok.c_str(), nullptr);
            }
            positions.push_back(p);
        } else if (line.compare(i, 2, "f ") == 0 || line.compare(i, 2, "f\t") == 0) {
            std::istringstream ss(line.substr(i + 2));
            std::string tok;
            poly.clear();
            while (ss >> tok) {
                long idx = std::strtol(tok.c_str(), nullptr, 10);
                if (idx == 0) return false;
                long v = idx > 0 ? idx - 1 : (long)positions.size() + idx;
                if (v < 0) return false;
                poly.push_back((int)v);
            }
            for (size_t k = 1; k + 1 < poly.size(); ++k) triangles.push_back({poly[0], poly[k], poly[k + 1]});
        }
    }
    for (const auto& t : triangles)
        for (int v : t)
            if (v >= (int)positions.size()) return false;
    return true;
}

// ---------------------------------------------------------------- scene
PhysicMesh::PhysicMesh(std::vector<std::array<double, 3>> rest, std::vector<std::array<int, 3>> triangles,
                       const PhysicParameters& params, std::size_t material_identifier)
    : m_rest(std::move(rest)), m_triangles(std::move(triangles)), m_params(params), m_material(material_identifier) {
    m_q = m_rest;
    m_v.assign(m_rest.size(), {0., 0., 0.});
}

void PhysicMesh::setPositions(const std::vector<std::array<double, 3>>& q) {
    if (q.size() != m_q.size()) throw std::invalid_argument("position count does not match the mesh");
    m_q = q;
}
void PhysicMesh::setSpeeds(const std::vector<std::array<double, 3>>& v) {
    if (v.size() != m_v.size()) throw std::invalid_argument("speed count does not match the mesh");
    m_v = v;
}
void PhysicMesh::writeObj(std::ofstream& of) const {
    for (const auto& v : m_q) of << "v " << v[0] << " " << v[1] << " " << v[2] << std::endl;
    for (const auto& t : m_triangles) of << "f " << t[0] + 1 << " " << t[1] + 1 << " " << t[2] + 1 << std::endl;
}

PhysicScene::PhysicScene(FrictionCoefficientTable friction) : m_friction(std::move(friction)) {}
void PhysicScene::addMesh(std::shared_ptr<PhysicMesh> mesh) { m_meshes.push_back(std::move(mesh)); }
void PhysicScene::addObstacle(const pdfc_obstacle& obstacle) { m_obstacles.push_back(obstacle); }
void PhysicScene::addConstraint(const std::shared_ptr<AttachmentConstraint>& constraint) {
    if (!constraint) throw std::invalid_argument("null constraint");
    if (constraint->index >= getNumberOfVertices()) throw std::invalid_argument("attachment vertex index out of range");
    if (!(constraint->weight > 0.)) throw std::invalid_argument("attachment weight must be positive");
    m_user_constraints.push_back(constraint);
}
void PhysicScene::addMeshObstacle(pdfc_obstacle ob, std::shared_ptr<const MeshObstacleData> data) {
    ob.mesh_triangles = data->triangles.data();
    ob.n_mesh_triangles = data->triangles.size() / 3;
    ob.n_frames = data->times.empty() ? 1 : data->times.size();
    ob.n_mesh_vertices = data->frames.size() / (3 * ob.n_frames);
    ob.mesh_frames = data->frames.data();
    ob.frame_times = data->times.empty() ? nullptr : data->times.data();
    m_mesh_data.push_back(std::move(data));
    m_obstacles.push_back(ob);
}
void PhysicScene::finalize() {
    if (m_meshes.size() != 1)
        throw std::invalid_argument("JSON reader - Error : exactly one dynamic mesh is supported");
    const std::size_t nm = m_friction.size();
    if (m_meshes[0]->getMaterialIdentifier() >= nm) throw std::invalid_argument("mesh material identifier out of range");
    for (const auto& o : m_obstacles)
        if ((std::size_t)o.material >= nm) throw std::invalid_argument("obstacle material identifier out of range");
}
std::size_t PhysicScene::getNumberOfVertices() const {
    std::size_t n = 0;
    for (const auto& m : m_meshes) n += m->getNumberOfVertices();
    return n;
}
std::vector<double> PhysicScene::getGeneralizedPositions() const {
    std::vector<double> q;
    for (const auto& m : m_meshes)
        for (std::size_t i = 0; i < m->getNumberOfVertices(); ++i) {
            auto p = m->getVertexPosition(i);
            q.insert(q.end(), p.begin(), p.end());
        }
    return q;
}
std::vector<double> PhysicScene::getGeneralizedSpeeds() const {
    std::vector<double> v;
    for (const auto& m : m_meshes)
        for (std::size_t i = 0; i < m->getNumberOfVertices(); ++i) {
            auto p = m->getVertexSpeed(i);
            v.insert(v.end(), p.begin(), p.end());
        }
    return v;
}
void PhysicScene::setGeneralizedPositions(const std::vector<double>& q) {
    std::size_t off = 0;
    for (auto& m : m_meshes) {
        std::vector<std::array<double, 3>> p(m->getNumberOfVertices());
        for (auto& x : p) {
            x = {q.at(off), q.at(off + 1), q.at(off + 2)};
            off += 3;
        }
        m->setPositions(p);
    }
}
void PhysicScene::setGeneralizedSpeeds(const std::vector<double>& v) {
    std::size_t off = 0;
    for (auto& m : m_meshes) {
        std::vector<std::array<double, 3>> p(m->getNumberOfVertices());
        for (auto& x : p) {
            x = {v.at(off), v.at(off + 1), v.at(off + 2)};
            off += 3;
        }
        m->setSpeeds(p);
    }
}

// ------------------------------------------------------------------ JSON
json::Value getJSONDocumentFromFilename(const std::string& filename) { return json::parse_file(filename); }

// src/scene_building.cpp:49-193
PhysicScene getSceneFromJSONDocument(const json::Value& doc) {
    if (!doc.IsObject()) throw std::invalid_argument("JSON reader - Error : root must be an object");
    const json::Value& fr = require_array(doc, "Friction coefficients");
    const std::size_t nm = fr.Size();
    PhysicScene::FrictionCoefficientTable table(nm, std::vector<double>(nm));
    for (std::size_t i = 0; i < nm; ++i) {
        if (!fr[i].IsArray() || fr[i].Size() != nm)
            throw std::invalid_argument("JSON reader - Error : Friction coefficients must be a square matrix");
        for (std::size_t j = 0; j < nm; ++j) {
            if (!fr[i][j].IsDouble())
                throw std::invalid_argument(
                    "JSON reader - Error : Friction coefficients must be a square matrix of double");
            table[i][j] = fr[i][j].GetDouble();
        }
    }
    PhysicScene scene(table);
    for (const json::Value& m : require_array(doc, "Meshes").array()) {
        std::string rest_file = require_string(m, "Obj filename");
        std::vector<std::array<double, 3>> rest, cur;
        std::vector<std::array<int, 3>> tris, tris2;
        if (!read_obj(rest_file, rest, tris)) throw std::invalid_argument("Could not load obj file: " + rest_file);
        PhysicParameters pp;
        pp.area_density = require_double(m, "Area Density");
        pp.bend = require_double(m, "Bending");
        pp.stretch = require_double(m, "Stretching");
        const std::size_t mat = require_uint(m, "Material Identifier");
        auto mesh = std::make_shared<PhysicMesh>(rest, tris, pp, mat);
        if (m.HasMember("Current state obj filename")) {
            std::string cur_file = require_string(m, "Current state obj filename");
            if (!read_obj(cur_file, cur, tris2)) throw std::invalid_argument("Could not load obj file: " + cur_file);
            mesh->setPositions(cur);
        }
        scene.addMesh(mesh);
    }
    for (const json::Value& o : require_array(doc, "Obstacles").array()) {
        std::string type = require_string(o, "Type");
        pdfc_obstacle ob{};
        if (type == "Sphere") {
            ob.type = PDFC_OBSTACLE_SPHERE;
            auto p = vector3(require_array(o, "Position"));
            for (int k = 0; k < 3; ++k) ob.point[k] = p[k];
            ob.radius = require_double(o, "Radius");
            ob.material = (int32_t)require_uint(o, "Material Identifier");
        } else if (type == "Plane") {
            ob.type = PDFC_OBSTACLE_PLANE;
            auto p = vector3(require_array(o, "Point"));
            auto nrm = vector3(require_array(o, "Normal"));
            for (int k = 0; k < 3; ++k) {
                ob.point[k] = p[k];
                ob.normal[k] = nrm[k];
            }
            ob.material = (int32_t)require_uint(o, "Material Identifier");
        } else if (type == "Mesh") {
            // src/scene_building.cpp:222-246 (getStaticMeshFromObject)
            const std::string path = require_string(o, "Path");
            std::vector<std::array<double, 3>> pos;
            std::vector<std::array<int, 3>> tris;
            if (!read_obj(path, pos, tris)) throw std::invalid_argument("Could not load obj file: " + path);
            auto data = std::make_shared<PhysicScene::MeshObstacleData>();
            for (const auto& t : tris) data->triangles.insert(data->triangles.end(), t.begin(), t.end());
            for (const auto& x : pos) data->frames.insert(data->frames.end(), x.begin(), x.end());
            ob.type = PDFC_OBSTACLE_MESH;
            ob.tolerance = require_double(o, "Threshold");
            ob.material = (int32_t)require_uint(o, "Material Identifier");
            scene.addMeshObstacle(ob, data);
            continue;
        } else if (type == "Moving Mesh") {
            // src/scene_building.cpp:248-277 (getMovingMeshFromObject): frame k is
            // <Obj Prefix><k zero-padded to Suffix Size>.obj at time k * Time Between Frame;
            // the faces are frame 0's (MovingMesh keeps the first frames' topology).
            const std::string prefix = require_string(o, "Obj Prefix");
            const std::size_t suffix = require_uint(o, "Suffix Size");
            const std::size_t nf = require_uint(o, "Number Frame");
            const double between = require_double(o, "Time Between Frame");
            ob.type = PDFC_OBSTACLE_MESH;
            ob.tolerance = require_double(o, "Collision tolerance");
            ob.material = (int32_t)require_uint(o, "Material Identifier");
            if (nf < 2) throw std::invalid_argument("JSON reader - Error : a moving mesh needs at least two frames");
            auto data = std::make_shared<PhysicScene::MeshObstacleData>();
            std::size_t nv = 0;
            for (std::size_t k = 0; k < nf; ++k) {
                std::ostringstream name;
                name << prefix;
                if (suffix > 0) name << std::setfill('0') << std::setw((int)suffix);
                name << k << ".obj";
                std::vector<std::array<double, 3>> pos;
                std::vector<std::array<int, 3>> tris;
                if (!read_obj(name.str(), pos, tris))
                    throw std::invalid_argument("Could not load obj file: " + name.str());
                if (k == 0) {
                    nv = pos.size();
                    for (const auto& t : tris) data->triangles.insert(data->triangles.end(), t.begin(), t.end());
                } else if (pos.size() != nv) {
                    throw std::invalid_argument("JSON reader - Error : moving mesh frame " + name.str() +
                                                " has a different number of vertices than frame 0");
                }
                for (const auto& x : pos) data->frames.insert(data->frames.end(), x.begin(), x.end());
                data->times.push_back((double)k * between);
            }
            scene.addMeshObstacle(ob, data);
            continue;
        } else if (type == "Axis Aligned Box") {
            // src/scene_building.cpp:204-211 (getStaticAxisAlignedBoxFromObject)
            ob.type = PDFC_OBSTACLE_BOX;
            auto p = vector3(require_array(o, "Position"));
            ob.material = (int32_t)require_uint(o, "Material Identifier");
            auto sz = vector3(require_array(o, "Size"));
            for (int k = 0; k < 3; ++k) {
                ob.point[k] = p[k];
                ob.size[k] = sz[k];
            }
        } else if (type == "Rotating Sphere") {
            // src/scene_building.cpp:279-293 (getRotatingSphereFromObject)
            ob.type = PDFC_OBSTACLE_ROTATING_SPHERE;
            ob.radius = require_double(o, "Radius");
            auto c = vector3(require_array(o, "Position"));
            auto ax = vector3(require_array(o, "Rotation Axe"));
            for (int k = 0; k < 3; ++k) {
                ob.point[k] = c[k];
                ob.rotation_axis[k] = ax[k];
            }
            ob.angular_speed = require_double(o, "Radial Speed");
            ob.rotation_start = require_double(o, "Rotation Start Time");
            ob.material = (int32_t)require_uint(o, "Material Identifier");
        } else {
            throw std::invalid_argument("JSON reader - Error : Invalid type for static object type");
        }
        scene.addObstacle(ob);
    }
    scene.finalize();
    return scene;
}

// src/simulation_building.cpp:37-78
SimulationParameters getSimulationParametersFromJSONDocument(const json::Value& doc) {
    SimulationParameters p;
    const json::Value& pv = require(doc, "Parameters");
    p.max_output_frame = require_uint(pv, "Frame number");
    p.local_global_iterations = require_uint(pv, "LocalGlobal");
    p.time_step = require_double(pv, "Timestep");
    p.sim_steps_per_frame = require_uint(pv, "Steps per frame");
    p.self_collision = require_bool(pv, "Self-collision");
    p.self_collision_tolerance = p.self_collision ? require_double(pv, "Self-collision tolerance") : 0.;
    const json::Value& g = require_array(pv, "Gravity");
    for (std::size_t c = 0; c < 3; ++c) p.gravity[c] = g[c].GetDouble();
    try {
        p.air_damping = require_double(pv, "Air damping");
    } catch (const std::invalid_argument&) {
        p.air_damping = 0.;
    }
    if (const json::Value* t = pv.FindMember("Type"))
        if (t->IsString()) p.simulation_type = t->GetString();
    return p;
}

// src/log_building.cpp:29-118
StepCallBacks getLogCallBacksFromJSONDocument(const json::Value& doc) {
    if (!doc.IsObject()) throw std::invalid_argument("JSON reader - Error : root must be an object");
    StepCallBacks cbs;
    if (!doc.HasMember("log")) return cbs;
    for (const json::Value& lo : require_array(doc, "log").array()) {
        std::string type = require_string(lo, "type");
        std::vector<std::size_t> idx;
        for (const json::Value& x : require_array(lo, "vertices").array()) idx.push_back((std::size_t)x.GetInt());
        if (type == "position") {
            cbs.addCallback([idx](std::size_t, const PhysicScene& scene) {
                std::cout << "Position " << idx.size() << "\n";
                for (std::size_t i : idx) {
                    auto p = scene.getMesh(0)->getVertexPosition(i);
                    std::cout << "v" << i << "\n" << p[0] << "\n" << p[1] << "\n" << p[2] << "\n";
                }
            });
        } else if (type == "velocity") {
            cbs.addCallback([idx](std::size_t, const PhysicScene& scene) {
                std::cout << "Velocity " << idx.size() << "\n";
                for (std::size_t i : idx) {
                    auto p = scene.getMesh(0)->getVertexSpeed(i);
                    std::cout << "v" << i << "\n" << p[0] << "\n" << p[1] << "\n" << p[2] << "\n";
                }
            });
        } else {
            throw std::invalid_argument("Log type \"" + type + "\" is unsupported");
        }
    }
    return cbs;
}

// template/physics/io.tpp:23-63
void computeAndWriteStats(const std::string& name, std::ofstream& of, const std::vector<double>& vec) {
    of << name << std::endl;
    if (vec.empty()) {
        of << "\t(No data)" << std::endl;
    } else if (vec.size() == 1) {
        of << "\t Value : " << vec[0] << std::endl;
    } else {
        double mn = std::numeric_limits<double>::max(), mx = std::numeric_limits<double>::min(), sum = 0.;
        for (double x : vec) {
            mn = std::min(mn, x);
            mx = std::max(mx, x);
            sum += x;
        }
        of << "\t Number      : " << vec.size() << std::endl
           << "\t Min value   : " << mn << std::endl
           << "\t Mean value  : " << sum / vec.size() << std::endl
           << "\t Max value   : " << mx << std::endl;
    }
    of << std::endl << "-------------------" << std::endl << std::endl;
}
