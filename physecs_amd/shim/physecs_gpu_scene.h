// physecs::Scene over libphysecs_gpu.so: the drop-in host side of the MI355X step.
//
// This header takes the place of the reference's include/Physecs/Physecs.h (build
// physecs_gpu_scene.cpp instead of src/Physecs.cpp). It keeps the reference's public surface
// (include/Physecs/Physecs.h:14-230): construction on an app-owned entt::registry, the
// substep / iteration / gravity knobs, simulate(float), joints, pairs that do not collide,
// kinematic switches, collider edits, the contact filter, trigger listeners and the scene
// queries. The components are the reference's own (include/Physecs/Components.h,
// include/Physecs/Colliders.h) and the app's TransformComponent with src/Transform.h's
// layout; simulate() gathers them into the library's SoA descriptor when bodies or
// colliders were added or removed, steps on the GPU, and writes poses and velocities back
// into the components in place, as the reference's simulate does (src/Physecs.cpp:501-530).
//
// Joints: a built-in joint is created by its type (PGX_JOINT_*) and its data members
// (pgx_joint_desc), which is what createJoint<T> and the joints' setters carry
// (src/Joints/*.cpp), and runs entirely on the device. A user's own Joint subclass
// (include/Physecs/Joint.h) is created with createJoint<T> as in the reference; its
// getSolverDesc and MakeConstraintsFunc write rows through this folder's
// Constraint1DContainer.h and run on the host at every step and substep
// (pgx_world_set_joint_callbacks).
#pragma once

#include <cstdint>
#include <functional>
#include <unordered_map>
#include <vector>

#include <entt.hpp>
#include <glm/glm.hpp>
#include <glm/gtc/quaternion.hpp>

#include "Components.h"
#include "Constraint1DContainer.h"
#include "Joint.h"
#include "Transform.h"
#include "physecs_gpu.h"

namespace physecs {

// include/Physecs/Physecs.h:18-43
struct OverlapHit {
    entt::entity entity;
    int colIndex;
};

struct OverlapMtdHit {
    entt::entity entity;
    int colIndex;
    glm::vec3 normal;
    float mtd;
};

class OnTriggerEnterListener {
public:
    virtual void onTriggerEnter(entt::entity, int, entt::entity, int) = 0;
    virtual ~OnTriggerEnterListener() = default;
};

class OnTriggerExitListener {
public:
    virtual void onTriggerExit(entt::entity, int, entt::entity, int) = 0;
    virtual ~OnTriggerExitListener() = default;
};

enum ContactType { COLLISION, TRIGGER };
ContactType defaultContactFilter(bool isTrigger0, int data0, bool isTrigger1, int data1);

class Scene {
public:
    // Scene(registry, numThreads) (Physecs.h:199): the thread count has no role on the GPU
    Scene(entt::registry& registry, int numThreads, int device = 0);
    explicit Scene(entt::registry& registry) : Scene(registry, 0) {}
    ~Scene();
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    void setNumSubSteps(int numSubSteps);        // Physecs.h:201
    void setNumIterations(int numIterations);    // Physecs.h:202
    void setGravity(float gravity);              // Physecs.h:203
    void simulate(float timeStep);               // Physecs.h:204

    // Physecs.h:205-206. The unfiltered overload never reports a hit, as in the reference
    // (its leaf test calls the empty filter, src/Physecs.cpp:581). The filtered overload
    // takes the filter this boundary carries: one entity it rejects.
    entt::entity raycastClosest(glm::vec3 rayOrig, glm::vec3 rayDir, float maxDistance, glm::vec3* hitPos = nullptr);
    entt::entity raycastClosest(glm::vec3 rayOrig, glm::vec3 rayDir, float maxDistance, entt::entity reject,
                                glm::vec3* hitPos = nullptr);
    std::vector<OverlapHit> overlap(glm::vec3 pos, glm::quat ori, Geometry geometry, int filter);   // Physecs.h:207
    std::vector<OverlapMtdHit> overlapWithMinTranslationalDistance(glm::vec3 pos, glm::quat ori,
                                                                   Geometry geometry);              // Physecs.h:208

    // createJoint<T>(entity0, anchor0Pos, anchor0Or, entity1, anchor1Pos, anchor1Or)
    // (Physecs.h:209-214) for built-in type `type` (PGX_JOINT_*); `params` are the type's
    // data members (pgx_joint_desc: limits, drive, target, ratio), NULL for the defaults.
    // Returns the joint's handle (its creation index).
    int createJoint(int type, entt::entity entity0, glm::vec3 anchor0Pos, glm::quat anchor0Or, entt::entity entity1,
                    glm::vec3 anchor1Pos, glm::quat anchor1Or, const pgx_joint_desc* params = nullptr);
    void setJointParams(int joint, const pgx_joint_desc& params);   // the joints' setters
    void destroyJoint(int joint);                                    // Physecs.h:215

    // createJoint<T> (Physecs.h:209-214) for a user Joint subclass T: the Scene calls T's
    // getSolverDesc at every step and the MakeConstraintsFunc it returns at every substep
    // (src/Physecs.cpp:321-355,430-439) with the world-space data of the device's poses.
    template <typename T>
    T* createJoint(entt::entity entity0, glm::vec3 anchor0Pos, glm::quat anchor0Or, entt::entity entity1,
                   glm::vec3 anchor1Pos, glm::quat anchor1Or) {
        T* joint = new T(entity0, anchor0Pos, anchor0Or, entity1, anchor1Pos, anchor1Or);
        addUserJoint(joint);
        return joint;
    }
    // Scene::destroyJoint(Joint*) (Physecs.h:215, src/Physecs.cpp:712-723) for a joint made by
    // createJoint<T>: the joint leaves the world and is deleted. Joints still alive when the
    // Scene is destroyed are not deleted, as in the reference.
    void destroyJoint(Joint* joint);

    // Physecs.h:216-217: collider edits of one entity (the world is rebuilt from the
    // registry at the next simulate; the contact cache starts over)
    void clearColliders(entt::entity entity);
    void addCollider(entt::entity entity, const Collider& collider);
    void setIsKinematic(entt::entity entity, bool isKinematic);     // Physecs.h:218
    void addOnTriggerEnterCallback(OnTriggerEnterListener* callback);      // Physecs.h:219-222
    void addOnTriggerExitCallback(OnTriggerExitListener* callback);
    void removeOnTriggerEnterCallback(OnTriggerEnterListener* callback);
    void removeOnTriggerExitCallback(OnTriggerExitListener* callback);
    void setCanCollide(entt::entity entity0, entt::entity entity1, bool canCollide);   // Physecs.h:223
    void setContactFilter(ContactType (*filter)(bool, int, bool, int));               // Physecs.h:224
    entt::registry& getRegistry() { return registry; }                                 // Physecs.h:225
    const std::vector<glm::vec3>& getContactPoints();                                  // Physecs.h:226

    // By default every simulate() first copies the registry's poses and velocities to the
    // device, so edits the app made since the last step count (the reference reads the
    // components directly). An app that edits only through patch<TransformComponent> (the
    // hooked path) can turn the copy off.
    void setSyncFromRegistry(bool enabled) { syncFromRegistry = enabled; }

    // The descriptor gathered from the registry as it is now: what world creation uploads.
    const pgx_scene_desc& gatherDescriptor();
    pgx_world* deviceWorld() { return world; }

private:
    entt::registry& registry;
    int device = 0;
    int subSteps = 8, iterations = 2;            // Physecs.h:156-157
    float gravity = 9.81f;                       // Physecs.h:159
    pgx_world* world = nullptr;
    bool dirty = true;                           // bodies or colliders added or removed: rebuild
                                                 // the world, carrying its caches across
    bool syncFromRegistry = true;
    ContactType (*contactFilter)(bool, int, bool, int) = defaultContactFilter;
    std::vector<OnTriggerEnterListener*> enterListeners;
    std::vector<OnTriggerExitListener*> exitListeners;
    std::vector<glm::vec3> contactPoints;
    // trigger pairs (entity0, index0, entity1, index1) that lost a collider in a rebuild: the
    // reference reports them as exits at the next step
    std::vector<uint32_t> pendingExits;

    struct JointRecord {
        pgx_joint_desc desc;
        entt::entity e0, e1;
        bool alive;
        Joint* user;             // a user subclass (PGX_JOINT_CUSTOM), or NULL
        JointSolverDesc solver;  // what its getSolverDesc returned for this step
    };
    std::vector<JointRecord> joints;
    std::vector<int> jointLog;                   // creations (index) and destructions (-1 - index)
    std::vector<std::pair<entt::entity, entt::entity>> nonColliding;

    // descriptor storage
    pgx_scene_desc desc{};
    std::vector<entt::entity> actors, dynEntities;
    std::unordered_map<entt::entity, int> actorOf;
    std::vector<uint32_t> entity;
    std::vector<float> pos, quat, vel, angvel, invMass, com, invI, colPos, colQuat, colParams, colMat;
    std::vector<int32_t> dynOfActor, actorOfDyn, colActor, colIndex, colType, colData;
    std::vector<uint8_t> kinematic, colTrigger, colEnable;
    // convex and triangle meshes, each app-owned mesh once
    std::unordered_map<const ConvexMesh*, int> convexOf;
    std::unordered_map<const TriangleMesh*, int> trimeshOf;
    std::vector<int32_t> cvVertStart, cvFaceStart, faceIdxStart, faceIdx, triVertStart, triStart, triIdx;
    std::vector<float> cvVerts, faceNormal, faceCentroid, triVerts;
    std::vector<std::vector<uint8_t>> triCache;
    std::vector<const void*> triCachePtr;
    std::vector<size_t> triCacheBytes;

    void addUserJoint(Joint* joint);
    static int userJointLayout(void* scene, int joint, const pgx_joint_world_space* ws, int32_t* flags);
    static void userJointRows(void* scene, int joint, const pgx_joint_world_space* ws, pgx_joint_row* rows, int n);
    void connectSignals();
    void markDirty(entt::registry&, entt::entity) { dirty = true; }
    void onTransformPatched(entt::registry&, entt::entity e);
    int meshIndex(const ConvexMesh* mesh);
    int meshIndex(const TriangleMesh* mesh);
    void rebuild();
    void applyContactFilter();
    void pushState();
    void pullState();
    void fireTriggers();
    int actorIndex(entt::entity e) const;
    void queryShape(const Geometry& geometry, int32_t& type, float* params);
};

}  // namespace physecs
