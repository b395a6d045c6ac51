// physecs::Scene over libphysecs_gpu.so (see physecs_gpu_scene.h).
#include "physecs_gpu_scene.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace physecs {

namespace {

void check(int rc, const char* what) {
    if (rc != PGX_OK) throw std::runtime_error(std::string(what) + ": " + pgx_last_error());
}

void put3(std::vector<float>& v, const glm::vec3& x) { v.insert(v.end(), {x.x, x.y, x.z}); }
void put4(std::vector<float>& v, const glm::quat& q) { v.insert(v.end(), {q.x, q.y, q.z, q.w}); }

}  // namespace

// src/Physecs.cpp:20-23
ContactType defaultContactFilter(bool isTrigger0, int /*data0*/, bool isTrigger1, int /*data1*/) {
    return isTrigger0 || isTrigger1 ? TRIGGER : COLLISION;
}

Scene::Scene(entt::registry& r, int /*numThreads*/, int dev) : registry(r), device(dev) { connectSignals(); }

Scene::~Scene() {
    // the hooks the constructor connected (src/Physecs.cpp:92-98 connects the reference's)
    registry.on_construct<RigidBodyCollisionComponent>().disconnect(this);
    registry.on_destroy<RigidBodyCollisionComponent>().disconnect(this);
    registry.on_construct<RigidBodyDynamicComponent>().disconnect(this);
    registry.on_destroy<RigidBodyDynamicComponent>().disconnect(this);
    registry.on_update<TransformComponent>().disconnect(this);
    if (world) pgx_world_destroy(world);
}

void Scene::connectSignals() {
    // bodies and colliders created or destroyed: the device world is rebuilt at the next
    // simulate; a patched transform moves its actor's bounds at once (onRigidBodyMove)
    registry.on_construct<RigidBodyCollisionComponent>().connect<&Scene::markDirty>(*this);
    registry.on_destroy<RigidBodyCollisionComponent>().connect<&Scene::markDirty>(*this);
    registry.on_construct<RigidBodyDynamicComponent>().connect<&Scene::markDirty>(*this);
    registry.on_destroy<RigidBodyDynamicComponent>().connect<&Scene::markDirty>(*this);
    registry.on_update<TransformComponent>().connect<&Scene::onTransformPatched>(*this);
}

void Scene::setNumSubSteps(int n) {
    subSteps = n;
    if (world) check(pgx_world_set_params(world, subSteps, iterations, gravity), "setNumSubSteps");
}
void Scene::setNumIterations(int n) {
    iterations = n;
    if (world) check(pgx_world_set_params(world, subSteps, iterations, gravity), "setNumIterations");
}
void Scene::setGravity(float g) {
    gravity = g;
    if (world) check(pgx_world_set_params(world, subSteps, iterations, gravity), "setGravity");
}

int Scene::actorIndex(entt::entity e) const {
    auto it = actorOf.find(e);
    return it == actorOf.end() ? -1 : it->second;
}

int Scene::meshIndex(const ConvexMesh* mesh) {
    auto it = convexOf.find(mesh);
    if (it != convexOf.end()) return it->second;
    const int m = (int)convexOf.size();
    convexOf.emplace(mesh, m);
    // ConvexMesh keeps its vertices padded to a multiple of 4 (src/ConvexMesh.cpp:3-11); the
    // library pads them the same way, so the unpadded ones go in
    const int nv = mesh->vertices.size();
    for (int i = 0; i < nv; ++i) put3(cvVerts, mesh->vertices[i]);
    cvVertStart.push_back(cvVertStart.back() + nv);
    for (const ConvexMeshFace& f : mesh->faces) {
        faceIdx.insert(faceIdx.end(), f.indices.begin(), f.indices.end());
        faceIdxStart.push_back((int32_t)faceIdx.size());
        put3(faceNormal, f.normal);
        put3(faceCentroid, f.centroid);
    }
    cvFaceStart.push_back(cvFaceStart.back() + (int32_t)mesh->faces.size());
    return m;
}

int Scene::meshIndex(const TriangleMesh* mesh) {
    auto it = trimeshOf.find(mesh);
    if (it != trimeshOf.end()) return it->second;
    const int m = (int)trimeshOf.size();
    trimeshOf.emplace(mesh, m);
    // The TriangleMesh constructor left its triangles in build order and its BVH built
    // (src/TriangleMesh.cpp:7-164); both go to the device as they are, as a cache, so the
    // triangle indices (the contact cache's keys) and the BVH are the reference's own.
    const int nv = (int)mesh->vertices.size(), nt = (int)mesh->triangles.size(), nn = (int)mesh->bvh.size();
    std::vector<float> v, tn, tc, nb;
    std::vector<int32_t> ti, ncount, nindex;
    for (const glm::vec3& p : mesh->vertices) put3(v, p);
    for (const Triangle& t : mesh->triangles) {
        ti.insert(ti.end(), {(int32_t)t.indices[0], (int32_t)t.indices[1], (int32_t)t.indices[2]});
        put3(tn, t.normal);
        put3(tc, t.centroid);
    }
    for (const TriangleMeshBVHNode& n : mesh->bvh) {
        put3(nb, n.bounds.min);
        put3(nb, n.bounds.max);
        ncount.push_back(n.triCount);
        nindex.push_back(n.index);
    }
    size_t bytes = 0;
    check(pgx_trimesh_cache_pack(v.data(), nv, ti.data(), nt, tn.data(), tc.data(), nb.data(), ncount.data(),
                                 nindex.data(), nn, nullptr, 0, &bytes), "TriangleMesh cache size");
    triCache.emplace_back(bytes);
    check(pgx_trimesh_cache_pack(v.data(), nv, ti.data(), nt, tn.data(), tc.data(), nb.data(), ncount.data(),
                                 nindex.data(), nn, triCache.back().data(), bytes, &bytes), "TriangleMesh cache");
    triCacheBytes.push_back(bytes);
    triVerts.insert(triVerts.end(), v.begin(), v.end());
    triIdx.insert(triIdx.end(), ti.begin(), ti.end());
    triVertStart.push_back(triVertStart.back() + nv);
    triStart.push_back(triStart.back() + nt);
    return m;
}

const pgx_scene_desc& Scene::gatherDescriptor() {
    actors.clear(); dynEntities.clear(); actorOf.clear(); entity.clear();
    pos.clear(); quat.clear(); vel.clear(); angvel.clear(); invMass.clear(); com.clear(); invI.clear();
    colPos.clear(); colQuat.clear(); colParams.clear(); colMat.clear();
    dynOfActor.clear(); actorOfDyn.clear(); colActor.clear(); colIndex.clear(); colType.clear(); colData.clear();
    kinematic.clear(); colTrigger.clear(); colEnable.clear();
    convexOf.clear(); trimeshOf.clear();
    cvVertStart.assign(1, 0); cvFaceStart.assign(1, 0); faceIdxStart.assign(1, 0); faceIdx.clear();
    cvVerts.clear(); faceNormal.clear(); faceCentroid.clear();
    triVertStart.assign(1, 0); triStart.assign(1, 0); triIdx.clear(); triVerts.clear();
    triCache.clear(); triCachePtr.clear(); triCacheBytes.clear();

    auto& dynStorage = registry.storage<RigidBodyDynamicComponent>();
    auto& colStorage = registry.storage<RigidBodyCollisionComponent>();
    // actors: the entities with a dynamic or a collision component, by entity value
    std::vector<entt::entity> ents;
    for (size_t i = 0; i < dynStorage.size(); ++i) ents.push_back(dynStorage.data()[i]);
    for (size_t i = 0; i < colStorage.size(); ++i)
        if (!dynStorage.contains(colStorage.data()[i])) ents.push_back(colStorage.data()[i]);
    std::sort(ents.begin(), ents.end(), [](entt::entity a, entt::entity b) { return (uint32_t)a < (uint32_t)b; });
    for (entt::entity e : ents) {
        const auto& t = registry.get<TransformComponent>(e);
        actorOf.emplace(e, (int)actors.size());
        actors.push_back(e);
        entity.push_back((uint32_t)e);
        put3(pos, t.position);
        put4(quat, t.orientation);
        dynOfActor.push_back(-1);
    }
    // the dynamic storage's packed order is the reference's body index (the component's
    // offset from the storage's first page, src/Physecs.cpp:117,333-334)
    for (size_t i = 0; i < dynStorage.size(); ++i) {
        const entt::entity e = dynStorage.data()[i];
        const RigidBodyDynamicComponent& d = dynStorage.get(e);
        const int a = actorOf.at(e);
        dynOfActor[a] = (int)dynEntities.size();
        dynEntities.push_back(e);
        actorOfDyn.push_back(a);
        kinematic.push_back(d.isKinematic ? 1 : 0);
        put3(vel, d.velocity);
        put3(angvel, d.angularVelocity);
        invMass.push_back(d.invMass);
        put3(com, d.com);
        for (int c = 0; c < 3; ++c)          // column-major, as glm::mat3 stores it
            for (int r = 0; r < 3; ++r) invI.push_back(d.invInertiaTensor[c][r]);
    }
    // colliders in broadphase-entry creation order: the collision storage's packed order,
    // each entity's colliders in list order (src/Physecs.cpp:25-37)
    for (size_t i = 0; i < colStorage.size(); ++i) {
        const entt::entity e = colStorage.data()[i];
        const RigidBodyCollisionComponent& cc = colStorage.get(e);
        const int a = actorOf.at(e);
        for (int k = 0; k < (int)cc.colliders.size(); ++k) {
            const Collider& c = cc.colliders[k];
            colActor.push_back(a);
            colIndex.push_back(k);
            put3(colPos, c.position);
            put4(colQuat, c.orientation);
            colType.push_back((int32_t)c.geometry.type);
            float p[4] = {0.f, 0.f, 0.f, 0.f};
            switch (c.geometry.type) {
                case SPHERE: p[0] = c.geometry.sphere.radius; break;
                case CAPSULE: p[0] = c.geometry.capsule.halfHeight; p[1] = c.geometry.capsule.radius; break;
                case BOX:
                    p[0] = c.geometry.box.halfExtents.x; p[1] = c.geometry.box.halfExtents.y;
                    p[2] = c.geometry.box.halfExtents.z;
                    break;
                case CONVEX_MESH:
                    p[0] = c.geometry.convex.scale.x; p[1] = c.geometry.convex.scale.y;
                    p[2] = c.geometry.convex.scale.z; p[3] = (float)meshIndex(c.geometry.convex.mesh);
                    break;
                case TRIANGLE_MESH: p[3] = (float)meshIndex(c.geometry.triangleMesh.mesh); break;
            }
            colParams.insert(colParams.end(), p, p + 4);
            colMat.insert(colMat.end(), {c.material.friction, c.material.restitution, c.material.damping});
            colTrigger.push_back(c.isTrigger ? 1 : 0);
            colEnable.push_back(c.enableSimulation ? 1 : 0);
            colData.push_back(c.data);
        }
    }
    for (const auto& b : triCache) triCachePtr.push_back(b.data());

    desc = pgx_scene_desc{};
    desc.n_actors = (int)actors.size();
    desc.entity = entity.data(); desc.pos = pos.data(); desc.quat = quat.data(); desc.dyn_of_actor = dynOfActor.data();
    desc.n_dyn = (int)dynEntities.size();
    desc.actor_of_dyn = actorOfDyn.data(); desc.kinematic = kinematic.data(); desc.vel = vel.data();
    desc.angvel = angvel.data(); desc.inv_mass = invMass.data(); desc.com = com.data(); desc.inv_inertia = invI.data();
    desc.n_colliders = (int)colActor.size();
    desc.col_actor = colActor.data(); desc.col_index = colIndex.data(); desc.col_pos = colPos.data();
    desc.col_quat = colQuat.data(); desc.col_type = colType.data(); desc.col_params = colParams.data();
    desc.col_material = colMat.data(); desc.col_trigger = colTrigger.data(); desc.col_enable = colEnable.data();
    desc.col_data = colData.data();
    desc.n_worlds = 1;
    desc.actor_world = nullptr;
    desc.n_convex = (int)convexOf.size();
    desc.convex_vert_start = cvVertStart.data(); desc.convex_verts = cvVerts.data();
    desc.convex_face_start = cvFaceStart.data(); desc.face_idx_start = faceIdxStart.data();
    desc.face_idx = faceIdx.data(); desc.face_normal = faceNormal.data(); desc.face_centroid = faceCentroid.data();
    desc.n_trimesh = (int)trimeshOf.size();
    desc.tri_vert_start = triVertStart.data(); desc.tri_verts = triVerts.data();
    desc.tri_start = triStart.data(); desc.tri_idx = triIdx.data();
    desc.tri_bvh_cache = triCachePtr.empty() ? nullptr : triCachePtr.data();
    desc.tri_bvh_cache_bytes = triCacheBytes.empty() ? nullptr : triCacheBytes.data();
    return desc;
}

void Scene::rebuild() {
    // The old world's state goes to the new world's colliders by entity and collider index,
    // the keys of the reference's contactCache and triggerCache, which survive addCollider,
    // clearColliders and the creation or destruction of other bodies (src/Physecs.cpp:725-752).
    std::vector<char> carried;
    if (world) {
        size_t n = 0;
        check(pgx_world_state_size(world, &n), "state size");
        carried.resize(n);
        check(pgx_world_save_state(world, carried.data(), n), "save state");
        pgx_world_destroy(world);
        world = nullptr;
    }
    gatherDescriptor();
    pgx_config cfg{subSteps, iterations, gravity, device, 0, 0};
    check(pgx_world_create(&cfg, &desc, &world), "pgx_world_create");
    check(pgx_world_set_params(world, subSteps, iterations, gravity), "pgx_world_set_params");
    for (const auto& p : nonColliding)
        check(pgx_world_set_can_collide(world, (uint32_t)p.first, (uint32_t)p.second, 0), "setCanCollide");
    // the joint creations and destructions replayed in their order, so every joint takes
    // the colour it took when it was created
    for (int op : jointLog) {
        if (op < 0) {
            check(pgx_world_destroy_joint(world, -1 - op), "destroyJoint");
            continue;
        }
        pgx_joint_desc d = joints[op].desc;
        d.actor0 = actorIndex(joints[op].e0);
        d.actor1 = actorIndex(joints[op].e1);
        check(pgx_world_add_joints(world, 1, &d), "createJoint");
    }
    check(pgx_world_set_joint_callbacks(world, &Scene::userJointLayout, &Scene::userJointRows, this), "joint callbacks");
    applyContactFilter();
    if (!carried.empty()) {
        int32_t nGone = 0;
        std::vector<uint32_t> gone(4 * 64);
        int rc = pgx_world_load_state_matched(world, carried.data(), carried.size(), 64, gone.data(), &nGone);
        if (rc == PGX_ERR_CAPACITY && nGone > 64) {
            gone.resize(4 * (size_t)nGone);
            rc = pgx_world_load_state_matched(world, carried.data(), carried.size(), nGone, gone.data(), &nGone);
        }
        check(rc, "carry the world's state");
        pendingExits.insert(pendingExits.end(), gone.begin(), gone.begin() + 4 * (size_t)nGone);
    }
    dirty = false;
}

void Scene::applyContactFilter() {
    if (!world) return;
    if (contactFilter == defaultContactFilter) {
        check(pgx_world_set_contact_filter(world, 0, nullptr, nullptr), "setContactFilter");
        return;
    }
    // the filter is a pure function of (isTrigger, data) on both sides: tabulate it over the
    // scene's distinct values
    std::vector<std::pair<bool, int>> distinct;
    std::vector<int32_t> cls(colActor.size());
    for (size_t c = 0; c < colActor.size(); ++c) {
        const std::pair<bool, int> v{colTrigger[c] != 0, colData[c]};
        auto it = std::find(distinct.begin(), distinct.end(), v);
        cls[c] = (int32_t)(it - distinct.begin());
        if (it == distinct.end()) distinct.push_back(v);
    }
    const int k = (int)distinct.size();
    std::vector<uint8_t> table((size_t)k * k);
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j)
            table[(size_t)i * k + j] =
                contactFilter(distinct[i].first, distinct[i].second, distinct[j].first, distinct[j].second) == TRIGGER;
    check(pgx_world_set_contact_filter(world, k, table.data(), cls.data()), "setContactFilter");
}

void Scene::pushState() {
    for (size_t a = 0; a < actors.size(); ++a) {
        const auto& t = registry.get<TransformComponent>(actors[a]);
        std::memcpy(&pos[3 * a], &t.position, 3 * sizeof(float));
        quat[4 * a] = t.orientation.x; quat[4 * a + 1] = t.orientation.y;
        quat[4 * a + 2] = t.orientation.z; quat[4 * a + 3] = t.orientation.w;
    }
    for (size_t i = 0; i < dynEntities.size(); ++i) {
        const auto& d = registry.get<RigidBodyDynamicComponent>(dynEntities[i]);
        std::memcpy(&vel[3 * i], &d.velocity, 3 * sizeof(float));
        std::memcpy(&angvel[3 * i], &d.angularVelocity, 3 * sizeof(float));
    }
    check(pgx_world_write_state(world, pos.data(), quat.data(), vel.data(), angvel.data()), "write state");
}

void Scene::pullState() {
    check(pgx_world_read_state(world, pos.data(), quat.data(), vel.data(), angvel.data()), "read state");
    for (size_t i = 0; i < dynEntities.size(); ++i) {
        if (kinematic[i]) continue;
        const entt::entity e = dynEntities[i];
        const int a = actorOfDyn[i];
        auto& t = registry.get<TransformComponent>(e);
        t.position = glm::vec3(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]);
        t.orientation = glm::quat(quat[4 * a + 3], quat[4 * a], quat[4 * a + 1], quat[4 * a + 2]);
        auto& d = registry.get<RigidBodyDynamicComponent>(e);
        d.velocity = glm::vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        d.angularVelocity = glm::vec3(angvel[3 * i], angvel[3 * i + 1], angvel[3 * i + 2]);
    }
}

void Scene::fireTriggers() {
    std::vector<uint32_t> gone;
    gone.swap(pendingExits);
    if (enterListeners.empty() && exitListeners.empty()) return;
    int cap = 256;
    std::vector<int32_t> en, ex;
    int32_t ne = 0, nx = 0;
    while (true) {
        en.resize(2 * (size_t)cap);
        ex.resize(2 * (size_t)cap);
        const int rc = pgx_world_get_trigger_events(world, cap, en.data(), &ne, ex.data(), &nx);
        if (rc == PGX_ERR_CAPACITY) { cap = std::max(ne, nx); continue; }
        check(rc, "trigger events");
        break;
    }
    // src/Physecs.cpp:538-552: {entity, colIndex} of both sides
    for (int i = 0; i < ne; ++i)
        for (OnTriggerEnterListener* l : enterListeners)
            l->onTriggerEnter(actors[colActor[en[2 * i]]], colIndex[en[2 * i]], actors[colActor[en[2 * i + 1]]],
                              colIndex[en[2 * i + 1]]);
    for (int i = 0; i < nx; ++i)
        for (OnTriggerExitListener* l : exitListeners)
            l->onTriggerExit(actors[colActor[ex[2 * i]]], colIndex[ex[2 * i]], actors[colActor[ex[2 * i + 1]]],
                             colIndex[ex[2 * i + 1]]);
    // pairs whose collider went in a rebuild since the last step
    for (size_t i = 0; i + 3 < gone.size(); i += 4)
        for (OnTriggerExitListener* l : exitListeners)
            l->onTriggerExit((entt::entity)gone[i], (int)gone[i + 1], (entt::entity)gone[i + 2], (int)gone[i + 3]);
}

void Scene::simulate(float timeStep) {
    if (dirty || !world) rebuild();
    else if (syncFromRegistry) pushState();
    check(pgx_world_step(world, timeStep), "simulate");
    pullState();
    contactPoints.clear();
    fireTriggers();
}

void Scene::onTransformPatched(entt::registry&, entt::entity e) {
    // registry.patch<TransformComponent>: onRigidBodyMove -> updateBounds (src/Physecs.cpp:51-54,79-90)
    if (!world || dirty) return;
    const int a = actorIndex(e);
    if (a < 0) return;
    const auto& t = registry.get<TransformComponent>(e);
    const float p[3] = {t.position.x, t.position.y, t.position.z};
    const float q[4] = {t.orientation.x, t.orientation.y, t.orientation.z, t.orientation.w};
    check(pgx_world_patch_transforms(world, 1, &a, p, q), "patch transform");
}

int Scene::createJoint(int type, entt::entity entity0, glm::vec3 anchor0Pos, glm::quat anchor0Or, entt::entity entity1,
                       glm::vec3 anchor1Pos, glm::quat anchor1Or, const pgx_joint_desc* params) {
    JointRecord r{};
    if (params) r.desc = *params;
    else {   // the data members' defaults (PrismaticJoint.h:8-17, ServoJoint.h:7-11, GearJoint.h:7-14)
        r.desc.upper_limit = 1.f;
        r.desc.drive_stiffness = type == PGX_JOINT_SERVO ? 30.f : 5.f;
        r.desc.drive_damping = 1.f;
        r.desc.gear_ratio = 1.f;
    }
    r.desc.type = type;
    const float a0[7] = {anchor0Pos.x, anchor0Pos.y, anchor0Pos.z, anchor0Or.x, anchor0Or.y, anchor0Or.z, anchor0Or.w};
    const float a1[7] = {anchor1Pos.x, anchor1Pos.y, anchor1Pos.z, anchor1Or.x, anchor1Or.y, anchor1Or.z, anchor1Or.w};
    std::memcpy(r.desc.anchor0, a0, sizeof a0);
    std::memcpy(r.desc.anchor1, a1, sizeof a1);
    r.e0 = entity0;
    r.e1 = entity1;
    r.alive = true;
    joints.push_back(r);
    jointLog.push_back((int)joints.size() - 1);
    if (world && !dirty) {
        pgx_joint_desc d = r.desc;
        d.actor0 = actorIndex(entity0);
        d.actor1 = actorIndex(entity1);
        check(pgx_world_add_joints(world, 1, &d), "createJoint");
    }
    return (int)joints.size() - 1;
}

void Scene::setJointParams(int joint, const pgx_joint_desc& params) {
    JointRecord& r = joints.at(joint);
    const pgx_joint_desc keep = r.desc;
    r.desc = params;
    r.desc.type = keep.type;
    std::memcpy(r.desc.anchor0, keep.anchor0, sizeof keep.anchor0);
    std::memcpy(r.desc.anchor1, keep.anchor1, sizeof keep.anchor1);
    if (world && !dirty) {
        pgx_joint_desc d = r.desc;
        d.actor0 = actorIndex(r.e0);
        d.actor1 = actorIndex(r.e1);
        check(pgx_world_set_joint_params(world, joint, &d), "joint setter");
    }
}

void Scene::destroyJoint(int joint) {
    if (!joints.at(joint).alive) return;
    joints[joint].alive = false;
    jointLog.push_back(-1 - joint);
    if (world && !dirty) check(pgx_world_destroy_joint(world, joint), "destroyJoint");
}

void Scene::addUserJoint(Joint* joint) {
    JointRecord r{};
    r.desc.type = PGX_JOINT_CUSTOM;
    const glm::vec3 p0 = joint->getAnchor0Pos(), p1 = joint->getAnchor1Pos();
    const glm::quat q0 = joint->getAnchor0Or(), q1 = joint->getAnchor1Or();
    const float a0[7] = {p0.x, p0.y, p0.z, q0.x, q0.y, q0.z, q0.w};
    const float a1[7] = {p1.x, p1.y, p1.z, q1.x, q1.y, q1.z, q1.w};
    std::memcpy(r.desc.anchor0, a0, sizeof a0);
    std::memcpy(r.desc.anchor1, a1, sizeof a1);
    r.e0 = joint->getEntity0();
    r.e1 = joint->getEntity1();
    r.alive = true;
    r.user = joint;
    joints.push_back(r);
    jointLog.push_back((int)joints.size() - 1);
    if (world && !dirty) {
        pgx_joint_desc d = r.desc;
        d.actor0 = actorIndex(r.e0);
        d.actor1 = actorIndex(r.e1);
        check(pgx_world_add_joints(world, 1, &d), "createJoint");
    }
}

void Scene::destroyJoint(Joint* joint) {
    for (size_t j = 0; j < joints.size(); ++j)
        if (joints[j].alive && joints[j].user == joint) {
            destroyJoint((int)j);
            joints[j].user = nullptr;
            delete joint;
            return;
        }
}

// Joint::getSolverDesc of a user joint (src/Physecs.cpp:338-339), on the registry as the
// last step left it; the step fails when the joint declares more rows than the device holds
int Scene::userJointLayout(void* scene, int joint, const pgx_joint_world_space*, int32_t* flags) {
    Scene& s = *static_cast<Scene*>(scene);
    JointRecord& r = s.joints.at(joint);
    std::vector<int32_t> declared;
    Constraint1DLayout layout(declared);
    r.solver = r.user->getSolverDesc(s.registry, layout);
    if (declared.size() <= (size_t)PGX_JOINT_MAX_ROWS) std::copy(declared.begin(), declared.end(), flags);
    return (int)declared.size();
}

// the MakeConstraintsFunc that getSolverDesc returned, on JointWorldSpaceData (the same layout
// as pgx_joint_world_space), src/Physecs.cpp:431-437
void Scene::userJointRows(void* scene, int joint, const pgx_joint_world_space* ws, pgx_joint_row* rows, int n) {
    static_assert(sizeof(JointWorldSpaceData) == sizeof(pgx_joint_world_space), "JointWorldSpaceData layout");
    Scene& s = *static_cast<Scene*>(scene);
    JointRecord& r = s.joints.at(joint);
    JointWorldSpaceData data;
    std::memcpy(&data, ws, sizeof data);
    Constraint1DWriter writer(rows, n);
    if (r.solver.makeConstraintsFunc) r.solver.makeConstraintsFunc(data, r.solver.additionalData, writer);
}

void Scene::clearColliders(entt::entity e) {
    registry.get<RigidBodyCollisionComponent>(e).colliders.clear();
    dirty = true;
}

void Scene::addCollider(entt::entity e, const Collider& collider) {
    registry.get<RigidBodyCollisionComponent>(e).colliders.push_back(collider);
    dirty = true;
}

void Scene::setIsKinematic(entt::entity e, bool isKinematic) {
    auto& d = registry.get<RigidBodyDynamicComponent>(e);
    d.isKinematic = isKinematic;
    if (isKinematic) {              // src/Physecs.cpp:757-760
        d.velocity = glm::vec3(0);
        d.angularVelocity = glm::vec3(0);
    }
    if (!world || dirty) return;
    // the flag changes the world's broadphase classes: rebuilt now, its contact cache, gear
    // angles and trigger lists carried across
    rebuild();
}

void Scene::setCanCollide(entt::entity e0, entt::entity e1, bool canCollide) {
    const auto key = (uint32_t)e0 < (uint32_t)e1 ? std::make_pair(e0, e1) : std::make_pair(e1, e0);
    auto it = std::find(nonColliding.begin(), nonColliding.end(), key);
    if (canCollide && it != nonColliding.end()) nonColliding.erase(it);
    if (!canCollide && it == nonColliding.end()) nonColliding.push_back(key);
    if (world && !dirty)
        check(pgx_world_set_can_collide(world, (uint32_t)e0, (uint32_t)e1, canCollide ? 1 : 0), "setCanCollide");
}

void Scene::setContactFilter(ContactType (*filter)(bool, int, bool, int)) {
    contactFilter = filter ? filter : defaultContactFilter;
    if (world && !dirty) applyContactFilter();
}

void Scene::addOnTriggerEnterCallback(OnTriggerEnterListener* c) { enterListeners.push_back(c); }
void Scene::addOnTriggerExitCallback(OnTriggerExitListener* c) { exitListeners.push_back(c); }
void Scene::removeOnTriggerEnterCallback(OnTriggerEnterListener* c) {
    enterListeners.erase(std::find(enterListeners.begin(), enterListeners.end(), c));
}
void Scene::removeOnTriggerExitCallback(OnTriggerExitListener* c) {
    exitListeners.erase(std::find(exitListeners.begin(), exitListeners.end(), c));
}

entt::entity Scene::raycastClosest(glm::vec3 rayOrig, glm::vec3 rayDir, float maxDistance, glm::vec3* hitPos) {
    // the reference's unfiltered overload calls an empty filter at every leaf, which never
    // accepts (src/Physecs.cpp:581,611-617)
    if (hitPos) *hitPos = rayOrig;
    (void)rayDir;
    (void)maxDistance;
    return entt::null;
}

entt::entity Scene::raycastClosest(glm::vec3 rayOrig, glm::vec3 rayDir, float maxDistance, entt::entity reject,
                                   glm::vec3* hitPos) {
    if (dirty || !world) rebuild();
    int32_t col = -1;
    float t = 0.f;
    const uint32_t ig = (uint32_t)reject;
    check(pgx_world_raycast(world, 1, &rayOrig.x, &rayDir.x, &maxDistance, nullptr, &ig, &col, &t, nullptr),
          "raycastClosest");
    if (col < 0) return entt::null;
    if (hitPos) *hitPos = rayOrig + rayDir * t;
    return actors[colActor[col]];
}

void Scene::queryShape(const Geometry& g, int32_t& type, float* p) {
    type = (int32_t)g.type;
    p[0] = p[1] = p[2] = p[3] = 0.f;
    switch (g.type) {
        case SPHERE: p[0] = g.sphere.radius; break;
        case CAPSULE: p[0] = g.capsule.halfHeight; p[1] = g.capsule.radius; break;
        case BOX: p[0] = g.box.halfExtents.x; p[1] = g.box.halfExtents.y; p[2] = g.box.halfExtents.z; break;
        case CONVEX_MESH: {
            auto it = convexOf.find(g.convex.mesh);
            if (it == convexOf.end()) throw std::runtime_error("query: convex mesh not in the scene");
            p[0] = g.convex.scale.x; p[1] = g.convex.scale.y; p[2] = g.convex.scale.z; p[3] = (float)it->second;
            break;
        }
        case TRIANGLE_MESH: {
            auto it = trimeshOf.find(g.triangleMesh.mesh);
            if (it == trimeshOf.end()) throw std::runtime_error("query: triangle mesh not in the scene");
            p[3] = (float)it->second;
            break;
        }
    }
}

std::vector<OverlapHit> Scene::overlap(glm::vec3 pos, glm::quat ori, Geometry geometry, int filter) {
    if (dirty || !world) rebuild();
    int32_t type, off[2], total = 0;
    float prm[4];
    queryShape(geometry, type, prm);
    const float q[4] = {ori.x, ori.y, ori.z, ori.w};
    std::vector<int32_t> cols(64);
    int rc;
    while ((rc = pgx_world_overlap(world, 1, &type, prm, &pos.x, q, nullptr, &filter, (int)cols.size(), off,
                                   cols.data(), &total, nullptr)) == PGX_ERR_CAPACITY)
        cols.resize(total);
    check(rc, "overlap");
    std::vector<OverlapHit> out;
    for (int k = 0; k < total; ++k) out.push_back({actors[colActor[cols[k]]], colIndex[cols[k]]});
    return out;
}

std::vector<OverlapMtdHit> Scene::overlapWithMinTranslationalDistance(glm::vec3 pos, glm::quat ori, Geometry geometry) {
    if (dirty || !world) rebuild();
    int32_t type, off[2], total = 0;
    float prm[4];
    queryShape(geometry, type, prm);
    const float q[4] = {ori.x, ori.y, ori.z, ori.w};
    std::vector<int32_t> cols(64);
    std::vector<float> nm(4 * 64);
    int rc;
    while ((rc = pgx_world_overlap_mtd(world, 1, &type, prm, &pos.x, q, nullptr, (int)cols.size(), off, cols.data(),
                                       nm.data(), &total, nullptr)) == PGX_ERR_CAPACITY) {
        cols.resize(total);
        nm.resize(4 * (size_t)total);
    }
    check(rc, "overlapWithMinTranslationalDistance");
    std::vector<OverlapMtdHit> out;
    for (int k = 0; k < total; ++k)
        out.push_back({actors[colActor[cols[k]]], colIndex[cols[k]], glm::vec3(nm[4 * k], nm[4 * k + 1], nm[4 * k + 2]),
                       nm[4 * k + 3]});
    return out;
}

const std::vector<glm::vec3>& Scene::getContactPoints() {
    // the last step's contact points (both sides of every point), as the reference's debug
    // list collects them (src/Physecs.cpp:278-279)
    if (contactPoints.empty() && world) {
        pgx_counts c{};
        check(pgx_world_get_counts(world, &c), "counts");
        int n = 0;
        std::vector<int32_t> keys(5 * (size_t)std::max(c.manifolds, 1)), ints(4 * (size_t)std::max(c.manifolds, 1));
        std::vector<float> f(56 * (size_t)std::max(c.manifolds, 1));
        check(pgx_world_get_manifolds(world, keys.data(), ints.data(), f.data(), c.manifolds, &n), "manifolds");
        for (int m = 0; m < n; ++m)
            for (int k = 0; k < ints[4 * m + 2]; ++k) {
                const float* q = &f[56 * (size_t)m + 4 + 6 * k];
                contactPoints.emplace_back(q[0], q[1], q[2]);
                contactPoints.emplace_back(q[3], q[4], q[5]);
            }
    }
    return contactPoints;
}

}  // namespace physecs
