// fcpw_amd.hpp -- C++ host API of the MI355X backend: the subset of FCPW's
// Scene<DIM> (include/fcpw/fcpw.h:13-255) that builds what the GPU consumes,
// and GPUScene<DIM> (include/fcpw/fcpw_gpu.h:9-93) implemented over the C ABI
// of libfcpw_hip.so (include/fcpw_hip.h).
//
// Differences from the reference, all at the API edge:
//  - Eigen matrices are replaced by row-major float/int arrays with a row count;
//    the std::vector overloads keep the reference record types.
//  - Errors throw std::runtime_error / std::invalid_argument instead of
//    printing and calling exit(EXIT_FAILURE).
//  - fcpwDirectoryPath is kept for signature compatibility and ignored: no
//    shader source is compiled at run time.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/fcpw_hip.h"
#include "bvh_build.hpp"
#include "instances.hpp"
#include "snch_build.hpp"

namespace fcpw_amd {

// scene_data.h:17-24
enum class AggregateType {
    Baseline = 0,
    Bvh_LongestAxisCenter = 1,
    Bvh_OverlapSurfaceArea = 2,
    Bvh_SurfaceArea = 3,
    Bvh_OverlapVolume = 4,
    Bvh_Volume = 5
};
// fcpw.h:7-10
enum class PrimitiveType { LineSegment, Triangle };

// GPU records with the reference layouts (bvh_interop_structures.h:14-28,730-774)
struct float2 {
    float x, y;
} __attribute__((aligned(8)));
struct float3 {
    float x, y, z;
};
struct GPURay {
    GPURay() : o{0, 0, 0}, d{0, 0, 0}, dInv{0, 0, 0}, tMax(std::numeric_limits<float>::max()) {}
    GPURay(const float3 &o_, const float3 &d_, float tMax_ = std::numeric_limits<float>::max())
        : o(o_), d(d_), dInv{1.0f / d_.x, 1.0f / d_.y, 1.0f / d_.z}, tMax(tMax_) {}
    float3 o, d, dInv;
    float tMax;
};
struct GPUBoundingSphere {
    GPUBoundingSphere() : c{0, 0, 0}, r2(0.0f) {}
    GPUBoundingSphere(const float3 &c_, float r2_) : c(c_), r2(r2_) {}
    float3 c;
    float r2;
};
struct GPUInteraction {
    GPUInteraction() : p{0, 0, 0}, n{0, 0, 0}, uv{0, 0}, d(std::numeric_limits<float>::max()), index(FCPW_HIP_UINT_MAX) {}
    float3 p, n;
    float2 uv;
    float d;
    uint32_t index;
};
static_assert(sizeof(GPURay) == sizeof(fcpw_hip_ray), "GPURay layout");
static_assert(sizeof(GPUBoundingSphere) == sizeof(fcpw_hip_bounding_sphere), "GPUBoundingSphere layout");
static_assert(sizeof(GPUInteraction) == sizeof(fcpw_hip_interaction) && sizeof(GPUInteraction) == 40,
              "GPUInteraction layout");
// GPUBoundingCone / GPUVertex / GPUEdge (bvh_interop_structures.h:41-52,121-162)
static_assert(sizeof(fcpw_hip_bounding_cone) == 20 && sizeof(fcpw_hip_vertex) == 44 && sizeof(fcpw_hip_edge) == 56,
              "silhouette record layouts");

inline CostHeuristic heuristic_of(AggregateType t)
{
    switch (t) {
    case AggregateType::Bvh_LongestAxisCenter: return CostHeuristic::LongestAxisCenter;
    case AggregateType::Bvh_OverlapSurfaceArea: return CostHeuristic::OverlapSurfaceArea;
    case AggregateType::Bvh_SurfaceArea: return CostHeuristic::SurfaceArea;
    case AggregateType::Bvh_OverlapVolume: return CostHeuristic::OverlapVolume;
    case AggregateType::Bvh_Volume: return CostHeuristic::Volume;
    default: throw std::invalid_argument("AggregateType::Baseline has no BVH for the GPU path");
    }
}

// static partition over hardware threads, as the reference's batched helpers
// (fcpw.inl:1547-1559)
template <typename F>
inline void parallel_rows(int64_t n, F &&f)
{
    int nt = (int)std::thread::hardware_concurrency();
    if (nt < 1 || n < 65536) { f((int64_t)0, n); return; }
    std::vector<std::thread> th;
    int64_t per = n / nt;
    for (int i = 0; i < nt; i++) {
        int64_t s = i * per, e = (i == nt - 1) ? n : (i + 1) * per;
        th.emplace_back([&f, s, e]() { f(s, e); });
    }
    for (auto &t : th) t.join();
}

// ---------------------------------------------------------------------------
// Scene<DIM>: geometry ingest, BVH build, vertex updates, CPU refit
// ---------------------------------------------------------------------------
template <int DIM>
class Scene {
public:
    Scene() = default;

    // fcpw.inl:206-216
    void setObjectCount(int nObjects)
    {
        if (nObjects < 0) throw std::invalid_argument("setObjectCount: negative count");
        objects_.assign((size_t)nObjects, Object());
        bvh_.reset();
        snch_.reset();
        object_bvh_.clear();
        instances_.clear();
        instance_tree_.reset();
        silhouettes_ = false;
    }
    void setObjectVertexCount(int nVertices, int objectIndex)
    {
        obj(objectIndex).positions.assign((size_t)nVertices * DIM, 0.0f);
    }
    void setObjectVertex(const float *position, int vertexIndex, int objectIndex)
    {
        Object &o = obj(objectIndex);
        check_vertex(o, vertexIndex);
        for (int d = 0; d < DIM; d++) o.positions[(size_t)vertexIndex * DIM + d] = position[d];
    }
    // fcpw.inl:218-238 (row-major nVertices x DIM)
    void setObjectVertices(const float *positions, int nVertices, int objectIndex)
    {
        Object &o = obj(objectIndex);
        o.positions.assign(positions, positions + (size_t)nVertices * DIM);
    }
    void setObjectPrimitiveCount(int nPrims, int objectIndex)
    {
        Object &o = obj(objectIndex);
        if (o.has_prims) throw std::runtime_error(DIM == 3 ? "Already set triangles!" : "Already set line segments!");
        o.indices.assign((size_t)nPrims * DIM, -1);
        o.has_prims = true;
    }
    void setObjectPrimitive(const int32_t *indices, int primIndex, int objectIndex)
    {
        Object &o = obj(objectIndex);
        if (primIndex < 0 || (size_t)primIndex * DIM >= o.indices.size())
            throw std::out_of_range("primitive index out of range");
        for (int k = 0; k < DIM; k++) o.indices[(size_t)primIndex * DIM + k] = indices[k];
    }
    // fcpw.inl:240-390 (row-major nPrims x DIM); triangles for DIM 3, line segments for DIM 2
    void setObjectPrimitives(const int32_t *indices, int nPrims, int objectIndex)
    {
        setObjectPrimitiveCount(nPrims, objectIndex);
        Object &o = obj(objectIndex);
        o.indices.assign(indices, indices + (size_t)nPrims * DIM);
    }

    // fcpw.inl:392-397: appends instance transforms to an object; each is a
    // (DIM+1) x (DIM+1) row-major affine matrix with uniform scaling (fcpw.h:68)
    void setObjectInstanceTransforms(const float *transforms, int nTransforms, int objectIndex)
    {
        Object &o = obj(objectIndex);
        for (int k = 0; k < nTransforms; k++) {
            float t[12];
            transform_from_matrix<DIM>(transforms + (size_t)k * (DIM + 1) * (DIM + 1), t);
            o.transforms.insert(o.transforms.end(), t, t + 12);
        }
    }

    // fcpw.inl:429-521: marks the scene for silhouette queries; the silhouette
    // data is built with the BVH by the next build() (as the reference, whose
    // build creates the SNCH variant when silhouettes were computed)
    void computeSilhouettes(const IgnoreSilhouetteFn &ignoreSilhouette = {})
    {
        if (objects_.size() != 1)
            throw std::runtime_error("silhouettes need a scene of exactly one object on the MI355X backend");
        silhouettes_ = true;
        ignore_ = ignoreSilhouette;
    }

    // Scene::build (fcpw.inl:1052-1126) with binary BVHs.  One object without
    // instance transforms gets one tree; otherwise every object gets its own and
    // the instances (one per transform, or the object itself) a tree above them.
    // vectorize is accepted and ignored: the GPU path needs the binary BVH
    // (fcpw_gpu.h:17-20), and no vectorized CPU traversal exists here.
    void build(AggregateType aggregateType, bool vectorize, bool printStats = false, bool reduceMemoryFootprint = false)
    {
        (void)vectorize;
        (void)reduceMemoryFootprint;
        if (objects_.empty()) throw std::runtime_error("Cannot construct a BVH without primitives");
        object_bvh_.clear();
        instances_.clear();
        instance_tree_.reset();
        if (objects_.size() > 1 || !objects_[0].transforms.empty()) {
            buildInstanced(aggregateType, printStats);
            return;
        }
        Object &o = objects_[0];
        if (!o.has_prims || o.indices.empty()) throw std::runtime_error("Cannot construct a BVH without primitives");
        aggregate_ = aggregateType;
        auto t0 = std::chrono::high_resolution_clock::now();
        bvh_ = std::make_unique<FlatBvh<DIM>>(build_bvh<DIM>(o.positions.data(), nVertices(), o.indices.data(),
                                                             nPrimitives(), heuristic_of(aggregateType)));
        snch_.reset();
        if (silhouettes_)
            snch_ = std::make_unique<SnchData<DIM>>(
                build_snch<DIM>(*bvh_, o.positions.data(), nVertices(), o.indices.data(), nPrimitives(), ignore_));
        if (printStats) {
            double s = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
            printf("BVH construction time: %g seconds\n", s);
            printf("BVH stats: %zu nodes, %d leaves, %d max depth, %d primitives\n", bvh_->nodes.size(),
                   bvh_->n_leaves, bvh_->max_depth, nPrimitives());
        }
        ++version_;
    }

    // fcpw.inl:1128-1143 (vertex ids in the caller's order)
    void updateObjectVertex(const float *position, int vertexIndex, int objectIndex)
    {
        setObjectVertex(position, vertexIndex, objectIndex);
    }
    void updateObjectVertices(const float *positions, int nVertices, int objectIndex)
    {
        Object &o = obj(objectIndex);
        if ((size_t)nVertices * DIM != o.positions.size())
            throw std::invalid_argument("updateObjectVertices: vertex count differs from the scene's");
        o.positions.assign(positions, positions + (size_t)nVertices * DIM);
    }
    // Scene::refit (fcpw.inl:1145-1158) on the host copy of the BVH (with
    // instances: every object's tree, then the tree over the instances)
    void refit(bool printStats = false)
    {
        auto t0 = std::chrono::high_resolution_clock::now();
        if (instance_tree_) {
            for (size_t i = 0; i < objects_.size(); i++)
                refit_bvh<DIM>(*object_bvh_[i], objects_[i].positions.data(), objects_[i].indices.data());
            refit_instance_tree<DIM>(*instance_tree_, objectBvhs(), instances_);
        } else {
            require_bvh();
            refit_bvh<DIM>(*bvh_, objects_[0].positions.data(), objects_[0].indices.data());
        }
        if (snch_) refit_snch_cones<DIM>(*snch_, *bvh_, objects_[0].positions.data());
        if (printStats) {
            double s = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
            printf("refit time: %g seconds\n", s);
        }
    }

    // a tree built elsewhere (GPUScene::buildOnGPU); drops any silhouette data
    void adoptBvh(FlatBvh<DIM> &&bvh, AggregateType aggregateType)
    {
        bvh_ = std::make_unique<FlatBvh<DIM>>(std::move(bvh));
        snch_.reset();
        if (silhouettes_) {
            const Object &o = objects_.at(0);
            snch_ = std::make_unique<SnchData<DIM>>(
                build_snch<DIM>(*bvh_, o.positions.data(), nVertices(), o.indices.data(), nPrimitives(), ignore_));
        }
        aggregate_ = aggregateType;
        ++version_;
    }

    // trees built elsewhere, one per object (GPUScene::buildOnGPU on a scene of
    // several objects): the instance list and tree are then built on the host
    void adoptInstanced(std::vector<FlatBvh<DIM>> &&objectTrees, AggregateType aggregateType)
    {
        if (silhouettes_)
            throw std::runtime_error("silhouettes need a scene of exactly one object on the MI355X backend");
        if (objectTrees.size() != objects_.size()) throw std::invalid_argument("adoptInstanced: one tree per object");
        object_bvh_.clear();
        for (auto &t : objectTrees) object_bvh_.push_back(std::make_unique<FlatBvh<DIM>>(std::move(t)));
        finishInstanced(aggregateType);
    }

    // accessors used by GPUScene
    bool built() const { return (bool)bvh_ || (bool)instance_tree_; }
    // several objects or instance transforms: per-object trees and the instance tree
    bool instanced() const { return (bool)instance_tree_; }
    // build() will make an instance tree (several objects or instance transforms)
    bool needsInstances() const { return objects_.size() > 1 || (!objects_.empty() && !objects_[0].transforms.empty()); }
    int nObjects() const { return (int)objects_.size(); }
    const FlatBvh<DIM> &objectBvh(int i) const
    {
        if (!instance_tree_) return bvh();
        return *object_bvh_.at((size_t)i);
    }
    const std::vector<float> &objectPositions(int i) const { return objects_.at((size_t)i).positions; }
    const std::vector<int32_t> &objectIndices(int i) const { return objects_.at((size_t)i).indices; }
    int objectNVertices(int i) const { return (int)(objects_.at((size_t)i).positions.size() / DIM); }
    int objectNPrimitives(int i) const { return (int)(objects_.at((size_t)i).indices.size() / DIM); }
    const std::vector<fcpw_hip_instance> &instances() const { return instances_; }
    const FlatBvh<DIM> &instanceTree() const
    {
        if (!instance_tree_) throw std::runtime_error("the scene has no instance tree (one object, no transforms)");
        return *instance_tree_;
    }
    const FlatBvh<DIM> &bvh() const { require_bvh(); return *bvh_; }
    const std::vector<float> &positions() const { return objects_.at(0).positions; }
    const std::vector<int32_t> &indices() const { return objects_.at(0).indices; }
    int nVertices() const { return (int)(objects_.at(0).positions.size() / DIM); }
    int nPrimitives() const { return (int)(objects_.at(0).indices.size() / DIM); }
    uint64_t version() const { return version_; }
    bool hasSilhouettes() const { return (bool)snch_; }
    bool silhouettesRequested() const { return silhouettes_; }
    const SnchData<DIM> &snch() const
    {
        if (!snch_) throw std::runtime_error("the scene was built without silhouettes (call computeSilhouettes first)");
        return *snch_;
    }

private:
    struct Object {
        std::vector<float> positions;
        std::vector<int32_t> indices;
        std::vector<float> transforms; // 12 floats (3x4 row-major) per instance transform
        bool has_prims = false;
    };
    std::vector<const FlatBvh<DIM> *> objectBvhs() const
    {
        std::vector<const FlatBvh<DIM> *> v;
        for (const auto &b : object_bvh_) v.push_back(b.get());
        return v;
    }
    // fcpw.inl:1060-1112 for several objects or instances
    void buildInstanced(AggregateType aggregateType, bool printStats)
    {
        if (silhouettes_)
            throw std::runtime_error("silhouettes need a scene of exactly one object on the MI355X backend");
        auto t0 = std::chrono::high_resolution_clock::now();
        CostHeuristic h = heuristic_of(aggregateType);
        for (size_t i = 0; i < objects_.size(); i++) {
            const Object &o = objects_[i];
            if (!o.has_prims || o.indices.empty()) throw std::runtime_error("Cannot construct a BVH without primitives");
            object_bvh_.push_back(std::make_unique<FlatBvh<DIM>>(
                build_bvh<DIM>(o.positions.data(), (int)(o.positions.size() / DIM), o.indices.data(),
                               (int)(o.indices.size() / DIM), h)));
        }
        finishInstanced(aggregateType);
        if (printStats) {
            double s = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
            printf("BVH construction time: %g seconds (%zu objects, %zu instances)\n", s, objects_.size(),
                   instances_.size());
        }
    }
    // the instance list and the tree over the instances, once every object has its tree
    void finishInstanced(AggregateType aggregateType)
    {
        CostHeuristic h = heuristic_of(aggregateType);
        std::vector<ObjectGeometry> geo;
        for (size_t i = 0; i < objects_.size(); i++)
            geo.push_back(ObjectGeometry{objects_[i].positions.data(), objects_[i].indices.data()});
        instances_.clear();
        for (size_t i = 0; i < objects_.size(); i++) {
            const Object &o = objects_[i];
            size_t nt = o.transforms.size() / 12;
            if (nt == 0) {
                fcpw_hip_instance I{};
                I.object = (int32_t)i;
                instances_.push_back(I);
            }
            for (size_t k = 0; k < nt; k++) {
                fcpw_hip_instance I{};
                I.object = (int32_t)i;
                I.transformed = 1;
                for (int j = 0; j < 12; j++) I.t[j] = o.transforms[k * 12 + j];
                invert_transform<DIM>(I.t, I.tinv);
                instances_.push_back(I);
            }
        }
        instance_tree_ = std::make_unique<FlatBvh<DIM>>(build_instance_tree<DIM>(objectBvhs(), geo, instances_, h));
        bvh_.reset();
        snch_.reset();
        aggregate_ = aggregateType;
        ++version_;
    }
    Object &obj(int i)
    {
        if (i < 0 || (size_t)i >= objects_.size()) throw std::out_of_range("object index out of range");
        return objects_[i];
    }
    static void check_vertex(const Object &o, int v)
    {
        if (v < 0 || (size_t)v * DIM >= o.positions.size()) throw std::out_of_range("vertex index out of range");
    }
    void require_bvh() const
    {
        if (instance_tree_)
            throw std::runtime_error("the scene has several objects or instances: its trees are per object");
        if (!bvh_) throw std::runtime_error("the scene has not been built");
    }
    std::vector<Object> objects_;
    std::unique_ptr<FlatBvh<DIM>> bvh_;
    std::vector<std::unique_ptr<FlatBvh<DIM>>> object_bvh_;
    std::vector<fcpw_hip_instance> instances_;
    std::unique_ptr<FlatBvh<DIM>> instance_tree_;
    std::unique_ptr<SnchData<DIM>> snch_;
    bool silhouettes_ = false;
    IgnoreSilhouetteFn ignore_;
    AggregateType aggregate_ = AggregateType::Bvh_SurfaceArea;
    uint64_t version_ = 0;
};

// ---------------------------------------------------------------------------
// GPUScene<DIM> over the C ABI
// ---------------------------------------------------------------------------
inline void hip_check(fcpw_hip_ctx *ctx, int status, const char *what)
{
    if (status == FCPW_HIP_OK) return;
    std::string msg = std::string(what) + ": " + fcpw_hip_last_error(ctx);
    if (status == FCPW_HIP_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <int DIM>
class GPUScene {
public:
    // fcpw_gpu.inl:10-16; the directory argument is ignored
    GPUScene(const std::string &fcpwDirectoryPath, bool printLogs = false) : dir_(fcpwDirectoryPath), logs_(printLogs)
    {
        const char *dev = std::getenv("FCPW_HIP_DEVICE");
        device_ = dev ? std::atoi(dev) : 0;
    }
    ~GPUScene()
    {
        if (ctx_) fcpw_hip_destroy(ctx_);
        for (fcpw_hip_ctx *o : objs_) fcpw_hip_destroy(o);
    }
    GPUScene(const GPUScene &) = delete;
    GPUScene &operator=(const GPUScene &) = delete;

    // fcpw_gpu.inl:18-62
    void transferToGPU(Scene<DIM> &scene, const std::string &deviceBackend = "default")
    {
        (void)deviceBackend; // HIP is the only backend
        if (!ctx_) hip_check(nullptr, fcpw_hip_create(device_, &ctx_), "fcpw_hip_create");
        if (scene.instanced()) {
            transferInstanced(scene);
            return;
        }
        const FlatBvh<DIM> &b = scene.bvh();
        uploadTree(ctx_, b, scene.positions(), scene.indices(), "transferToGPU");
        instanced_ = false;
        int n = (int)b.nodes.size(), np = scene.nPrimitives();
        if (scene.hasSilhouettes()) uploadSilhouettes(scene);
        if (logs_) {
            int64_t rec = 0, depth = 0;
            fcpw_hip_get_info(ctx_, "n_records", &rec);
            fcpw_hip_get_info(ctx_, "max_depth", &depth);
            printf("transferToGPU: %d nodes -> %lld records, depth %lld, %d primitives\n", n, (long long)rec,
                   (long long)depth, np);
        }
    }

    // Scene::build + transferToGPU in one step on the device (fcpw_hip_build_bvh):
    // the reference's binned SAH tree, built from the scene's current geometry.
    // With exportTree the tree is also copied back into the scene, as if
    // scene.build(aggregateType) had made it.  A scene marked for silhouettes
    // always exports: its silhouette references and cones are then built on the
    // host from that tree and uploaded.  A scene of several objects or instance
    // transforms builds every object's tree on the device and always exports
    // them: the tree over the instances is built from them on the host.
    void buildOnGPU(Scene<DIM> &scene, AggregateType aggregateType = AggregateType::Bvh_SurfaceArea,
                    bool exportTree = false)
    {
        if (scene.needsInstances()) {
            buildInstancedOnGPU(scene, aggregateType);
            return;
        }
        if (scene.silhouettesRequested()) exportTree = true;
        if (!ctx_) hip_check(nullptr, fcpw_hip_create(device_, &ctx_), "fcpw_hip_create");
        int h = (int)heuristic_of(aggregateType);
        instanced_ = false;
        if (exportTree) {
            scene.adoptBvh(buildTree(ctx_, scene.positions(), scene.indices(), h, true), aggregateType);
            if (scene.hasSilhouettes()) uploadSilhouettes(scene);
        } else {
            buildTree(ctx_, scene.positions(), scene.indices(), h, false);
        }
    }

    // fcpw_gpu.inl:64-80: with updateGeometry, upload the scene's current vertex
    // positions first; then refit every box on the device
    // (with silhouettes, updateGeometry also re-extracts the silhouette records from
    // the new positions, as allocateGeometryBuffers does, and the refit recomputes
    // the cones from them)
    void refit(Scene<DIM> &scene, bool updateGeometry = true)
    {
        require();
        if (instanced_) {
            if ((int)objs_.size() != scene.nObjects()) throw std::runtime_error("refit: the scene's objects changed");
            if (updateGeometry)
                for (int i = 0; i < scene.nObjects(); i++)
                    hip_check(objs_[i],
                              fcpw_hip_update_positions(objs_[i], scene.objectPositions(i).data(),
                                                        scene.objectNVertices(i)),
                              "refit");
            hip_check(ctx_, fcpw_hip_refit(ctx_, nullptr), "refit");
            return;
        }
        if (updateGeometry) {
            hip_check(ctx_, fcpw_hip_update_positions(ctx_, scene.positions().data(), scene.nVertices()), "refit");
            if (scene.hasSilhouettes()) uploadSilhouettes(scene);
        }
        hip_check(ctx_, fcpw_hip_refit(ctx_, nullptr), "refit");
    }

    // fcpw_gpu.inl:94-135 (row-major n x DIM origins and directions)
    void intersect(const float *rayOrigins, const float *rayDirections, const float *rayDistanceBounds, int64_t n,
                   std::vector<GPUInteraction> &interactions, bool checkForOcclusion = false,
                   std::vector<int32_t> *objectIndices = nullptr)
    {
        std::vector<GPURay> &rays = rayBuf_;
        rays.resize((size_t)n);
        parallel_rows(n, [&](int64_t s, int64_t e) {
            for (int64_t i = s; i < e; i++) {
                const float *o = rayOrigins + i * DIM, *d = rayDirections + i * DIM;
                float3 oo{o[0], o[1], DIM == 2 ? 0.0f : o[DIM - 1]};
                float3 dd{d[0], d[1], DIM == 2 ? 0.0f : d[DIM - 1]};
                rays[i] = GPURay(oo, dd, rayDistanceBounds[i]);
            }
        });
        intersect(rays, interactions, checkForOcclusion, objectIndices);
    }
    // fcpw_gpu.inl:137-156
    void intersect(const std::vector<GPURay> &rays, std::vector<GPUInteraction> &interactions,
                   bool checkForOcclusion = false, std::vector<int32_t> *objectIndices = nullptr)
    {
        require();
        interactions.resize(rays.size()); // the device writes every record, hit or miss
        if (objectIndices) objectIndices->assign(rays.size(), -1);
        hip_check(ctx_,
                  fcpw_hip_intersect_objects(ctx_, (const fcpw_hip_ray *)rays.data(), (int64_t)rays.size(),
                                             checkForOcclusion ? 1 : 0, (fcpw_hip_interaction *)interactions.data(),
                                             objectIndices ? objectIndices->data() : nullptr),
                  "intersect");
    }
    // fcpw_gpu.inl:158-196 (row-major n x DIM centers, n x 3 random numbers in [0, 1)); with raw
    // pointers this overload would collide with the ray one, hence the distinct name
    void intersectSpheres(const float *sphereCenters, const float *sphereSquaredRadii, const float *randNums,
                          int64_t n, std::vector<GPUInteraction> &interactions,
                          std::vector<int32_t> *objectIndices = nullptr)
    {
        std::vector<GPUBoundingSphere> &spheres = sphereBuf_;
        spheres.resize((size_t)n);
        std::vector<float3> rnd((size_t)n);
        parallel_rows(n, [&](int64_t s, int64_t e) {
            for (int64_t i = s; i < e; i++) {
                const float *c = sphereCenters + i * DIM;
                spheres[i] = GPUBoundingSphere(float3{c[0], c[1], DIM == 2 ? 0.0f : c[DIM - 1]}, sphereSquaredRadii[i]);
                const float *r = randNums + i * 3;
                rnd[i] = float3{r[0], r[1], DIM == 2 ? 0.0f : r[2]};
            }
        });
        intersect(spheres, rnd, interactions, objectIndices);
    }
    // fcpw_gpu.inl:198-217
    void intersect(const std::vector<GPUBoundingSphere> &spheres, const std::vector<float3> &randNums,
                   std::vector<GPUInteraction> &interactions, std::vector<int32_t> *objectIndices = nullptr)
    {
        if (randNums.size() != spheres.size())
            throw std::invalid_argument("intersect(spheres): one rand_nums entry per sphere");
        require();
        interactions.resize(spheres.size()); // the device writes every record, hit or miss
        if (objectIndices) objectIndices->assign(spheres.size(), -1);
        hip_check(ctx_,
                  fcpw_hip_intersect_spheres_objects(ctx_, (const fcpw_hip_bounding_sphere *)spheres.data(),
                                                     (const fcpw_hip_float3 *)randNums.data(),
                                                     (int64_t)spheres.size(),
                                                     (fcpw_hip_interaction *)interactions.data(),
                                                     objectIndices ? objectIndices->data() : nullptr),
                  "intersect(spheres)");
    }

    // fcpw_gpu.inl:219-253 (row-major n x DIM points)
    void findClosestPoints(const float *queryPoints, const float *squaredMaxRadii, int64_t n,
                           std::vector<GPUInteraction> &interactions, bool recordNormals = false,
                           std::vector<int32_t> *objectIndices = nullptr)
    {
        std::vector<GPUBoundingSphere> &spheres = sphereBuf_;
        spheres.resize((size_t)n);
        parallel_rows(n, [&](int64_t s, int64_t e) {
            for (int64_t i = s; i < e; i++) {
                const float *q = queryPoints + i * DIM;
                spheres[i] = GPUBoundingSphere(float3{q[0], q[1], DIM == 2 ? 0.0f : q[DIM - 1]}, squaredMaxRadii[i]);
            }
        });
        findClosestPoints(spheres, interactions, recordNormals, objectIndices);
    }
    // fcpw_gpu.inl:255-274
    // objectIndices (optional): Interaction::objectIndex per query, -1 on a miss
    void findClosestPoints(const std::vector<GPUBoundingSphere> &spheres, std::vector<GPUInteraction> &interactions,
                           bool recordNormals = false, std::vector<int32_t> *objectIndices = nullptr)
    {
        require();
        interactions.resize(spheres.size()); // the device writes every record, hit or miss
        if (objectIndices) objectIndices->assign(spheres.size(), -1);
        hip_check(ctx_,
                  fcpw_hip_closest_points_objects(ctx_, (const fcpw_hip_bounding_sphere *)spheres.data(),
                                                  (int64_t)spheres.size(), recordNormals ? 1 : 0,
                                                  (fcpw_hip_interaction *)interactions.data(),
                                                  objectIndices ? objectIndices->data() : nullptr),
                  "findClosestPoints");
    }

    // fcpw_gpu.inl:276-316 (row-major n x DIM points)
    void findClosestSilhouettePoints(const float *queryPoints, const float *squaredMaxRadii,
                                     const int32_t *flipNormalOrientation, int64_t n,
                                     std::vector<GPUInteraction> &interactions, float squaredMinRadius = 0.0f,
                                     float precision = 1e-3f)
    {
        std::vector<GPUBoundingSphere> &spheres = sphereBuf_;
        spheres.resize((size_t)n);
        std::vector<uint32_t> flip((size_t)n);
        parallel_rows(n, [&](int64_t s, int64_t e) {
            for (int64_t i = s; i < e; i++) {
                const float *q = queryPoints + i * DIM;
                spheres[i] = GPUBoundingSphere(float3{q[0], q[1], DIM == 2 ? 0.0f : q[DIM - 1]}, squaredMaxRadii[i]);
                flip[i] = (uint32_t)flipNormalOrientation[i];
            }
        });
        findClosestSilhouettePoints(spheres, flip, interactions, squaredMinRadius, precision);
    }
    // fcpw_gpu.inl:318-338
    void findClosestSilhouettePoints(const std::vector<GPUBoundingSphere> &spheres,
                                     const std::vector<uint32_t> &flipNormalOrientation,
                                     std::vector<GPUInteraction> &interactions, float squaredMinRadius = 0.0f,
                                     float precision = 1e-3f)
    {
        if (flipNormalOrientation.size() != spheres.size())
            throw std::invalid_argument("findClosestSilhouettePoints: one flipNormalOrientation entry per sphere");
        require();
        interactions.resize(spheres.size()); // the device writes every record, hit or miss
        hip_check(ctx_,
                  fcpw_hip_closest_silhouette_points(ctx_, (const fcpw_hip_bounding_sphere *)spheres.data(),
                                                     flipNormalOrientation.data(), (int64_t)spheres.size(),
                                                     squaredMinRadius, precision,
                                                     (fcpw_hip_interaction *)interactions.data()),
                  "findClosestSilhouettePoints");
    }

    // the device context, for callers that pass device pointers directly
    fcpw_hip_ctx *context()
    {
        require();
        return ctx_;
    }

    bool instanced() const { return instanced_; }

private:
    // query records packed from caller arrays, kept between calls so that a
    // repeated call of the same size reuses memory it has already touched
    std::vector<GPUBoundingSphere> sphereBuf_;
    std::vector<GPURay> rayBuf_;

    // fcpw_hip_build_bvh into ctx; with exportTree the tree is returned as the
    // host build would have made it
    FlatBvh<DIM> buildTree(fcpw_hip_ctx *ctx, const std::vector<float> &positions, const std::vector<int32_t> &indices,
                           int h, bool exportTree)
    {
        int np = (int)(indices.size() / DIM);
        if (np <= 0) throw std::runtime_error("Cannot construct a BVH without primitives");
        std::vector<float> pmin, pmax;
        std::vector<int32_t> nrefs, offset, order;
        if (exportTree) {
            size_t cap = 2 * (size_t)np - 1;
            pmin.resize(cap * 3);
            pmax.resize(cap * 3);
            nrefs.resize(cap);
            offset.resize(cap);
            order.resize((size_t)np);
        }
        int n_nodes = 0;
        hip_check(ctx,
                  fcpw_hip_build_bvh(ctx, DIM, positions.data(), (int)(positions.size() / DIM), indices.data(), np, h,
                                     exportTree ? pmin.data() : nullptr, exportTree ? pmax.data() : nullptr,
                                     exportTree ? nrefs.data() : nullptr, exportTree ? offset.data() : nullptr,
                                     exportTree ? order.data() : nullptr, &n_nodes),
                  "buildOnGPU");
        FlatBvh<DIM> b;
        if (!exportTree) return b;
        b.nodes.resize((size_t)n_nodes);
        int leaves = 0;
        for (int i = 0; i < n_nodes; i++) {
            for (int d = 0; d < DIM; d++) {
                b.nodes[i].box.lo[d] = pmin[(size_t)i * 3 + d];
                b.nodes[i].box.hi[d] = pmax[(size_t)i * 3 + d];
            }
            b.nodes[i].nrefs = nrefs[i];
            b.nodes[i].offset = offset[i];
            leaves += nrefs[i] > 0;
        }
        int64_t md = 0;
        fcpw_hip_get_info(ctx, "max_depth", &md);
        b.prim_order = std::move(order);
        b.n_leaves = leaves;
        b.max_depth = (int)md;
        return b;
    }
    // every object's tree built on the device into its own context (and copied
    // back: the instance tree needs each root box and primitive order), then the
    // instance list and tree on the host and the instance layer uploaded
    void buildInstancedOnGPU(Scene<DIM> &scene, AggregateType aggregateType)
    {
        if (scene.silhouettesRequested())
            throw std::runtime_error("silhouettes need a scene of exactly one object on the MI355X backend");
        if (!ctx_) hip_check(nullptr, fcpw_hip_create(device_, &ctx_), "fcpw_hip_create");
        int n = scene.nObjects();
        resizeObjects(n);
        int h = (int)heuristic_of(aggregateType);
        std::vector<FlatBvh<DIM>> trees;
        for (int i = 0; i < n; i++)
            trees.push_back(buildTree(objs_[i], scene.objectPositions(i), scene.objectIndices(i), h, true));
        scene.adoptInstanced(std::move(trees), aggregateType);
        uploadInstances(scene);
    }
    void resizeObjects(int n)
    {
        while ((int)objs_.size() > n) {
            fcpw_hip_destroy(objs_.back());
            objs_.pop_back();
        }
        while ((int)objs_.size() < n) {
            fcpw_hip_ctx *o = nullptr;
            hip_check(nullptr, fcpw_hip_create(device_, &o), "fcpw_hip_create");
            objs_.push_back(o);
        }
    }
    void require()
    {
        if (!ctx_) throw std::runtime_error("the BVH must be transferred to the GPU first (call transferToGPU)");
    }
    // one object's flat tree and primitives (extractBvhNodes / extractTriangles,
    // bvh_interop_structures.h:172-275)
    static void uploadTree(fcpw_hip_ctx *ctx, const FlatBvh<DIM> &b, const std::vector<float> &positions,
                           const std::vector<int32_t> &idx, const char *what)
    {
        size_t n = b.nodes.size();
        std::vector<float> pmin(n * 3, 0.0f), pmax(n * 3, 0.0f);
        std::vector<int32_t> nrefs(n), offset(n);
        flatten(b, pmin, pmax, nrefs, offset);
        int np = (int)b.prim_order.size();
        std::vector<int32_t> pv((size_t)np * DIM);
        for (int s = 0; s < np; s++)
            for (int k = 0; k < DIM; k++) pv[(size_t)s * DIM + k] = idx[(size_t)b.prim_order[s] * DIM + k];
        hip_check(ctx,
                  fcpw_hip_upload_bvh(ctx, DIM, (int)n, pmin.data(), pmax.data(), nrefs.data(), offset.data(), np,
                                      pv.data(), b.prim_order.data(), (int)(positions.size() / DIM),
                                      positions.data()),
                  what);
    }
    static void flatten(const FlatBvh<DIM> &b, std::vector<float> &pmin, std::vector<float> &pmax,
                        std::vector<int32_t> &nrefs, std::vector<int32_t> &offset)
    {
        for (size_t i = 0; i < b.nodes.size(); i++) {
            for (int d = 0; d < DIM; d++) {
                pmin[i * 3 + d] = b.nodes[i].box.lo[d];
                pmax[i * 3 + d] = b.nodes[i].box.hi[d];
            }
            nrefs[i] = b.nodes[i].nrefs;
            offset[i] = b.nodes[i].offset;
        }
    }
    // every object's tree into a context of its own, then the instance layer
    void transferInstanced(Scene<DIM> &scene)
    {
        int n = scene.nObjects();
        resizeObjects(n);
        for (int i = 0; i < n; i++)
            uploadTree(objs_[i], scene.objectBvh(i), scene.objectPositions(i), scene.objectIndices(i),
                       "transferToGPU(object)");
        uploadInstances(scene);
    }
    // the instance layer over the object contexts
    void uploadInstances(Scene<DIM> &scene)
    {
        int n = scene.nObjects();
        const FlatBvh<DIM> &t = scene.instanceTree();
        size_t nn = t.nodes.size();
        std::vector<float> pmin(nn * 3, 0.0f), pmax(nn * 3, 0.0f);
        std::vector<int32_t> nrefs(nn), offset(nn);
        flatten(t, pmin, pmax, nrefs, offset);
        const auto &inst = scene.instances();
        hip_check(ctx_,
                  fcpw_hip_upload_instances(ctx_, objs_.data(), n, inst.data(), (int)inst.size(), (int)nn,
                                            pmin.data(), pmax.data(), nrefs.data(), offset.data(),
                                            t.prim_order.data()),
                  "transferToGPU(instances)");
        instanced_ = true;
        if (logs_)
            printf("transferToGPU: %d objects, %zu instances, instance tree of %zu nodes\n", n, inst.size(), nn);
    }
    // the SNCH half of GPUBvhBuffers::allocate (bvh_interop_structures.h:193-300)
    void uploadSilhouettes(Scene<DIM> &scene)
    {
        const SnchData<DIM> &S = scene.snch();
        size_t n = S.sil_count.size();
        std::vector<fcpw_hip_bounding_cone> cones(n);
        for (size_t i = 0; i < n; i++) {
            const float *k = &S.cone[i * 5];
            cones[i] = fcpw_hip_bounding_cone{{k[0], k[1], k[2]}, k[3], k[4]};
        }
        std::vector<uint8_t> rec = snch_records<DIM>(S, scene.positions().data());
        hip_check(ctx_,
                  fcpw_hip_upload_silhouettes(ctx_, (int)n, cones.data(), S.sil_count.data(), S.sil_offset.data(),
                                              (int)S.ref_obj.size(), rec.data()),
                  "transferToGPU(silhouettes)");
    }
    std::string dir_;
    bool logs_;
    int device_ = 0;
    fcpw_hip_ctx *ctx_ = nullptr;
    std::vector<fcpw_hip_ctx *> objs_; // per-object contexts of an instanced scene
    bool instanced_ = false;
};

// Several GPUScenes queried in one persistent launch (no reference counterpart:
// the reference GPU path serves one object per GPUScene, fcpw_gpu.h:17-20).
// The scenes must outlive the batch.  Queries are concatenated in scene order,
// counts[s] for scene s; see fcpw_hip_batch_* in include/fcpw_hip.h.
template <int DIM>
class GPUSceneBatch {
public:
    explicit GPUSceneBatch(const std::vector<GPUScene<DIM> *> &scenes)
    {
        std::vector<fcpw_hip_ctx *> ctxs;
        for (GPUScene<DIM> *g : scenes) ctxs.push_back(g->context());
        check(fcpw_hip_batch_create(ctxs.data(), (int)ctxs.size(), &b_), "GPUSceneBatch");
        n_scenes_ = (int)scenes.size();
    }
    ~GPUSceneBatch()
    {
        if (b_) fcpw_hip_batch_destroy(b_);
    }
    GPUSceneBatch(const GPUSceneBatch &) = delete;
    GPUSceneBatch &operator=(const GPUSceneBatch &) = delete;

    int nScenes() const { return n_scenes_; }
    void setQueries(const std::vector<int64_t> &counts)
    {
        if ((int)counts.size() != n_scenes_) throw std::invalid_argument("setQueries: one count per scene");
        check(fcpw_hip_batch_set_queries(b_, counts.data()), "setQueries");
    }
    void findClosestPointsDevice(const GPUBoundingSphere *spheres, GPUInteraction *out, bool recordNormals,
                                 void *stream)
    {
        check(fcpw_hip_batch_closest_points_device(b_, (const fcpw_hip_bounding_sphere *)spheres,
                                                   recordNormals ? 1 : 0, (fcpw_hip_interaction *)out, stream),
              "findClosestPointsDevice");
    }
    void closestPointsCompactDevice(const float *points, const float *r2, float *d, float *p, int32_t *prim,
                                    int32_t *slot, void *stream)
    {
        check(fcpw_hip_batch_closest_points_compact_device(b_, points, r2, d, p, prim, slot, stream),
              "closestPointsCompactDevice");
    }
    void distanceBackwardDevice(const float *points, const int32_t *slot, const float *gradD, float *gradPoints,
                                float *gradVertices, void *stream)
    {
        check(fcpw_hip_batch_distance_backward_device(b_, points, slot, gradD, gradPoints, gradVertices, stream),
              "distanceBackwardDevice");
    }
    void setOption(const std::string &key, int value) { check(fcpw_hip_batch_set_option(b_, key.c_str(), value), "setOption"); }
    int getOption(const std::string &key) const
    {
        int v = 0;
        check(fcpw_hip_batch_get_option(b_, key.c_str(), &v), "getOption");
        return v;
    }

private:
    void check(int status, const char *what) const
    {
        if (status == FCPW_HIP_OK) return;
        std::string msg = std::string(what) + ": " + fcpw_hip_batch_last_error(b_);
        if (status == FCPW_HIP_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }
    fcpw_hip_batch *b_ = nullptr;
    int n_scenes_ = 0;
};

} // namespace fcpw_amd
