// C++ drop-in surface with the reference's host signatures, implemented over
// the C ABI of gcdt.h.  Each function keeps the reference's argument meaning
// and returns the elapsed GPU milliseconds like the reference does; a non-OK
// status becomes a std::runtime_error carrying gcdt_last_error().
//
//   reference declaration                              here
//   include/rawMesh.h:4   rawMeshBefore               rawMeshBefore
//   include/rawMesh.h:5   rawMeshAfter                rawMeshAfter
//   include/rawMesh.h:6   getSamplePoints             getSamplePoints
//   include/rawMesh.h:7   rawMesh                     rawMesh
//   include/boostmap.h:3  getBitmap                   getBitmap
//   include/addConstraints.h:5  addConstraints (14 args)   addConstraints
//   include/addConstraints.h:6  addConstraints (10 args)   addConstraints
//                               (the reference's body is a stub that returns 0
//                               and changes nothing, addConstraints.cu:1837-1840;
//                               kept as the same no-op so callers link)
//   include/fastDelaunay.h:5  tryFastDelaunay         tryFastDelaunay
//   kernel.cu:5038 delaunaySample, :4697 delaunayNew, :5293 filterTriangles,
//   :833 convertSample (header-less in the reference)
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "MemCache.h"
#include "gcdt.h"

typedef int2 Point2d;  // include/math_base.cuh:29
static_assert(sizeof(Point2d) == sizeof(gcdt_int2), "int2 layout");
static_assert(sizeof(int3) == sizeof(gcdt_int3), "int3 layout");

namespace gcdt_api {
inline void check(int status, const char* where)
{
    if (status != GCDT_OK) throw std::runtime_error(std::string(where) + ": " + gcdt_last_error());
}
inline gcdt_int2* P(Point2d* p) { return reinterpret_cast<gcdt_int2*>(p); }
inline gcdt_int3* T(int3* p) { return reinterpret_cast<gcdt_int3*>(p); }
}  // namespace gcdt_api

inline double getSamplePoints(void* tmp, size_t bytes, Point2d* points, int numPoints, Point2d*& samplePoints,
                              int*& sample, int& num)
{
    double ms = 0;
    gcdt_int2* sp = nullptr;
    gcdt_api::check(gcdt_get_sample_points(tmp, bytes, gcdt_api::P(points), numPoints, &sp, &sample, &num, &ms),
                    "getSamplePoints");
    samplePoints = reinterpret_cast<Point2d*>(sp);
    return ms;
}

inline double rawMeshBefore(void* tmp, size_t bytes, Point2d* points, int3* tris, int* sons, int& numTriangles,
                            int numPoints)
{
    double ms = 0;
    gcdt_api::check(gcdt_raw_mesh_before(tmp, bytes, gcdt_api::P(points), gcdt_api::T(tris), sons, &numTriangles,
                                         numPoints, &ms),
                    "rawMeshBefore");
    return ms;
}

inline double rawMeshAfter(void* tmp, size_t bytes, Point2d* points, int3* tris, int3* adjTris, int* sons,
                           int* bitmapOffsets, int* bitmapTriIds, int& numTriangles, int numPoints)
{
    double ms = 0;
    gcdt_api::check(gcdt_raw_mesh_after(tmp, bytes, gcdt_api::P(points), gcdt_api::T(tris), gcdt_api::T(adjTris), sons,
                                        bitmapOffsets, bitmapTriIds, &numTriangles, numPoints, &ms),
                    "rawMeshAfter");
    return ms;
}

inline void rawMesh(void* tmp, size_t bytes, Point2d* points, int3* tris, int3* adjTris, int* sons, int& numTriangles,
                    int numPoints)
{
    gcdt_api::check(gcdt_raw_mesh(tmp, bytes, gcdt_api::P(points), gcdt_api::T(tris), gcdt_api::T(adjTris), sons,
                                  &numTriangles, numPoints, nullptr),
                    "rawMesh");
}

inline double getBitmap(void* tmp, size_t bytes, Point2d* points, int3* tris, int*& bitmapOffsets, int*& bitmapTris,
                        int numTriangles)
{
    double ms = 0;
    gcdt_api::check(gcdt_get_bitmap(tmp, bytes, gcdt_api::P(points), gcdt_api::T(tris), &bitmapOffsets, &bitmapTris,
                                    numTriangles, &ms),
                    "getBitmap");
    return ms;
}

inline double addConstraints(void* tmp, size_t bytes, Point2d* points, int3* tris, int3* adjTris, int* sons,
                             int* bitmapOffsets, int* bitmapTris, int2* cons, int3*& newTris, int numCons,
                             int numPoints, int numTriangles, int& numNewTris)
{
    double ms = 0;
    gcdt_int3* nt = nullptr;
    gcdt_api::check(gcdt_add_constraints(tmp, bytes, gcdt_api::P(points), gcdt_api::T(tris), gcdt_api::T(adjTris), sons,
                                         bitmapOffsets, bitmapTris, reinterpret_cast<gcdt_int2*>(cons), &nt, numCons,
                                         numPoints, numTriangles, &numNewTris, &ms),
                    "addConstraints");
    newTris = reinterpret_cast<int3*>(nt);
    return ms;
}

// include/addConstraints.h:6 -- no-op, as in the reference (addConstraints.cu:1837-1840)
inline double addConstraints(void*, size_t, Point2d*, int3*, int3*, int*, int2*, int, int, int) { return 0.0; }

inline double tryFastDelaunay(void* tmp, size_t bytes, Point2d* points, int3* tris, int3* adjTris, int numPoints,
                              int numTriangles)
{
    double ms = 0;
    gcdt_api::check(gcdt_try_fast_delaunay(tmp, bytes, gcdt_api::P(points), gcdt_api::T(tris), gcdt_api::T(adjTris),
                                           numPoints, numTriangles, &ms),
                    "tryFastDelaunay");
    return ms;
}

inline double delaunaySample(void* tmp, size_t bytes, Point2d* points, int3* tris, int3* adjTris, int numPoints,
                             int numTriangles)
{
    double ms = 0;
    gcdt_api::check(gcdt_delaunay_sample(tmp, bytes, gcdt_api::P(points), gcdt_api::T(tris), gcdt_api::T(adjTris),
                                         numPoints, numTriangles, &ms),
                    "delaunaySample");
    return ms;
}

inline double delaunayNew(void* tmp, size_t bytes, Point2d*& points, int3*& tris, int2* cons, int numPoints,
                          int numTriangles, int numCons)
{
    double ms = 0;
    gcdt_api::check(gcdt_delaunay_new(tmp, bytes, reinterpret_cast<gcdt_int2**>(&points),
                                      reinterpret_cast<gcdt_int3**>(&tris), reinterpret_cast<gcdt_int2*>(cons),
                                      numPoints, numTriangles, numCons, &ms),
                    "delaunayNew");
    return ms;
}

inline double filterTriangles(void* tmp, size_t bytes, int3* tris, int3* newTris, int3* adjTris, int3* newAdjTris,
                              int* sons, int& numTriangles)
{
    double ms = 0;
    gcdt_api::check(gcdt_filter_triangles(tmp, bytes, gcdt_api::T(tris), gcdt_api::T(newTris), gcdt_api::T(adjTris),
                                          gcdt_api::T(newAdjTris), sons, &numTriangles, &ms),
                    "filterTriangles");
    return ms;
}

inline void convertSample(int3* tris, int* sampleId, int numSamples, int numPoints, int numT)
{
    gcdt_api::check(gcdt_convert_sample(gcdt_api::T(tris), sampleId, numSamples, numPoints, numT), "convertSample");
}

// utils::release(ptr, count) of the reference (include/utils.cuh:33-39)
template <class T>
inline void gcdtRelease(T*& ptr, size_t count)
{
    if (ptr) {
        gcdt_release(ptr, count * sizeof(T));
        ptr = nullptr;
    }
}
