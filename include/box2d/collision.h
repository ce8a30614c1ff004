// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the declarations below follow its public headers)
// SPDX-License-Identifier: MIT
// box2d_amd public header: geometry and contact manifold types.
//
// Drop-in counterpart of the subset of the reference's include/box2d/collision.h that the world
// step uses: shape primitives, hulls, mass properties, shape AABBs and the manifold layout that
// b2Contact_GetData / events report. Struct layouts match the reference (collision.h:124-197,
// 527-586).
#pragma once

#include "base.h"
#include "math_functions.h"

#include <stdbool.h>

#define B2_MAX_POLYGON_VERTICES 8

typedef struct b2ShapeProxy
{
	b2Vec2 points[B2_MAX_POLYGON_VERTICES];
	int count;
	float radius;
} b2ShapeProxy;

typedef struct b2MassData
{
	float mass;
	b2Vec2 center;
	float rotationalInertia;
} b2MassData;

typedef struct b2Circle
{
	b2Vec2 center;
	float radius;
} b2Circle;

typedef struct b2Capsule
{
	b2Vec2 center1;
	b2Vec2 center2;
	float radius;
} b2Capsule;

typedef struct b2Polygon
{
	b2Vec2 vertices[B2_MAX_POLYGON_VERTICES];
	b2Vec2 normals[B2_MAX_POLYGON_VERTICES];
	b2Vec2 centroid;
	float radius;
	int count;
} b2Polygon;

typedef struct b2Segment
{
	b2Vec2 point1;
	b2Vec2 point2;
} b2Segment;

typedef struct b2ChainSegment
{
	b2Vec2 ghost1;
	b2Segment segment;
	b2Vec2 ghost2;
	int chainId;
} b2ChainSegment;

typedef struct b2Hull
{
	b2Vec2 points[B2_MAX_POLYGON_VERTICES];
	int count;
} b2Hull;

/// Polygon construction (reference: src/geometry.c:57-252)
B2_API b2Polygon b2MakePolygon( const b2Hull* hull, float radius );
B2_API b2Polygon b2MakeOffsetPolygon( const b2Hull* hull, b2Vec2 position, b2Rot rotation );
B2_API b2Polygon b2MakeOffsetRoundedPolygon( const b2Hull* hull, b2Vec2 position, b2Rot rotation, float radius );
B2_API b2Polygon b2MakeSquare( float halfWidth );
B2_API b2Polygon b2MakeBox( float halfWidth, float halfHeight );
B2_API b2Polygon b2MakeRoundedBox( float halfWidth, float halfHeight, float radius );
B2_API b2Polygon b2MakeOffsetBox( float halfWidth, float halfHeight, b2Vec2 center, b2Rot rotation );
B2_API b2Polygon b2MakeOffsetRoundedBox( float halfWidth, float halfHeight, b2Vec2 center, b2Rot rotation, float radius );
B2_API b2Polygon b2TransformPolygon( b2Transform transform, const b2Polygon* polygon );

/// Mass properties (reference: src/geometry.c:254-400)
B2_API b2MassData b2ComputeCircleMass( const b2Circle* shape, float density );
B2_API b2MassData b2ComputeCapsuleMass( const b2Capsule* shape, float density );
B2_API b2MassData b2ComputePolygonMass( const b2Polygon* shape, float density );

/// Shape bounding boxes (reference: src/geometry.c:402-482)
B2_API b2AABB b2ComputeCircleAABB( const b2Circle* shape, b2WorldTransform transform );
B2_API b2AABB b2ComputeCapsuleAABB( const b2Capsule* shape, b2WorldTransform transform );
B2_API b2AABB b2ComputePolygonAABB( const b2Polygon* shape, b2WorldTransform transform );
B2_API b2AABB b2ComputeSegmentAABB( const b2Segment* shape, b2WorldTransform transform );

/// Convex hull with quickhull and collinear point removal (reference: src/hull.c:87-328)
B2_API b2Hull b2ComputeHull( const b2Vec2* points, int count );
B2_API bool b2ValidateHull( const b2Hull* hull );

typedef struct b2SegmentDistanceResult
{
	b2Vec2 closest1;
	b2Vec2 closest2;
	float fraction1;
	float fraction2;
	float distanceSquared;
} b2SegmentDistanceResult;

typedef struct b2SimplexCache
{
	uint16_t count;
	uint8_t indexA[3];
	uint8_t indexB[3];
} b2SimplexCache;

static const b2SimplexCache b2_emptySimplexCache = B2_ZERO_INIT;

/// A contact point in world space relative to each body's center of mass
/// (reference: include/box2d/collision.h:527-560)
typedef struct b2ManifoldPoint
{
	b2Vec2 anchorA;
	b2Vec2 anchorB;
	float separation;
	float baseSeparation;
	float normalImpulse;
	float tangentImpulse;
	float totalNormalImpulse;
	float normalVelocity;
	uint16_t id;
	bool persisted;
} b2ManifoldPoint;

typedef struct b2Manifold
{
	b2Vec2 normal;
	float rollingImpulse;
	b2ManifoldPoint points[2];
	int pointCount;
} b2Manifold;

typedef struct b2LocalManifoldPoint
{
	b2Vec2 point;
	float separation;
	uint16_t id;
} b2LocalManifoldPoint;

typedef struct b2LocalManifold
{
	b2Vec2 normal;
	b2LocalManifoldPoint points[2];
	int pointCount;
} b2LocalManifold;

/// Closest points between segments p1-q1 and p2-q2 (reference include/box2d/collision.h, src/distance.c:33-106).
B2_API b2SegmentDistanceResult b2SegmentDistance( b2Vec2 p1, b2Vec2 q1, b2Vec2 p2, b2Vec2 q2 );

/// Manifold functions: contact manifold of shape B, posed by xf in the frame of shape A, against
/// shape A (reference include/box2d/collision.h:615-640). These are the host forms of the
/// narrowphase the GPU step evaluates one lane per contact.
B2_API b2LocalManifold b2CollideCircles( const b2Circle* circleA, const b2Circle* circleB, b2Transform xf );
B2_API b2LocalManifold b2CollideCapsuleAndCircle( const b2Capsule* capsuleA, const b2Circle* circleB, b2Transform xf );
B2_API b2LocalManifold b2CollideSegmentAndCircle( const b2Segment* segmentA, const b2Circle* circleB, b2Transform xf );
B2_API b2LocalManifold b2CollidePolygonAndCircle( const b2Polygon* polygonA, const b2Circle* circleB, b2Transform xf );
B2_API b2LocalManifold b2CollideCapsules( const b2Capsule* capsuleA, const b2Capsule* capsuleB, b2Transform xf );
B2_API b2LocalManifold b2CollideSegmentAndCapsule( const b2Segment* segmentA, const b2Capsule* capsuleB, b2Transform xf );
B2_API b2LocalManifold b2CollidePolygonAndCapsule( const b2Polygon* polygonA, const b2Capsule* capsuleB, b2Transform xf );
B2_API b2LocalManifold b2CollidePolygons( const b2Polygon* polygonA, const b2Polygon* polygonB, b2Transform xf );
B2_API b2LocalManifold b2CollideSegmentAndPolygon( const b2Segment* segmentA, const b2Polygon* polygonB, b2Transform xf );

/// Input for b2ShapeDistance (reference include/box2d/collision.h:382-396).
typedef struct b2DistanceInput
{
	b2ShapeProxy proxyA;
	b2ShapeProxy proxyB;
	/// pose of B in A's frame
	b2Transform transform;
	bool useRadii;
} b2DistanceInput;

/// Output of b2ShapeDistance, in shape A's frame.
typedef struct b2DistanceOutput
{
	b2Vec2 pointA;
	b2Vec2 pointB;
	b2Vec2 normal;
	float distance;
	int iterations;
	int simplexCount;
} b2DistanceOutput;

typedef struct b2SimplexVertex
{
	b2Vec2 wA;
	b2Vec2 wB;
	b2Vec2 w;
	float a;
	int indexA;
	int indexB;
} b2SimplexVertex;

typedef struct b2Simplex
{
	b2SimplexVertex v1, v2, v3;
	int count;
} b2Simplex;

/// Center-of-mass sweep of a body over a step (reference include/box2d/collision.h:457-464).
typedef struct b2Sweep
{
	b2Vec2 localCenter;
	b2Vec2 c1;
	b2Vec2 c2;
	b2Rot q1;
	b2Rot q2;
} b2Sweep;

typedef struct b2TOIInput
{
	b2ShapeProxy proxyA;
	b2ShapeProxy proxyB;
	b2Sweep sweepA;
	b2Sweep sweepB;
	float maxFraction;
} b2TOIInput;

typedef enum b2TOIState
{
	b2_toiStateUnknown,
	b2_toiStateFailed,
	b2_toiStateOverlapped,
	b2_toiStateHit,
	b2_toiStateSeparated
} b2TOIState;

typedef struct b2TOIOutput
{
	b2TOIState state;
	b2Vec2 point;
	b2Vec2 normal;
	float fraction;
} b2TOIOutput;

/// Distance proxies and GJK distance (reference src/distance.c:108-611).
B2_API b2ShapeProxy b2MakeProxy( const b2Vec2* points, int count, float radius );
B2_API b2ShapeProxy b2MakeOffsetProxy( const b2Vec2* points, int count, float radius, b2Vec2 position, b2Rot rotation );
B2_API b2DistanceOutput b2ShapeDistance( const b2DistanceInput* input, b2SimplexCache* cache, b2Simplex* simplexes,
										 int simplexCapacity );

/// Sweep interpolation and conservative-advancement time of impact (reference src/distance.c:14-30, 1148-1441).
B2_API b2Transform b2GetSweepTransform( const b2Sweep* sweep, float time );
B2_API b2TOIOutput b2TimeOfImpact( const b2TOIInput* input );

/// One-sided chain segment manifolds; capsule and polygon forms carry the contact's GJK simplex
/// cache (reference include/box2d/collision.h:643-652).
B2_API b2LocalManifold b2CollideChainSegmentAndCircle( const b2ChainSegment* segmentA, const b2Circle* circleB, b2Transform xf );
B2_API b2LocalManifold b2CollideChainSegmentAndCapsule( const b2ChainSegment* segmentA, const b2Capsule* capsuleB, b2Transform xf,
														b2SimplexCache* cache );
B2_API b2LocalManifold b2CollideChainSegmentAndPolygon( const b2ChainSegment* segmentA, const b2Polygon* polygonB, b2Transform xf,
														b2SimplexCache* cache );
