// CPU restatement of the Box2D v3.2 arithmetic the GPU world step evaluates per lane.
//
// TEST INFRASTRUCTURE ONLY. This library is a checker: only tests/, the smoke entry point and
// the benchmark's CPU-baseline leg may load it, and nothing in the product links or calls it.
//
// Every function restates a reference routine with the reference's operation order so that,
// compiled with -ffp-contract=off, its float results are bit-identical to the reference's. Each
// definition cites the reference file and lines it follows. Value types mirror the reference's
// public structs field for field (include/box2d/math_functions.h, collision.h), so arrays can be
// passed in from the same buffers the GPU library takes.
//
// Pinning (what checks each function independently of this file):
//   o_atan2, o_cos_sin          - reference library outputs over a dense grid, libm atan2/cos/sin
//                                 within the approximations' published error, known angles
//   o_segment_distance          - reference library outputs, the exact values of
//                                 test/test_distance.c:12-43, brute-force float64 minimum
//   o_collide_* / batch         - reference library outputs on the 5400-pair fixture, the
//                                 known answer of test/test_collision.c:202-216, invariants
//                                 (unit normal, at most 2 points, ids in feature range)
//   o_make_soft                 - closed form of the soft constraint coefficients in float64
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ov2
{
	float x, y;
} ov2;

typedef struct orot
{
	float c, s;
} orot;

typedef struct oxf
{
	ov2 p;
	orot q;
} oxf;

typedef struct ocircle
{
	ov2 center;
	float radius;
} ocircle;

typedef struct ocapsule
{
	ov2 c1, c2;
	float radius;
} ocapsule;

typedef struct osegment
{
	ov2 p1, p2;
} osegment;

typedef struct opolygon
{
	ov2 v[8];
	ov2 n[8];
	ov2 centroid;
	float radius;
	int count;
} opolygon;

// Same layout as b2amd_Shape: b2ShapeType then the geometry.
typedef struct oshape
{
	int type;
	int padding;
	union
	{
		ocircle circle;
		ocapsule capsule;
		opolygon polygon;
		osegment segment;
	} u;
} oshape;

typedef struct omanifold_point
{
	ov2 point;
	float separation;
	uint16_t id;
} omanifold_point;

typedef struct omanifold
{
	ov2 normal;
	omanifold_point pts[2];
	int count;
} omanifold;

typedef struct osegdist
{
	ov2 closest1, closest2;
	float f1, f2, distSq;
} osegdist;

typedef struct osoft
{
	float biasRate, massScale, impulseScale;
} osoft;

float o_atan2( float y, float x );
void o_cos_sin( float radians, float* c, float* s );
osegdist o_segment_distance( ov2 p1, ov2 q1, ov2 p2, ov2 q2 );
osoft o_make_soft( float hertz, float zeta, float h );

omanifold o_collide_circles( const ocircle* a, const ocircle* b, oxf xf );
omanifold o_collide_capsule_circle( const ocapsule* a, const ocircle* b, oxf xf );
omanifold o_collide_polygon_circle( const opolygon* a, const ocircle* b, oxf xf );
omanifold o_collide_capsules( const ocapsule* a, const ocapsule* b, oxf xf );
omanifold o_collide_polygons( const opolygon* a, const opolygon* b, oxf xf );
omanifold o_collide_polygon_capsule( const opolygon* a, const ocapsule* b, oxf xf );
omanifold o_collide_segment_circle( const osegment* a, const ocircle* b, oxf xf );
omanifold o_collide_segment_capsule( const osegment* a, const ocapsule* b, oxf xf );
omanifold o_collide_segment_polygon( const osegment* a, const opolygon* b, oxf xf );

// Manifold i of shapesA[i] against shapesB[i] posed by xfs[i] in A's frame, pairs in the
// reference's primary order. Returns 0, or -1 for an unsupported pair.
int b2o_collide_batch( const oshape* shapesA, const oshape* shapesB, const oxf* xfs, int count, omanifold* out );

// Whole arrays for test sweeps.
void b2o_atan2_batch( const float* y, const float* x, float* out, int count );
void b2o_cos_sin_batch( const float* radians, float* c, float* s, int count );
void b2o_segment_distance_batch( const ov2* p1, const ov2* q1, const ov2* p2, const ov2* q2, osegdist* out, int count );

#ifdef __cplusplus
}
#endif
