// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the manifold arithmetic below restates its
// src/manifold.c and src/distance.c in the same operation order)
// SPDX-License-Identifier: MIT
// GPU narrowphase: contact manifolds for one shape pair per lane.
//
// Restated from the reference manifold routines (src/manifold.c) with the same operation order
// and the same feature ids, so a lane produces the bit-identical b2LocalManifold the reference
// computes for the same pair. All arithmetic is FP32 with contraction disabled at compile time.
//   circles                 src/manifold.c:36-66
//   capsule vs circle       src/manifold.c:68-125
//   polygon vs circle       src/manifold.c:127-234
//   capsule vs capsule      src/manifold.c:237-496
//   polygon vs capsule      src/manifold.c:504-508 (capsule as a two-vertex polygon, :16-34)
//   polygon clipping        src/manifold.c:511-643
//   max separation          src/manifold.c:646-682
//   polygon vs polygon      src/manifold.c:702-1028
//   segment distance        src/distance.c:33-106
#pragma once

#include "layout.h"

#include <cfloat>

namespace b2g
{

#define B2G_DEV __host__ __device__ __forceinline__
#define B2G_MAKE_ID( A, B ) ( (uint16_t)( ( (uint8_t)( A ) << 8 ) | (uint8_t)( B ) ) )

constexpr float kLinearSlop = 0.005f;
constexpr float kSpeculativeDistance = 4.0f * kLinearSlop;

B2G_DEV b2LocalManifold EmptyManifold()
{
	b2LocalManifold m;
	m.normal = { 0.0f, 0.0f };
	m.points[0].point = { 0.0f, 0.0f };
	m.points[0].separation = 0.0f;
	m.points[0].id = 0;
	m.points[1] = m.points[0];
	m.pointCount = 0;
	return m;
}

B2G_DEV b2SegmentDistanceResult SegmentDistance( b2Vec2 p1, b2Vec2 q1, b2Vec2 p2, b2Vec2 q2 )
{
	b2SegmentDistanceResult result;
	result.fraction1 = 0.0f;
	result.fraction2 = 0.0f;

	b2Vec2 d1 = b2Sub( q1, p1 );
	b2Vec2 d2 = b2Sub( q2, p2 );
	b2Vec2 r = b2Sub( p1, p2 );
	float dd1 = b2Dot( d1, d1 );
	float dd2 = b2Dot( d2, d2 );
	float rd1 = b2Dot( r, d1 );
	float rd2 = b2Dot( r, d2 );

	const float epsSqr = FLT_EPSILON * FLT_EPSILON;

	if ( dd1 < epsSqr || dd2 < epsSqr )
	{
		if ( dd1 >= epsSqr )
		{
			result.fraction1 = b2ClampFloat( -rd1 / dd1, 0.0f, 1.0f );
			result.fraction2 = 0.0f;
		}
		else if ( dd2 >= epsSqr )
		{
			result.fraction1 = 0.0f;
			result.fraction2 = b2ClampFloat( rd2 / dd2, 0.0f, 1.0f );
		}
	}
	else
	{
		float d12 = b2Dot( d1, d2 );
		float denominator = dd1 * dd2 - d12 * d12;

		float f1 = 0.0f;
		if ( denominator != 0.0f )
		{
			f1 = b2ClampFloat( ( d12 * rd2 - rd1 * dd2 ) / denominator, 0.0f, 1.0f );
		}

		float f2 = ( d12 * f1 + rd2 ) / dd2;

		if ( f2 < 0.0f )
		{
			f2 = 0.0f;
			f1 = b2ClampFloat( -rd1 / dd1, 0.0f, 1.0f );
		}
		else if ( f2 > 1.0f )
		{
			f2 = 1.0f;
			f1 = b2ClampFloat( ( d12 - rd1 ) / dd1, 0.0f, 1.0f );
		}

		result.fraction1 = f1;
		result.fraction2 = f2;
	}

	result.closest1 = b2MulAdd( p1, result.fraction1, d1 );
	result.closest2 = b2MulAdd( p2, result.fraction2, d2 );
	result.distanceSquared = b2DistanceSquared( result.closest1, result.closest2 );
	return result;
}

B2G_DEV b2LocalManifold CollideCircles( const b2Circle& circleA, const b2Circle& circleB, b2Transform xf )
{
	b2LocalManifold manifold = EmptyManifold();

	b2Vec2 pointA = circleA.center;
	b2Vec2 pointB = b2TransformPoint( xf, circleB.center );

	float distance;
	b2Vec2 normal = b2GetLengthAndNormalize( &distance, b2Sub( pointB, pointA ) );

	float radiusA = circleA.radius;
	float radiusB = circleB.radius;
	float separation = distance - radiusA - radiusB;
	if ( separation > kSpeculativeDistance )
	{
		return manifold;
	}

	b2Vec2 cA = b2MulAdd( pointA, radiusA, normal );
	b2Vec2 cB = b2MulAdd( pointB, -radiusB, normal );
	manifold.normal = normal;
	manifold.points[0].point = b2Lerp( cA, cB, 0.5f );
	manifold.points[0].separation = separation;
	manifold.points[0].id = 0;
	manifold.pointCount = 1;
	return manifold;
}

B2G_DEV b2LocalManifold CollideCapsuleAndCircle( const b2Capsule& capsuleA, const b2Circle& circleB, b2Transform xf )
{
	b2LocalManifold manifold = EmptyManifold();

	b2Vec2 pB = b2TransformPoint( xf, circleB.center );
	b2Vec2 p1 = capsuleA.center1;
	b2Vec2 p2 = capsuleA.center2;
	b2Vec2 e = b2Sub( p2, p1 );

	b2Vec2 pA;
	float s1 = b2Dot( b2Sub( pB, p1 ), e );
	float s2 = b2Dot( b2Sub( p2, pB ), e );
	if ( s1 < 0.0f )
	{
		pA = p1;
	}
	else if ( s2 < 0.0f )
	{
		pA = p2;
	}
	else
	{
		float s = s1 / b2Dot( e, e );
		pA = b2MulAdd( p1, s, e );
	}

	float distance;
	b2Vec2 normal = b2GetLengthAndNormalize( &distance, b2Sub( pB, pA ) );

	float radiusA = capsuleA.radius;
	float radiusB = circleB.radius;
	float separation = distance - radiusA - radiusB;
	if ( separation > kSpeculativeDistance )
	{
		return manifold;
	}

	b2Vec2 cA = b2MulAdd( pA, radiusA, normal );
	b2Vec2 cB = b2MulAdd( pB, -radiusB, normal );
	manifold.normal = normal;
	manifold.points[0].point = b2Lerp( cA, cB, 0.5f );
	manifold.points[0].separation = separation;
	manifold.points[0].id = 0;
	manifold.pointCount = 1;
	return manifold;
}

B2G_DEV b2LocalManifold CollidePolygonAndCircle( const b2Polygon& polygonA, const b2Circle& circleB, b2Transform xf )
{
	b2LocalManifold manifold = EmptyManifold();

	b2Vec2 center = b2TransformPoint( xf, circleB.center );
	float radiusA = polygonA.radius;
	float radiusB = circleB.radius;
	float radius = radiusA + radiusB;

	int normalIndex = 0;
	float separation = -FLT_MAX;
	int vertexCount = polygonA.count;
	for ( int i = 0; i < vertexCount; ++i )
	{
		float s = b2Dot( polygonA.normals[i], b2Sub( center, polygonA.vertices[i] ) );
		if ( s > separation )
		{
			separation = s;
			normalIndex = i;
		}
	}

	if ( separation > radius + kSpeculativeDistance )
	{
		return manifold;
	}

	int vertIndex1 = normalIndex;
	int vertIndex2 = vertIndex1 + 1 < vertexCount ? vertIndex1 + 1 : 0;
	b2Vec2 v1 = polygonA.vertices[vertIndex1];
	b2Vec2 v2 = polygonA.vertices[vertIndex2];

	float u1 = b2Dot( b2Sub( center, v1 ), b2Sub( v2, v1 ) );
	float u2 = b2Dot( b2Sub( center, v2 ), b2Sub( v1, v2 ) );

	if ( u1 < 0.0f && separation > FLT_EPSILON )
	{
		// closest to v1, outside the polygon
		b2Vec2 normal = b2Normalize( b2Sub( center, v1 ) );
		separation = b2Dot( b2Sub( center, v1 ), normal );
		if ( separation > radius + kSpeculativeDistance )
		{
			return manifold;
		}
		b2Vec2 cA = b2MulAdd( v1, radiusA, normal );
		b2Vec2 cB = b2MulSub( center, radiusB, normal );
		manifold.normal = normal;
		manifold.points[0].point = b2Lerp( cA, cB, 0.5f );
		manifold.points[0].separation = b2Dot( b2Sub( cB, cA ), normal );
		manifold.points[0].id = 0;
		manifold.pointCount = 1;
	}
	else if ( u2 < 0.0f && separation > FLT_EPSILON )
	{
		// closest to v2, outside the polygon
		b2Vec2 normal = b2Normalize( b2Sub( center, v2 ) );
		separation = b2Dot( b2Sub( center, v2 ), normal );
		if ( separation > radius + kSpeculativeDistance )
		{
			return manifold;
		}
		b2Vec2 cA = b2MulAdd( v2, radiusA, normal );
		b2Vec2 cB = b2MulSub( center, radiusB, normal );
		manifold.normal = normal;
		manifold.points[0].point = b2Lerp( cA, cB, 0.5f );
		manifold.points[0].separation = b2Dot( b2Sub( cB, cA ), normal );
		manifold.points[0].id = 0;
		manifold.pointCount = 1;
	}
	else
	{
		// center projects onto the face
		b2Vec2 normal = polygonA.normals[normalIndex];
		manifold.normal = normal;
		b2Vec2 cA = b2MulAdd( center, radiusA - b2Dot( b2Sub( center, v1 ), normal ), normal );
		b2Vec2 cB = b2MulSub( center, radiusB, normal );
		manifold.points[0].point = b2Lerp( cA, cB, 0.5f );
		manifold.points[0].separation = separation - radius;
		manifold.points[0].id = 0;
		manifold.pointCount = 1;
	}

	return manifold;
}

B2G_DEV b2LocalManifold CollideCapsules( const b2Capsule& capsuleA, const b2Capsule& capsuleB, b2Transform xf )
{
	b2Vec2 origin = capsuleA.center1;
	b2Transform xfs = { b2Sub( xf.p, origin ), xf.q };

	b2Vec2 p1 = { 0.0f, 0.0f };
	b2Vec2 q1 = b2Sub( capsuleA.center2, origin );
	b2Vec2 p2 = b2TransformPoint( xfs, capsuleB.center1 );
	b2Vec2 q2 = b2TransformPoint( xfs, capsuleB.center2 );

	b2Vec2 d1 = b2Sub( q1, p1 );
	b2Vec2 d2 = b2Sub( q2, p2 );
	float dd1 = b2Dot( d1, d1 );
	float dd2 = b2Dot( d2, d2 );
	const float epsSqr = FLT_EPSILON * FLT_EPSILON;

	b2Vec2 r = b2Sub( p1, p2 );
	float rd1 = b2Dot( r, d1 );
	float rd2 = b2Dot( r, d2 );
	float d12 = b2Dot( d1, d2 );
	float denom = dd1 * dd2 - d12 * d12;

	float f1 = 0.0f;
	if ( denom != 0.0f )
	{
		f1 = b2ClampFloat( ( d12 * rd2 - rd1 * dd2 ) / denom, 0.0f, 1.0f );
	}

	float f2 = ( d12 * f1 + rd2 ) / dd2;
	if ( f2 < 0.0f )
	{
		f2 = 0.0f;
		f1 = b2ClampFloat( -rd1 / dd1, 0.0f, 1.0f );
	}
	else if ( f2 > 1.0f )
	{
		f2 = 1.0f;
		f1 = b2ClampFloat( ( d12 - rd1 ) / dd1, 0.0f, 1.0f );
	}

	b2Vec2 closest1 = b2MulAdd( p1, f1, d1 );
	b2Vec2 closest2 = b2MulAdd( p2, f2, d2 );
	float distanceSquared = b2DistanceSquared( closest1, closest2 );

	b2LocalManifold manifold = EmptyManifold();
	float radiusA = capsuleA.radius;
	float radiusB = capsuleB.radius;
	float radius = radiusA + radiusB;
	float maxDistance = radius + kSpeculativeDistance;

	if ( distanceSquared > maxDistance * maxDistance )
	{
		return manifold;
	}

	float distance = sqrtf( distanceSquared );

	float length1, length2;
	b2Vec2 u1 = b2GetLengthAndNormalize( &length1, d1 );
	b2Vec2 u2 = b2GetLengthAndNormalize( &length2, d2 );

	float fp2 = b2Dot( b2Sub( p2, p1 ), u1 );
	float fq2 = b2Dot( b2Sub( q2, p1 ), u1 );
	bool outsideA = ( fp2 <= 0.0f && fq2 <= 0.0f ) || ( fp2 >= length1 && fq2 >= length1 );

	float fp1 = b2Dot( b2Sub( p1, p2 ), u2 );
	float fq1 = b2Dot( b2Sub( q1, p2 ), u2 );
	bool outsideB = ( fp1 <= 0.0f && fq1 <= 0.0f ) || ( fp1 >= length2 && fq1 >= length2 );

	if ( outsideA == false && outsideB == false )
	{
		b2Vec2 normalA;
		float separationA;
		{
			normalA = b2LeftPerp( u1 );
			float ss1 = b2Dot( b2Sub( p2, p1 ), normalA );
			float ss2 = b2Dot( b2Sub( q2, p1 ), normalA );
			float s1p = ss1 < ss2 ? ss1 : ss2;
			float s1n = -ss1 < -ss2 ? -ss1 : -ss2;
			if ( s1p > s1n )
			{
				separationA = s1p;
			}
			else
			{
				separationA = s1n;
				normalA = b2Neg( normalA );
			}
		}

		b2Vec2 normalB;
		float separationB;
		{
			normalB = b2LeftPerp( u2 );
			float ss1 = b2Dot( b2Sub( p1, p2 ), normalB );
			float ss2 = b2Dot( b2Sub( q1, p2 ), normalB );
			float s1p = ss1 < ss2 ? ss1 : ss2;
			float s1n = -ss1 < -ss2 ? -ss1 : -ss2;
			if ( s1p > s1n )
			{
				separationB = s1p;
			}
			else
			{
				separationB = s1n;
				normalB = b2Neg( normalB );
			}
		}

		if ( separationA + 0.1f * kLinearSlop >= separationB )
		{
			manifold.normal = normalA;
			b2Vec2 cp = p2;
			b2Vec2 cq = q2;

			if ( fp2 < 0.0f && fq2 > 0.0f )
			{
				cp = b2Lerp( p2, q2, ( 0.0f - fp2 ) / ( fq2 - fp2 ) );
			}
			else if ( fq2 < 0.0f && fp2 > 0.0f )
			{
				cq = b2Lerp( q2, p2, ( 0.0f - fq2 ) / ( fp2 - fq2 ) );
			}

			if ( fp2 > length1 && fq2 < length1 )
			{
				cp = b2Lerp( p2, q2, ( fp2 - length1 ) / ( fp2 - fq2 ) );
			}
			else if ( fq2 > length1 && fp2 < length1 )
			{
				cq = b2Lerp( q2, p2, ( fq2 - length1 ) / ( fq2 - fp2 ) );
			}

			float sp = b2Dot( b2Sub( cp, p1 ), normalA );
			float sq = b2Dot( b2Sub( cq, p1 ), normalA );

			if ( sp <= distance + kLinearSlop || sq <= distance + kLinearSlop )
			{
				manifold.points[0].point = b2MulAdd( cp, 0.5f * ( radiusA - radiusB - sp ), normalA );
				manifold.points[0].separation = sp - radius;
				manifold.points[0].id = B2G_MAKE_ID( 0, 0 );
				manifold.points[1].point = b2MulAdd( cq, 0.5f * ( radiusA - radiusB - sq ), normalA );
				manifold.points[1].separation = sq - radius;
				manifold.points[1].id = B2G_MAKE_ID( 0, 1 );
				manifold.pointCount = 2;
			}
		}
		else
		{
			manifold.normal = b2Neg( normalB );
			b2Vec2 cp = p1;
			b2Vec2 cq = q1;

			if ( fp1 < 0.0f && fq1 > 0.0f )
			{
				cp = b2Lerp( p1, q1, ( 0.0f - fp1 ) / ( fq1 - fp1 ) );
			}
			else if ( fq1 < 0.0f && fp1 > 0.0f )
			{
				cq = b2Lerp( q1, p1, ( 0.0f - fq1 ) / ( fp1 - fq1 ) );
			}

			if ( fp1 > length2 && fq1 < length2 )
			{
				cp = b2Lerp( p1, q1, ( fp1 - length2 ) / ( fp1 - fq1 ) );
			}
			else if ( fq1 > length2 && fp1 < length2 )
			{
				cq = b2Lerp( q1, p1, ( fq1 - length2 ) / ( fq1 - fp1 ) );
			}

			float sp = b2Dot( b2Sub( cp, p2 ), normalB );
			float sq = b2Dot( b2Sub( cq, p2 ), normalB );

			if ( sp <= distance + kLinearSlop || sq <= distance + kLinearSlop )
			{
				manifold.points[0].point = b2MulAdd( cp, 0.5f * ( radiusB - radiusA - sp ), normalB );
				manifold.points[0].separation = sp - radius;
				manifold.points[0].id = B2G_MAKE_ID( 0, 0 );
				manifold.points[1].point = b2MulAdd( cq, 0.5f * ( radiusB - radiusA - sq ), normalB );
				manifold.points[1].separation = sq - radius;
				manifold.points[1].id = B2G_MAKE_ID( 1, 0 );
				manifold.pointCount = 2;
			}
		}
	}

	if ( manifold.pointCount == 0 )
	{
		b2Vec2 normal = b2Sub( closest2, closest1 );
		if ( b2Dot( normal, normal ) > epsSqr )
		{
			normal = b2Normalize( normal );
		}
		else
		{
			normal = b2LeftPerp( u1 );
		}

		b2Vec2 c1 = b2MulAdd( closest1, radiusA, normal );
		b2Vec2 c2 = b2MulAdd( closest2, -radiusB, normal );
		int i1 = f1 == 0.0f ? 0 : 1;
		int i2 = f2 == 0.0f ? 0 : 1;

		manifold.normal = normal;
		manifold.points[0].point = b2Lerp( c1, c2, 0.5f );
		manifold.points[0].separation = sqrtf( distanceSquared ) - radius;
		manifold.points[0].id = B2G_MAKE_ID( i1, i2 );
		manifold.pointCount = 1;
	}

	for ( int i = 0; i < manifold.pointCount; ++i )
	{
		manifold.points[i].point = b2Add( manifold.points[i].point, origin );
	}
	return manifold;
}

// A capsule as a two-vertex rounded polygon (reference src/manifold.c:16-34).
B2G_DEV void MakeCapsulePolygon( b2Polygon* shape, b2Vec2 p1, b2Vec2 p2, float radius )
{
	shape->vertices[0] = p1;
	shape->vertices[1] = p2;
	shape->centroid = b2Lerp( p1, p2, 0.5f );
	b2Vec2 d = b2Sub( p2, p1 );
	b2Vec2 axis = b2Normalize( d );
	b2Vec2 normal = b2RightPerp( axis );
	shape->normals[0] = normal;
	shape->normals[1] = b2Neg( normal );
	shape->count = 2;
	shape->radius = radius;
}

// Max separation of poly2 from the edge normals of poly1.
B2G_DEV float FindMaxSeparation( int* edgeIndex, const b2Polygon& poly1, const b2Polygon& poly2 )
{
	int count1 = poly1.count;
	int count2 = poly2.count;
	int bestIndex = 0;
	float maxSeparation = -FLT_MAX;
	for ( int i = 0; i < count1; ++i )
	{
		b2Vec2 n = poly1.normals[i];
		b2Vec2 v1 = poly1.vertices[i];
		float si = FLT_MAX;
		for ( int j = 0; j < count2; ++j )
		{
			float sij = b2Dot( n, b2Sub( poly2.vertices[j], v1 ) );
			if ( sij < si )
			{
				si = sij;
			}
		}
		if ( si > maxSeparation )
		{
			maxSeparation = si;
			bestIndex = i;
		}
	}
	*edgeIndex = bestIndex;
	return maxSeparation;
}

B2G_DEV b2LocalManifold ClipPolygons( const b2Polygon& polyA, const b2Polygon& polyB, int edgeA, int edgeB, bool flip )
{
	b2LocalManifold manifold = EmptyManifold();

	const b2Polygon* poly1;
	const b2Polygon* poly2;
	int i11, i12, i21, i22;

	if ( flip )
	{
		poly1 = &polyB;
		poly2 = &polyA;
		i11 = edgeB;
		i12 = edgeB + 1 < polyB.count ? edgeB + 1 : 0;
		i21 = edgeA;
		i22 = edgeA + 1 < polyA.count ? edgeA + 1 : 0;
	}
	else
	{
		poly1 = &polyA;
		poly2 = &polyB;
		i11 = edgeA;
		i12 = edgeA + 1 < polyA.count ? edgeA + 1 : 0;
		i21 = edgeB;
		i22 = edgeB + 1 < polyB.count ? edgeB + 1 : 0;
	}

	b2Vec2 normal = poly1->normals[i11];
	b2Vec2 v11 = poly1->vertices[i11];
	b2Vec2 v12 = poly1->vertices[i12];
	b2Vec2 v21 = poly2->vertices[i21];
	b2Vec2 v22 = poly2->vertices[i22];

	b2Vec2 tangent = b2CrossSV( 1.0f, normal );

	float lower1 = 0.0f;
	float upper1 = b2Dot( b2Sub( v12, v11 ), tangent );
	float upper2 = b2Dot( b2Sub( v21, v11 ), tangent );
	float lower2 = b2Dot( b2Sub( v22, v11 ), tangent );

	if ( upper2 < lower1 || upper1 < lower2 )
	{
		return manifold;
	}

	b2Vec2 vLower;
	if ( lower2 < lower1 && upper2 - lower2 > FLT_EPSILON )
	{
		vLower = b2Lerp( v22, v21, ( lower1 - lower2 ) / ( upper2 - lower2 ) );
	}
	else
	{
		vLower = v22;
	}

	b2Vec2 vUpper;
	if ( upper2 > upper1 && upper2 - lower2 > FLT_EPSILON )
	{
		vUpper = b2Lerp( v22, v21, ( upper1 - lower2 ) / ( upper2 - lower2 ) );
	}
	else
	{
		vUpper = v21;
	}

	float separationLower = b2Dot( b2Sub( vLower, v11 ), normal );
	float separationUpper = b2Dot( b2Sub( vUpper, v11 ), normal );

	float r1 = poly1->radius;
	float r2 = poly2->radius;
	vLower = b2MulAdd( vLower, 0.5f * ( r1 - r2 - separationLower ), normal );
	vUpper = b2MulAdd( vUpper, 0.5f * ( r1 - r2 - separationUpper ), normal );
	float radius = r1 + r2;

	if ( flip == false )
	{
		manifold.normal = normal;
		manifold.points[0].point = vLower;
		manifold.points[0].separation = separationLower - radius;
		manifold.points[0].id = B2G_MAKE_ID( i11, i22 );
		manifold.points[1].point = vUpper;
		manifold.points[1].separation = separationUpper - radius;
		manifold.points[1].id = B2G_MAKE_ID( i12, i21 );
		manifold.pointCount = 2;
	}
	else
	{
		manifold.normal = b2Neg( normal );
		manifold.points[0].point = vUpper;
		manifold.points[0].separation = separationUpper - radius;
		manifold.points[0].id = B2G_MAKE_ID( i21, i12 );
		manifold.points[1].point = vLower;
		manifold.points[1].separation = separationLower - radius;
		manifold.points[1].id = B2G_MAKE_ID( i22, i11 );
		manifold.pointCount = 2;
	}

	return manifold;
}

B2G_DEV b2LocalManifold CollidePolygons( const b2Polygon& polygonA, const b2Polygon& polygonB, b2Transform xf )
{
	b2Vec2 origin = polygonA.vertices[0];
	const float linearSlop = kLinearSlop;
	const float speculativeDistance = kSpeculativeDistance;

	b2Transform xfs = { b2Sub( xf.p, origin ), xf.q };

	b2Polygon localPolyA;
	localPolyA.count = polygonA.count;
	localPolyA.radius = polygonA.radius;
	localPolyA.vertices[0] = { 0.0f, 0.0f };
	localPolyA.normals[0] = polygonA.normals[0];
	for ( int i = 1; i < localPolyA.count; ++i )
	{
		localPolyA.vertices[i] = b2Sub( polygonA.vertices[i], origin );
		localPolyA.normals[i] = polygonA.normals[i];
	}

	b2Polygon localPolyB;
	localPolyB.count = polygonB.count;
	localPolyB.radius = polygonB.radius;
	for ( int i = 0; i < localPolyB.count; ++i )
	{
		localPolyB.vertices[i] = b2TransformPoint( xfs, polygonB.vertices[i] );
		localPolyB.normals[i] = b2RotateVector( xfs.q, polygonB.normals[i] );
	}

	int edgeA = 0;
	float separationA = FindMaxSeparation( &edgeA, localPolyA, localPolyB );
	int edgeB = 0;
	float separationB = FindMaxSeparation( &edgeB, localPolyB, localPolyA );

	float radius = localPolyA.radius + localPolyB.radius;

	if ( separationA > speculativeDistance + radius || separationB > speculativeDistance + radius )
	{
		return EmptyManifold();
	}

	bool flip;
	if ( separationA >= separationB )
	{
		flip = false;
		b2Vec2 searchDirection = localPolyA.normals[edgeA];
		edgeB = 0;
		float minDot = FLT_MAX;
		for ( int i = 0; i < localPolyB.count; ++i )
		{
			float dot = b2Dot( searchDirection, localPolyB.normals[i] );
			if ( dot < minDot )
			{
				minDot = dot;
				edgeB = i;
			}
		}
	}
	else
	{
		flip = true;
		b2Vec2 searchDirection = localPolyB.normals[edgeB];
		edgeA = 0;
		float minDot = FLT_MAX;
		for ( int i = 0; i < localPolyA.count; ++i )
		{
			float dot = b2Dot( searchDirection, localPolyA.normals[i] );
			if ( dot < minDot )
			{
				minDot = dot;
				edgeA = i;
			}
		}
	}

	b2LocalManifold manifold = EmptyManifold();

	if ( separationA > 0.1f * linearSlop || separationB > 0.1f * linearSlop )
	{
		// Disjoint: closest features of the reference and incident edges.
		int i11 = edgeA;
		int i12 = edgeA + 1 < localPolyA.count ? edgeA + 1 : 0;
		int i21 = edgeB;
		int i22 = edgeB + 1 < localPolyB.count ? edgeB + 1 : 0;

		b2Vec2 v11 = localPolyA.vertices[i11];
		b2Vec2 v12 = localPolyA.vertices[i12];
		b2Vec2 v21 = localPolyB.vertices[i21];
		b2Vec2 v22 = localPolyB.vertices[i22];

		b2SegmentDistanceResult result = SegmentDistance( v11, v12, v21, v22 );
		float distance = sqrtf( result.distanceSquared );
		float separation = distance - radius;

		if ( distance - radius > speculativeDistance )
		{
			return manifold;
		}

		manifold = ClipPolygons( localPolyA, localPolyB, edgeA, edgeB, flip );

		float minSeparation = FLT_MAX;
		for ( int i = 0; i < manifold.pointCount; ++i )
		{
			minSeparation = b2MinFloat( minSeparation, manifold.points[i].separation );
		}

		if ( separation + 0.1f * linearSlop < minSeparation )
		{
			// A vertex-vertex pair is much closer than the clipped edges.
			int iA = -1, iB = -1;
			b2Vec2 vA = {}, vB = {};
			if ( result.fraction1 == 0.0f && result.fraction2 == 0.0f )
			{
				iA = i11;
				iB = i21;
				vA = v11;
				vB = v21;
			}
			else if ( result.fraction1 == 0.0f && result.fraction2 == 1.0f )
			{
				iA = i11;
				iB = i22;
				vA = v11;
				vB = v22;
			}
			else if ( result.fraction1 == 1.0f && result.fraction2 == 0.0f )
			{
				iA = i12;
				iB = i21;
				vA = v12;
				vB = v21;
			}
			else if ( result.fraction1 == 1.0f && result.fraction2 == 1.0f )
			{
				iA = i12;
				iB = i22;
				vA = v12;
				vB = v22;
			}

			if ( iA >= 0 )
			{
				b2Vec2 normal = b2Sub( vB, vA );
				float invDistance = 1.0f / distance;
				normal.x *= invDistance;
				normal.y *= invDistance;
				b2Vec2 c1 = b2MulAdd( vA, localPolyA.radius, normal );
				b2Vec2 c2 = b2MulAdd( vB, -localPolyB.radius, normal );
				manifold.normal = normal;
				manifold.points[0].point = b2Lerp( c1, c2, 0.5f );
				manifold.points[0].separation = distance - radius;
				manifold.points[0].id = B2G_MAKE_ID( iA, iB );
				manifold.pointCount = 1;
			}
		}
	}
	else
	{
		manifold = ClipPolygons( localPolyA, localPolyB, edgeA, edgeB, flip );
	}

	for ( int i = 0; i < manifold.pointCount; ++i )
	{
		manifold.points[i].point = b2Add( manifold.points[i].point, origin );
	}
	return manifold;
}

B2G_DEV b2LocalManifold CollidePolygonAndCapsule( const b2Polygon& polygonA, const b2Capsule& capsuleB, b2Transform xf )
{
	b2Polygon polyB;
	MakeCapsulePolygon( &polyB, capsuleB.center1, capsuleB.center2, capsuleB.radius );
	return CollidePolygons( polygonA, polyB, xf );
}

B2G_DEV b2LocalManifold CollideSegmentAndCircle( const b2Segment& segmentA, const b2Circle& circleB, b2Transform xf )
{
	b2Capsule capsuleA = { segmentA.point1, segmentA.point2, 0.0f };
	return CollideCapsuleAndCircle( capsuleA, circleB, xf );
}

B2G_DEV b2LocalManifold CollideSegmentAndCapsule( const b2Segment& segmentA, const b2Capsule& capsuleB, b2Transform xf )
{
	b2Capsule capsuleA = { segmentA.point1, segmentA.point2, 0.0f };
	return CollideCapsules( capsuleA, capsuleB, xf );
}

B2G_DEV b2LocalManifold CollideSegmentAndPolygon( const b2Segment& segmentA, const b2Polygon& polygonB, b2Transform xf )
{
	b2Polygon polygonA;
	MakeCapsulePolygon( &polygonA, segmentA.point1, segmentA.point2, 0.0f );
	return CollidePolygons( polygonA, polygonB, xf );
}

// Manifold dispatch for a primary-ordered shape pair (reference src/contact.c:188-206).
// Chain segments are not handled here yet; the host refuses to create such contacts.
B2G_DEV b2LocalManifold ComputeManifold( const ShapeDev& shapeA, const ShapeDev& shapeB, b2Transform xf )
{
	switch ( shapeA.type )
	{
		case b2_circleShape:
			return CollideCircles( shapeA.circle, shapeB.circle, xf );
		case b2_capsuleShape:
			if ( shapeB.type == b2_circleShape )
			{
				return CollideCapsuleAndCircle( shapeA.capsule, shapeB.circle, xf );
			}
			return CollideCapsules( shapeA.capsule, shapeB.capsule, xf );
		case b2_polygonShape:
			if ( shapeB.type == b2_circleShape )
			{
				return CollidePolygonAndCircle( shapeA.polygon, shapeB.circle, xf );
			}
			if ( shapeB.type == b2_capsuleShape )
			{
				return CollidePolygonAndCapsule( shapeA.polygon, shapeB.capsule, xf );
			}
			return CollidePolygons( shapeA.polygon, shapeB.polygon, xf );
		case b2_segmentShape:
			if ( shapeB.type == b2_circleShape )
			{
				return CollideSegmentAndCircle( shapeA.segment, shapeB.circle, xf );
			}
			if ( shapeB.type == b2_capsuleShape )
			{
				return CollideSegmentAndCapsule( shapeA.segment, shapeB.capsule, xf );
			}
			return CollideSegmentAndPolygon( shapeA.segment, shapeB.polygon, xf );
		default:
			return EmptyManifold();
	}
}

} // namespace b2g
