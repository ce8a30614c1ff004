// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the GJK and time-of-impact arithmetic below
// restates its src/distance.c in the same operation order)
// SPDX-License-Identifier: MIT
// GJK distance and time of impact, host and device.
//
// The continuous-collision kernel runs these per fast body; the library also exports the host
// forms (b2ShapeDistance, b2TimeOfImpact). The arithmetic follows the reference's
// src/distance.c operation for operation (GJK with a simplex cache, lines 138-611; the
// separating-axis conservative advancement, lines 936-1441) so the results are bit-identical.
#pragma once

#include "narrowphase.h"

#include <cfloat>

namespace b2g
{

B2G_DEV b2ShapeProxy MakeProxy( const b2Vec2* points, int count, float radius )
{
	count = count < B2_MAX_POLYGON_VERTICES ? count : B2_MAX_POLYGON_VERTICES;
	b2ShapeProxy proxy;
	for ( int i = 0; i < count; ++i )
	{
		proxy.points[i] = points[i];
	}
	proxy.count = count;
	proxy.radius = radius;
	return proxy;
}

// Index of the proxy point farthest along direction (first maximum wins).
B2G_DEV int FindSupport( const b2ShapeProxy& proxy, b2Vec2 direction )
{
	int best = 0;
	float bestValue = b2Dot( proxy.points[0], direction );
	for ( int i = 1; i < proxy.count; ++i )
	{
		float value = b2Dot( proxy.points[i], direction );
		if ( value > bestValue )
		{
			best = i;
			bestValue = value;
		}
	}
	return best;
}

// Simplex vertices addressed by position (v1, v2, v3 of b2Simplex).
B2G_DEV b2SimplexVertex& SimplexAt( b2Simplex& s, int i )
{
	return i == 0 ? s.v1 : ( i == 1 ? s.v2 : s.v3 );
}

B2G_DEV b2Simplex SimplexFromCache( const b2SimplexCache& cache, const b2ShapeProxy& proxyA, const b2ShapeProxy& proxyB )
{
	b2Simplex s;
	s.count = cache.count;
	for ( int i = 0; i < s.count; ++i )
	{
		b2SimplexVertex& v = SimplexAt( s, i );
		v.indexA = cache.indexA[i];
		v.indexB = cache.indexB[i];
		v.wA = proxyA.points[v.indexA];
		v.wB = proxyB.points[v.indexB];
		v.w = b2Sub( v.wA, v.wB );
		v.a = -1.0f;
	}

	if ( s.count == 0 )
	{
		b2SimplexVertex& v = s.v1;
		v.indexA = 0;
		v.indexB = 0;
		v.wA = proxyA.points[0];
		v.wB = proxyB.points[0];
		v.w = b2Sub( v.wA, v.wB );
		v.a = 1.0f;
		s.count = 1;
	}
	return s;
}

B2G_DEV b2SimplexCache CacheFromSimplex( b2Simplex& s )
{
	b2SimplexCache cache;
	cache.count = (uint16_t)s.count;
	for ( int i = 0; i < 3; ++i )
	{
		cache.indexA[i] = 0;
		cache.indexB[i] = 0;
	}
	for ( int i = 0; i < s.count; ++i )
	{
		cache.indexA[i] = (uint8_t)SimplexAt( s, i ).indexA;
		cache.indexB[i] = (uint8_t)SimplexAt( s, i ).indexB;
	}
	return cache;
}

B2G_DEV void WitnessPoints( const b2Simplex& s, b2Vec2* a, b2Vec2* b )
{
	if ( s.count == 1 )
	{
		*a = s.v1.wA;
		*b = s.v1.wB;
	}
	else if ( s.count == 2 )
	{
		*a = { s.v1.a * s.v1.wA.x + s.v2.a * s.v2.wA.x, s.v1.a * s.v1.wA.y + s.v2.a * s.v2.wA.y };
		*b = { s.v1.a * s.v1.wB.x + s.v2.a * s.v2.wB.x, s.v1.a * s.v1.wB.y + s.v2.a * s.v2.wB.y };
	}
	else if ( s.count == 3 )
	{
		*a = { s.v1.a * s.v1.wA.x + s.v2.a * s.v2.wA.x + s.v3.a * s.v3.wA.x,
			   s.v1.a * s.v1.wA.y + s.v2.a * s.v2.wA.y + s.v3.a * s.v3.wA.y };
		*b = *a;
	}
	else
	{
		*a = { 0.0f, 0.0f };
		*b = { 0.0f, 0.0f };
	}
}

// Closest point of a segment simplex to the origin; returns the search direction.
B2G_DEV b2Vec2 SolveSimplex2( b2Simplex& s )
{
	b2Vec2 w1 = s.v1.w;
	b2Vec2 w2 = s.v2.w;
	b2Vec2 e12 = b2Sub( w2, w1 );

	float d12_2 = -b2Dot( w1, e12 );
	if ( d12_2 <= 0.0f )
	{
		s.v1.a = 1.0f;
		s.count = 1;
		return b2Neg( w1 );
	}

	float d12_1 = b2Dot( w2, e12 );
	if ( d12_1 <= 0.0f )
	{
		s.v2.a = 1.0f;
		s.count = 1;
		s.v1 = s.v2;
		return b2Neg( w2 );
	}

	float inv = 1.0f / ( d12_1 + d12_2 );
	s.v1.a = d12_1 * inv;
	s.v2.a = d12_2 * inv;
	s.count = 2;
	return b2CrossSV( b2Cross( b2Add( w1, w2 ), e12 ), e12 );
}

// Closest feature of a triangle simplex to the origin (Voronoi regions in the reference's order).
B2G_DEV b2Vec2 SolveSimplex3( b2Simplex& s )
{
	b2Vec2 w1 = s.v1.w;
	b2Vec2 w2 = s.v2.w;
	b2Vec2 w3 = s.v3.w;

	b2Vec2 e12 = b2Sub( w2, w1 );
	float d12_1 = b2Dot( w2, e12 );
	float d12_2 = -b2Dot( w1, e12 );

	b2Vec2 e13 = b2Sub( w3, w1 );
	float d13_1 = b2Dot( w3, e13 );
	float d13_2 = -b2Dot( w1, e13 );

	b2Vec2 e23 = b2Sub( w3, w2 );
	float d23_1 = b2Dot( w3, e23 );
	float d23_2 = -b2Dot( w2, e23 );

	float n123 = b2Cross( e12, e13 );
	float d123_1 = n123 * b2Cross( w2, w3 );
	float d123_2 = n123 * b2Cross( w3, w1 );
	float d123_3 = n123 * b2Cross( w1, w2 );

	if ( d12_2 <= 0.0f && d13_2 <= 0.0f )
	{
		s.v1.a = 1.0f;
		s.count = 1;
		return b2Neg( w1 );
	}

	if ( d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f )
	{
		float inv = 1.0f / ( d12_1 + d12_2 );
		s.v1.a = d12_1 * inv;
		s.v2.a = d12_2 * inv;
		s.count = 2;
		return b2CrossSV( b2Cross( b2Add( w1, w2 ), e12 ), e12 );
	}

	if ( d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f )
	{
		float inv = 1.0f / ( d13_1 + d13_2 );
		s.v1.a = d13_1 * inv;
		s.v3.a = d13_2 * inv;
		s.count = 2;
		s.v2 = s.v3;
		return b2CrossSV( b2Cross( b2Add( w1, w3 ), e13 ), e13 );
	}

	if ( d12_1 <= 0.0f && d23_2 <= 0.0f )
	{
		s.v2.a = 1.0f;
		s.count = 1;
		s.v1 = s.v2;
		return b2Neg( w2 );
	}

	if ( d13_1 <= 0.0f && d23_1 <= 0.0f )
	{
		s.v3.a = 1.0f;
		s.count = 1;
		s.v1 = s.v3;
		return b2Neg( w3 );
	}

	if ( d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f )
	{
		float inv = 1.0f / ( d23_1 + d23_2 );
		s.v2.a = d23_1 * inv;
		s.v3.a = d23_2 * inv;
		s.count = 2;
		s.v1 = s.v3;
		return b2CrossSV( b2Cross( b2Add( w2, w3 ), e23 ), e23 );
	}

	float inv = 1.0f / ( d123_1 + d123_2 + d123_3 );
	s.v1.a = d123_1 * inv;
	s.v2.a = d123_2 * inv;
	s.v3.a = d123_3 * inv;
	s.count = 3;
	return { 0.0f, 0.0f };
}

// GJK distance between two proxies, B posed in A's frame. The cache seeds and receives the simplex.
B2G_DEV b2DistanceOutput ShapeDistance( const b2DistanceInput& input, b2SimplexCache* cache )
{
	b2DistanceOutput output;
	output.pointA = { 0.0f, 0.0f };
	output.pointB = { 0.0f, 0.0f };
	output.normal = { 0.0f, 0.0f };
	output.distance = 0.0f;
	output.iterations = 0;
	output.simplexCount = 0;

	const b2ShapeProxy& proxyA = input.proxyA;
	b2ShapeProxy proxyB;
	proxyB.count = input.proxyB.count;
	proxyB.radius = input.proxyB.radius;
	for ( int i = 0; i < proxyB.count; ++i )
	{
		proxyB.points[i] = b2TransformPoint( input.transform, input.proxyB.points[i] );
	}

	b2Simplex simplex = SimplexFromCache( *cache, proxyA, proxyB );
	b2Vec2 nonUnitNormal = { 0.0f, 0.0f };
	int saveA[3], saveB[3];

	const int maxIterations = 20;
	int iteration = 0;
	while ( iteration < maxIterations )
	{
		int saveCount = simplex.count;
		for ( int i = 0; i < saveCount; ++i )
		{
			saveA[i] = SimplexAt( simplex, i ).indexA;
			saveB[i] = SimplexAt( simplex, i ).indexB;
		}

		b2Vec2 d = { 0.0f, 0.0f };
		if ( simplex.count == 1 )
		{
			d = b2Neg( simplex.v1.w );
		}
		else if ( simplex.count == 2 )
		{
			d = SolveSimplex2( simplex );
		}
		else if ( simplex.count == 3 )
		{
			d = SolveSimplex3( simplex );
		}

		if ( simplex.count == 3 || b2Dot( d, d ) < 1000.0f * FLT_MIN )
		{
			// overlap (origin inside the triangle, or a degenerate search direction)
			*cache = CacheFromSimplex( simplex );
			b2Vec2 a, b;
			WitnessPoints( simplex, &a, &b );
			output.distance = 0.0f;
			output.pointA = a;
			output.pointB = b;
			return output;
		}

		nonUnitNormal = d;

		b2SimplexVertex& vertex = SimplexAt( simplex, simplex.count );
		vertex.indexA = FindSupport( proxyA, d );
		vertex.wA = proxyA.points[vertex.indexA];
		vertex.indexB = FindSupport( proxyB, b2Neg( d ) );
		vertex.wB = proxyB.points[vertex.indexB];
		vertex.w = b2Sub( vertex.wA, vertex.wB );

		++iteration;

		bool duplicate = false;
		for ( int i = 0; i < saveCount; ++i )
		{
			if ( vertex.indexA == saveA[i] && vertex.indexB == saveB[i] )
			{
				duplicate = true;
				break;
			}
		}
		if ( duplicate )
		{
			break;
		}

		simplex.count += 1;
	}

	b2Vec2 normal = b2Normalize( nonUnitNormal );
	b2Vec2 a, b;
	WitnessPoints( simplex, &a, &b );
	output.normal = normal;
	output.distance = b2Distance( a, b );
	output.pointA = a;
	output.pointB = b;
	output.iterations = iteration;
	output.simplexCount = 0;

	*cache = CacheFromSimplex( simplex );

	if ( input.useRadii )
	{
		float radiusA = input.proxyA.radius;
		float radiusB = input.proxyB.radius;
		output.distance = b2MaxFloat( 0.0f, output.distance - radiusA - radiusB );
		output.pointA = b2MulAdd( output.pointA, radiusA, normal );
		output.pointB = b2MulSub( output.pointB, radiusB, normal );
	}
	return output;
}

// Transform of a sweep at a fraction of the step: lerped center and normalized lerped rotation.
B2G_DEV b2Transform SweepTransform( const b2Sweep& sweep, float time )
{
	b2Transform xf;
	xf.p = b2Add( b2MulSV( 1.0f - time, sweep.c1 ), b2MulSV( time, sweep.c2 ) );
	b2Rot q = { ( 1.0f - time ) * sweep.q1.c + time * sweep.q2.c, ( 1.0f - time ) * sweep.q1.s + time * sweep.q2.s };
	xf.q = b2NormalizeRot( q );
	xf.p = b2Sub( xf.p, b2RotateVector( xf.q, sweep.localCenter ) );
	return xf;
}

enum SeparationType
{
	kSeparationPoints,
	kSeparationFaceA,
	kSeparationFaceB,
};

struct SeparationFunction
{
	const b2ShapeProxy* proxyA;
	const b2ShapeProxy* proxyB;
	b2Sweep sweepA, sweepB;
	b2Vec2 localPoint;
	b2Vec2 axis;
	int type;
};

B2G_DEV SeparationFunction MakeSeparationFunction( const b2SimplexCache& cache, const b2ShapeProxy* proxyA, const b2Sweep& sweepA,
												   const b2ShapeProxy* proxyB, const b2Sweep& sweepB, float t1 )
{
	SeparationFunction f;
	f.proxyA = proxyA;
	f.proxyB = proxyB;
	f.sweepA = sweepA;
	f.sweepB = sweepB;

	b2Transform xfA = SweepTransform( sweepA, t1 );
	b2Transform xfB = SweepTransform( sweepB, t1 );

	if ( cache.count == 1 )
	{
		f.type = kSeparationPoints;
		b2Vec2 pointA = b2TransformPoint( xfA, proxyA->points[cache.indexA[0]] );
		b2Vec2 pointB = b2TransformPoint( xfB, proxyB->points[cache.indexB[0]] );
		f.axis = b2Normalize( b2Sub( pointB, pointA ) );
		f.localPoint = { 0.0f, 0.0f };
		return f;
	}

	if ( cache.indexA[0] == cache.indexA[1] )
	{
		// two points on B, one on A
		f.type = kSeparationFaceB;
		b2Vec2 b1 = proxyB->points[cache.indexB[0]];
		b2Vec2 b2 = proxyB->points[cache.indexB[1]];
		f.axis = b2Normalize( b2CrossVS( b2Sub( b2, b1 ), 1.0f ) );
		b2Vec2 normal = b2RotateVector( xfB.q, f.axis );
		f.localPoint = { 0.5f * ( b1.x + b2.x ), 0.5f * ( b1.y + b2.y ) };
		b2Vec2 pointB = b2TransformPoint( xfB, f.localPoint );
		b2Vec2 pointA = b2TransformPoint( xfA, proxyA->points[cache.indexA[0]] );
		float s = b2Dot( b2Sub( pointA, pointB ), normal );
		if ( s < 0.0f )
		{
			f.axis = b2Neg( f.axis );
		}
		return f;
	}

	// two points on A, one or two on B
	f.type = kSeparationFaceA;
	b2Vec2 a1 = proxyA->points[cache.indexA[0]];
	b2Vec2 a2 = proxyA->points[cache.indexA[1]];
	f.axis = b2Normalize( b2CrossVS( b2Sub( a2, a1 ), 1.0f ) );
	b2Vec2 normal = b2RotateVector( xfA.q, f.axis );
	f.localPoint = { 0.5f * ( a1.x + a2.x ), 0.5f * ( a1.y + a2.y ) };
	b2Vec2 pointA = b2TransformPoint( xfA, f.localPoint );
	b2Vec2 pointB = b2TransformPoint( xfB, proxyB->points[cache.indexB[0]] );
	float s = b2Dot( b2Sub( pointB, pointA ), normal );
	if ( s < 0.0f )
	{
		f.axis = b2Neg( f.axis );
	}
	return f;
}

// Deepest points along the separating axis at time t and their separation.
B2G_DEV float FindMinSeparation( const SeparationFunction& f, int* indexA, int* indexB, float t )
{
	b2Transform xfA = SweepTransform( f.sweepA, t );
	b2Transform xfB = SweepTransform( f.sweepB, t );

	if ( f.type == kSeparationPoints )
	{
		b2Vec2 axisA = b2InvRotateVector( xfA.q, f.axis );
		b2Vec2 axisB = b2InvRotateVector( xfB.q, b2Neg( f.axis ) );
		*indexA = FindSupport( *f.proxyA, axisA );
		*indexB = FindSupport( *f.proxyB, axisB );
		b2Vec2 pointA = b2TransformPoint( xfA, f.proxyA->points[*indexA] );
		b2Vec2 pointB = b2TransformPoint( xfB, f.proxyB->points[*indexB] );
		return b2Dot( b2Sub( pointB, pointA ), f.axis );
	}

	if ( f.type == kSeparationFaceA )
	{
		b2Vec2 normal = b2RotateVector( xfA.q, f.axis );
		b2Vec2 pointA = b2TransformPoint( xfA, f.localPoint );
		b2Vec2 axisB = b2InvRotateVector( xfB.q, b2Neg( normal ) );
		*indexA = -1;
		*indexB = FindSupport( *f.proxyB, axisB );
		b2Vec2 pointB = b2TransformPoint( xfB, f.proxyB->points[*indexB] );
		return b2Dot( b2Sub( pointB, pointA ), normal );
	}

	b2Vec2 normal = b2RotateVector( xfB.q, f.axis );
	b2Vec2 pointB = b2TransformPoint( xfB, f.localPoint );
	b2Vec2 axisA = b2InvRotateVector( xfA.q, b2Neg( normal ) );
	*indexB = -1;
	*indexA = FindSupport( *f.proxyA, axisA );
	b2Vec2 pointA = b2TransformPoint( xfA, f.proxyA->points[*indexA] );
	return b2Dot( b2Sub( pointA, pointB ), normal );
}

B2G_DEV float EvaluateSeparation( const SeparationFunction& f, int indexA, int indexB, float t )
{
	b2Transform xfA = SweepTransform( f.sweepA, t );
	b2Transform xfB = SweepTransform( f.sweepB, t );

	if ( f.type == kSeparationPoints )
	{
		b2Vec2 pointA = b2TransformPoint( xfA, f.proxyA->points[indexA] );
		b2Vec2 pointB = b2TransformPoint( xfB, f.proxyB->points[indexB] );
		return b2Dot( b2Sub( pointB, pointA ), f.axis );
	}

	if ( f.type == kSeparationFaceA )
	{
		b2Vec2 normal = b2RotateVector( xfA.q, f.axis );
		b2Vec2 pointA = b2TransformPoint( xfA, f.localPoint );
		b2Vec2 pointB = b2TransformPoint( xfB, f.proxyB->points[indexB] );
		return b2Dot( b2Sub( pointB, pointA ), normal );
	}

	b2Vec2 normal = b2RotateVector( xfB.q, f.axis );
	b2Vec2 pointB = b2TransformPoint( xfB, f.localPoint );
	b2Vec2 pointA = b2TransformPoint( xfA, f.proxyA->points[indexA] );
	return b2Dot( b2Sub( pointA, pointB ), normal );
}

// Conservative advancement along successive separating axes; fraction in [0, maxFraction].
B2G_DEV b2TOIOutput TimeOfImpact( const b2TOIInput& input )
{
	b2TOIOutput output;
	output.state = b2_toiStateUnknown;
	output.point = { 0.0f, 0.0f };
	output.normal = { 0.0f, 0.0f };
	output.fraction = input.maxFraction;

	const b2Sweep& sweepA = input.sweepA;
	const b2Sweep& sweepB = input.sweepB;
	const b2ShapeProxy* proxyA = &input.proxyA;
	const b2ShapeProxy* proxyB = &input.proxyB;

	float tMax = input.maxFraction;
	float totalRadius = proxyA->radius + proxyB->radius;
	float target = b2MaxFloat( kLinearSlop, totalRadius - kLinearSlop );
	float tolerance = 0.25f * kLinearSlop;

	float t1 = 0.0f;
	const int maxIterations = 20;
	int distanceIterations = 0;

	b2SimplexCache cache;
	cache.count = 0;
	for ( int i = 0; i < 3; ++i )
	{
		cache.indexA[i] = 0;
		cache.indexB[i] = 0;
	}
	b2DistanceInput distanceInput;
	distanceInput.proxyA = input.proxyA;
	distanceInput.proxyB = input.proxyB;
	distanceInput.useRadii = false;

	for ( ;; )
	{
		b2Transform xfA = SweepTransform( sweepA, t1 );
		b2Transform xfB = SweepTransform( sweepB, t1 );
		distanceInput.transform = b2InvMulTransforms( xfA, xfB );
		b2DistanceOutput distanceOutput = ShapeDistance( distanceInput, &cache );

		b2Vec2 worldNormal = b2RotateVector( xfA.q, distanceOutput.normal );
		b2Vec2 worldPointA = b2TransformPoint( xfA, distanceOutput.pointA );
		b2Vec2 worldPointB = b2TransformPoint( xfA, distanceOutput.pointB );

		distanceIterations += 1;

		if ( distanceOutput.distance <= 0.0f )
		{
			output.state = b2_toiStateOverlapped;
			output.fraction = 0.0f;
			break;
		}

		if ( distanceOutput.distance <= target + tolerance )
		{
			output.state = b2_toiStateHit;
			b2Vec2 pA = b2MulAdd( worldPointA, proxyA->radius, worldNormal );
			b2Vec2 pB = b2MulAdd( worldPointB, -proxyB->radius, worldNormal );
			output.point = b2Lerp( pA, pB, 0.5f );
			output.normal = worldNormal;
			output.fraction = t1;
			break;
		}

		SeparationFunction fcn = MakeSeparationFunction( cache, proxyA, sweepA, proxyB, sweepB, t1 );

		bool done = false;
		float t2 = tMax;
		int pushBackIterations = 0;
		for ( ;; )
		{
			int indexA, indexB;
			float s2 = FindMinSeparation( fcn, &indexA, &indexB, t2 );

			if ( s2 > target + tolerance )
			{
				output.state = b2_toiStateSeparated;
				output.fraction = tMax;
				done = true;
				break;
			}

			if ( s2 > target - tolerance )
			{
				t1 = t2;
				break;
			}

			float s1 = EvaluateSeparation( fcn, indexA, indexB, t1 );

			if ( s1 < target - tolerance )
			{
				output.state = b2_toiStateFailed;
				output.fraction = t1;
				done = true;
				break;
			}

			if ( s1 <= target + tolerance )
			{
				output.state = b2_toiStateHit;
				b2Vec2 pA = b2MulAdd( worldPointA, proxyA->radius, worldNormal );
				b2Vec2 pB = b2MulAdd( worldPointB, -proxyB->radius, worldNormal );
				output.point = b2Lerp( pA, pB, 0.5f );
				output.normal = worldNormal;
				output.fraction = t1;
				done = true;
				break;
			}

			// 1D root of separation(t) = target by alternating bisection and false position
			int rootIterationCount = 0;
			float a1 = t1, a2 = t2;
			for ( ;; )
			{
				float t;
				if ( rootIterationCount & 1 )
				{
					t = a1 + ( target - s1 ) * ( a2 - a1 ) / ( s2 - s1 );
				}
				else
				{
					t = 0.5f * ( a1 + a2 );
				}
				rootIterationCount += 1;

				float s = EvaluateSeparation( fcn, indexA, indexB, t );

				if ( b2AbsFloat( s - target ) < tolerance )
				{
					t2 = t;
					break;
				}

				if ( s > target )
				{
					a1 = t;
					s1 = s;
				}
				else
				{
					a2 = t;
					s2 = s;
				}

				if ( rootIterationCount == 50 )
				{
					break;
				}
			}

			pushBackIterations += 1;
			if ( pushBackIterations == B2_MAX_POLYGON_VERTICES )
			{
				break;
			}
		}

		if ( done )
		{
			break;
		}

		if ( distanceIterations == maxIterations )
		{
			output.state = b2_toiStateFailed;
			b2Vec2 pA = b2MulAdd( worldPointA, proxyA->radius, worldNormal );
			b2Vec2 pB = b2MulAdd( worldPointB, -proxyB->radius, worldNormal );
			output.point = b2Lerp( pA, pB, 0.5f );
			output.normal = worldNormal;
			output.fraction = t1;
			break;
		}
	}

	return output;
}

} // namespace b2g
