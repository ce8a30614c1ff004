// Polygon manifold with derivatives, for the adjoint of the narrowphase.
//
// The backward pass needs the Jacobian of a polygon-polygon manifold (normal, contact points,
// separations in A's frame) with respect to the relative transform of B in A's frame. This header
// evaluates the polygon manifold on forward-mode dual numbers with four derivative components
// (d/d p.x, d/d p.y, d/d q.c, d/d q.s). Every value part is computed with the same IEEE float
// operations in the same order as the float manifold in narrowphase.h (reference
// src/manifold.c:511-1028 and the segment distance of src/distance.c:33-106), so branches taken on
// value parts reproduce the recorded forward's feature choices exactly; the derivative parts are
// the derivatives of the branch that was taken.
#pragma once

#include "narrowphase.h"

namespace b2g
{

struct D4
{
	float v;
	float d[4];
};

#define B2G_DD __device__ __forceinline__

B2G_DD D4 DConst( float v )
{
	return D4{ v, { 0.0f, 0.0f, 0.0f, 0.0f } };
}

B2G_DD D4 operator+( const D4& a, const D4& b )
{
	D4 r;
	r.v = a.v + b.v;
	for ( int k = 0; k < 4; ++k )
	{
		r.d[k] = a.d[k] + b.d[k];
	}
	return r;
}

B2G_DD D4 operator-( const D4& a, const D4& b )
{
	D4 r;
	r.v = a.v - b.v;
	for ( int k = 0; k < 4; ++k )
	{
		r.d[k] = a.d[k] - b.d[k];
	}
	return r;
}

B2G_DD D4 operator-( const D4& a )
{
	D4 r;
	r.v = -a.v;
	for ( int k = 0; k < 4; ++k )
	{
		r.d[k] = -a.d[k];
	}
	return r;
}

B2G_DD D4 operator*( const D4& a, const D4& b )
{
	D4 r;
	r.v = a.v * b.v;
	for ( int k = 0; k < 4; ++k )
	{
		r.d[k] = a.d[k] * b.v + a.v * b.d[k];
	}
	return r;
}

B2G_DD D4 operator*( float s, const D4& b )
{
	D4 r;
	r.v = s * b.v;
	for ( int k = 0; k < 4; ++k )
	{
		r.d[k] = s * b.d[k];
	}
	return r;
}

B2G_DD D4 operator*( const D4& a, float s )
{
	D4 r;
	r.v = a.v * s;
	for ( int k = 0; k < 4; ++k )
	{
		r.d[k] = a.d[k] * s;
	}
	return r;
}

B2G_DD D4 operator/( const D4& a, const D4& b )
{
	D4 r;
	r.v = a.v / b.v;
	for ( int k = 0; k < 4; ++k )
	{
		r.d[k] = ( a.d[k] - r.v * b.d[k] ) / b.v;
	}
	return r;
}

B2G_DD D4 operator/( float s, const D4& b )
{
	return DConst( s ) / b;
}

B2G_DD D4 DSqrt( const D4& a )
{
	D4 r;
	r.v = sqrtf( a.v );
	for ( int k = 0; k < 4; ++k )
	{
		r.d[k] = r.v > 0.0f ? a.d[k] / ( 2.0f * r.v ) : 0.0f;
	}
	return r;
}

struct DV2
{
	D4 x, y;
};

B2G_DD DV2 DVConst( b2Vec2 v )
{
	return DV2{ DConst( v.x ), DConst( v.y ) };
}

B2G_DD DV2 DSub( const DV2& a, const DV2& b )
{
	return DV2{ a.x - b.x, a.y - b.y };
}

B2G_DD DV2 DAdd( const DV2& a, const DV2& b )
{
	return DV2{ a.x + b.x, a.y + b.y };
}

B2G_DD D4 DDot( const DV2& a, const DV2& b )
{
	return a.x * b.x + a.y * b.y;
}

// b2MulAdd( a, s, b ) = a + s * b
B2G_DD DV2 DMulAdd( const DV2& a, const D4& s, const DV2& b )
{
	return DV2{ a.x + s * b.x, a.y + s * b.y };
}

// b2Lerp( a, b, t ) = ( 1 - t ) * a + t * b
B2G_DD DV2 DLerp( const DV2& a, const DV2& b, const D4& t )
{
	D4 omt = DConst( 1.0f ) - t;
	return DV2{ omt * a.x + t * b.x, omt * a.y + t * b.y };
}

B2G_DD D4 DClamp01( const D4& a )
{
	// b2ClampFloat( a, 0, 1 ): a < 0 ? 0 : ( a > 1 ? 1 : a )
	if ( a.v < 0.0f )
	{
		return DConst( 0.0f );
	}
	if ( a.v > 1.0f )
	{
		return DConst( 1.0f );
	}
	return a;
}

struct DPolygon
{
	DV2 vertices[B2_MAX_POLYGON_VERTICES];
	DV2 normals[B2_MAX_POLYGON_VERTICES];
	int count;
	float radius;
};

struct DManifoldPoint
{
	DV2 point;
	D4 separation;
	uint16_t id;
};

struct DManifold
{
	DV2 normal;
	DManifoldPoint points[2];
	int pointCount;
};

B2G_DD DManifold DEmptyManifold()
{
	DManifold m;
	m.normal = DVConst( { 0.0f, 0.0f } );
	m.points[0].point = DVConst( { 0.0f, 0.0f } );
	m.points[0].separation = DConst( 0.0f );
	m.points[0].id = 0;
	m.points[1] = m.points[0];
	m.pointCount = 0;
	return m;
}

B2G_DD float DFindMaxSeparation( int* edgeIndex, const DPolygon& poly1, const DPolygon& poly2 )
{
	// only the value parts matter here: the result picks features and is compared with bounds
	int count1 = poly1.count;
	int count2 = poly2.count;
	int bestIndex = 0;
	float maxSeparation = -FLT_MAX;
	for ( int i = 0; i < count1; ++i )
	{
		b2Vec2 n = { poly1.normals[i].x.v, poly1.normals[i].y.v };
		b2Vec2 v1 = { poly1.vertices[i].x.v, poly1.vertices[i].y.v };
		float si = FLT_MAX;
		for ( int j = 0; j < count2; ++j )
		{
			b2Vec2 v2 = { poly2.vertices[j].x.v, poly2.vertices[j].y.v };
			float sij = b2Dot( n, b2Sub( v2, v1 ) );
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

struct DSegmentDistance
{
	DV2 closest1, closest2;
	D4 fraction1, fraction2;
	D4 distanceSquared;
};

B2G_DD DSegmentDistance DSegmentDistanceOf( const DV2& p1, const DV2& q1, const DV2& p2, const DV2& q2 )
{
	DSegmentDistance result;
	result.fraction1 = DConst( 0.0f );
	result.fraction2 = DConst( 0.0f );

	DV2 d1 = DSub( q1, p1 );
	DV2 d2 = DSub( q2, p2 );
	DV2 r = DSub( p1, p2 );
	D4 dd1 = DDot( d1, d1 );
	D4 dd2 = DDot( d2, d2 );
	D4 rd1 = DDot( r, d1 );
	D4 rd2 = DDot( r, d2 );

	const float epsSqr = FLT_EPSILON * FLT_EPSILON;

	if ( dd1.v < epsSqr || dd2.v < epsSqr )
	{
		if ( dd1.v >= epsSqr )
		{
			result.fraction1 = DClamp01( -rd1 / dd1 );
			result.fraction2 = DConst( 0.0f );
		}
		else if ( dd2.v >= epsSqr )
		{
			result.fraction1 = DConst( 0.0f );
			result.fraction2 = DClamp01( rd2 / dd2 );
		}
	}
	else
	{
		D4 d12 = DDot( d1, d2 );
		D4 denominator = dd1 * dd2 - d12 * d12;

		D4 f1 = DConst( 0.0f );
		if ( denominator.v != 0.0f )
		{
			f1 = DClamp01( ( d12 * rd2 - rd1 * dd2 ) / denominator );
		}

		D4 f2 = ( d12 * f1 + rd2 ) / dd2;

		if ( f2.v < 0.0f )
		{
			f2 = DConst( 0.0f );
			f1 = DClamp01( -rd1 / dd1 );
		}
		else if ( f2.v > 1.0f )
		{
			f2 = DConst( 1.0f );
			f1 = DClamp01( ( d12 - rd1 ) / dd1 );
		}

		result.fraction1 = f1;
		result.fraction2 = f2;
	}

	result.closest1 = DMulAdd( p1, result.fraction1, d1 );
	result.closest2 = DMulAdd( p2, result.fraction2, d2 );
	DV2 c = DSub( result.closest2, result.closest1 );
	result.distanceSquared = c.x * c.x + c.y * c.y;
	return result;
}

B2G_DD DManifold DClipPolygons( const DPolygon& polyA, const DPolygon& polyB, int edgeA, int edgeB, bool flip )
{
	DManifold manifold = DEmptyManifold();

	const DPolygon* poly1;
	const DPolygon* poly2;
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

	DV2 normal = poly1->normals[i11];
	DV2 v11 = poly1->vertices[i11];
	DV2 v12 = poly1->vertices[i12];
	DV2 v21 = poly2->vertices[i21];
	DV2 v22 = poly2->vertices[i22];

	// b2CrossSV( 1, normal ) = { -normal.y, normal.x }
	DV2 tangent = { -( 1.0f * normal.y ), 1.0f * normal.x };

	D4 lower1 = DConst( 0.0f );
	D4 upper1 = DDot( DSub( v12, v11 ), tangent );
	D4 upper2 = DDot( DSub( v21, v11 ), tangent );
	D4 lower2 = DDot( DSub( v22, v11 ), tangent );

	if ( upper2.v < lower1.v || upper1.v < lower2.v )
	{
		return manifold;
	}

	DV2 vLower;
	if ( lower2.v < lower1.v && ( upper2 - lower2 ).v > FLT_EPSILON )
	{
		vLower = DLerp( v22, v21, ( lower1 - lower2 ) / ( upper2 - lower2 ) );
	}
	else
	{
		vLower = v22;
	}

	DV2 vUpper;
	if ( upper2.v > upper1.v && ( upper2 - lower2 ).v > FLT_EPSILON )
	{
		vUpper = DLerp( v22, v21, ( upper1 - lower2 ) / ( upper2 - lower2 ) );
	}
	else
	{
		vUpper = v21;
	}

	D4 separationLower = DDot( DSub( vLower, v11 ), normal );
	D4 separationUpper = DDot( DSub( vUpper, v11 ), normal );

	float r1 = poly1->radius;
	float r2 = poly2->radius;
	vLower = DMulAdd( vLower, 0.5f * ( DConst( r1 - r2 ) - separationLower ), normal );
	vUpper = DMulAdd( vUpper, 0.5f * ( DConst( r1 - r2 ) - separationUpper ), normal );
	float radius = r1 + r2;

	if ( flip == false )
	{
		manifold.normal = normal;
		manifold.points[0].point = vLower;
		manifold.points[0].separation = separationLower - DConst( radius );
		manifold.points[0].id = B2G_MAKE_ID( i11, i22 );
		manifold.points[1].point = vUpper;
		manifold.points[1].separation = separationUpper - DConst( radius );
		manifold.points[1].id = B2G_MAKE_ID( i12, i21 );
		manifold.pointCount = 2;
	}
	else
	{
		manifold.normal = DV2{ -normal.x, -normal.y };
		manifold.points[0].point = vUpper;
		manifold.points[0].separation = separationUpper - DConst( radius );
		manifold.points[0].id = B2G_MAKE_ID( i21, i12 );
		manifold.points[1].point = vLower;
		manifold.points[1].separation = separationLower - DConst( radius );
		manifold.points[1].id = B2G_MAKE_ID( i22, i11 );
		manifold.pointCount = 2;
	}

	return manifold;
}

// Polygon manifold of B in A's frame, xf = (p, q) the pose of B relative to A with dual parts
// seeded by the caller. Follows CollidePolygons in narrowphase.h step for step.
B2G_DD DManifold DCollidePolygons( const b2Polygon& polygonA, const b2Polygon& polygonB, const DV2& xfp, const DV2& xfq )
{
	b2Vec2 origin = polygonA.vertices[0];
	const float linearSlop = kLinearSlop;
	const float speculativeDistance = kSpeculativeDistance;

	// xfs = { xf.p - origin, xf.q }
	DV2 xsp = { xfp.x - DConst( origin.x ), xfp.y - DConst( origin.y ) };

	DPolygon localPolyA;
	localPolyA.count = polygonA.count;
	localPolyA.radius = polygonA.radius;
	localPolyA.vertices[0] = DVConst( { 0.0f, 0.0f } );
	localPolyA.normals[0] = DVConst( polygonA.normals[0] );
	for ( int i = 1; i < localPolyA.count; ++i )
	{
		localPolyA.vertices[i] = DVConst( b2Sub( polygonA.vertices[i], origin ) );
		localPolyA.normals[i] = DVConst( polygonA.normals[i] );
	}

	DPolygon localPolyB;
	localPolyB.count = polygonB.count;
	localPolyB.radius = polygonB.radius;
	for ( int i = 0; i < localPolyB.count; ++i )
	{
		// b2TransformPoint: ( q.c * v.x - q.s * v.y ) + p.x, ( q.s * v.x + q.c * v.y ) + p.y
		b2Vec2 v = polygonB.vertices[i];
		localPolyB.vertices[i].x = ( xfq.x * v.x - xfq.y * v.y ) + xsp.x;
		localPolyB.vertices[i].y = ( xfq.y * v.x + xfq.x * v.y ) + xsp.y;
		// b2RotateVector: q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y
		b2Vec2 n = polygonB.normals[i];
		localPolyB.normals[i].x = xfq.x * n.x - xfq.y * n.y;
		localPolyB.normals[i].y = xfq.y * n.x + xfq.x * n.y;
	}

	int edgeA = 0;
	float separationA = DFindMaxSeparation( &edgeA, localPolyA, localPolyB );
	int edgeB = 0;
	float separationB = DFindMaxSeparation( &edgeB, localPolyB, localPolyA );

	float radius = localPolyA.radius + localPolyB.radius;

	if ( separationA > speculativeDistance + radius || separationB > speculativeDistance + radius )
	{
		return DEmptyManifold();
	}

	bool flip;
	if ( separationA >= separationB )
	{
		flip = false;
		b2Vec2 searchDirection = { localPolyA.normals[edgeA].x.v, localPolyA.normals[edgeA].y.v };
		edgeB = 0;
		float minDot = FLT_MAX;
		for ( int i = 0; i < localPolyB.count; ++i )
		{
			b2Vec2 n = { localPolyB.normals[i].x.v, localPolyB.normals[i].y.v };
			float dot = b2Dot( searchDirection, n );
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
		b2Vec2 searchDirection = { localPolyB.normals[edgeB].x.v, localPolyB.normals[edgeB].y.v };
		edgeA = 0;
		float minDot = FLT_MAX;
		for ( int i = 0; i < localPolyA.count; ++i )
		{
			b2Vec2 n = { localPolyA.normals[i].x.v, localPolyA.normals[i].y.v };
			float dot = b2Dot( searchDirection, n );
			if ( dot < minDot )
			{
				minDot = dot;
				edgeA = i;
			}
		}
	}

	DManifold manifold = DEmptyManifold();

	if ( separationA > 0.1f * linearSlop || separationB > 0.1f * linearSlop )
	{
		int i11 = edgeA;
		int i12 = edgeA + 1 < localPolyA.count ? edgeA + 1 : 0;
		int i21 = edgeB;
		int i22 = edgeB + 1 < localPolyB.count ? edgeB + 1 : 0;

		DV2 v11 = localPolyA.vertices[i11];
		DV2 v12 = localPolyA.vertices[i12];
		DV2 v21 = localPolyB.vertices[i21];
		DV2 v22 = localPolyB.vertices[i22];

		DSegmentDistance result = DSegmentDistanceOf( v11, v12, v21, v22 );
		D4 distance = DSqrt( result.distanceSquared );
		D4 separation = distance - DConst( radius );

		if ( ( distance - DConst( radius ) ).v > speculativeDistance )
		{
			return manifold;
		}

		manifold = DClipPolygons( localPolyA, localPolyB, edgeA, edgeB, flip );

		float minSeparation = FLT_MAX;
		for ( int i = 0; i < manifold.pointCount; ++i )
		{
			minSeparation = b2MinFloat( minSeparation, manifold.points[i].separation.v );
		}

		if ( separation.v + 0.1f * linearSlop < minSeparation )
		{
			int iA = -1, iB = -1;
			DV2 vA = DVConst( { 0.0f, 0.0f } ), vB = DVConst( { 0.0f, 0.0f } );
			float f1 = result.fraction1.v, f2 = result.fraction2.v;
			if ( f1 == 0.0f && f2 == 0.0f )
			{
				iA = i11;
				iB = i21;
				vA = v11;
				vB = v21;
			}
			else if ( f1 == 0.0f && f2 == 1.0f )
			{
				iA = i11;
				iB = i22;
				vA = v11;
				vB = v22;
			}
			else if ( f1 == 1.0f && f2 == 0.0f )
			{
				iA = i12;
				iB = i21;
				vA = v12;
				vB = v21;
			}
			else if ( f1 == 1.0f && f2 == 1.0f )
			{
				iA = i12;
				iB = i22;
				vA = v12;
				vB = v22;
			}

			if ( iA >= 0 )
			{
				DV2 normal = DSub( vB, vA );
				D4 invDistance = 1.0f / distance;
				normal.x = normal.x * invDistance;
				normal.y = normal.y * invDistance;
				DV2 c1 = DMulAdd( vA, DConst( localPolyA.radius ), normal );
				DV2 c2 = DMulAdd( vB, DConst( -localPolyB.radius ), normal );
				manifold.normal = normal;
				manifold.points[0].point = DLerp( c1, c2, DConst( 0.5f ) );
				manifold.points[0].separation = distance - DConst( radius );
				manifold.points[0].id = B2G_MAKE_ID( iA, iB );
				manifold.pointCount = 1;
			}
		}
	}
	else
	{
		manifold = DClipPolygons( localPolyA, localPolyB, edgeA, edgeB, flip );
	}

	for ( int i = 0; i < manifold.pointCount; ++i )
	{
		manifold.points[i].point = DAdd( manifold.points[i].point, DVConst( origin ) );
	}
	return manifold;
}

} // namespace b2g
