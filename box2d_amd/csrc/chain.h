// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the chain-segment manifold arithmetic below
// restates its src/manifold.c in the same operation order)
// SPDX-License-Identifier: MIT
// One-sided chain segments against round and convex shapes, and the manifold dispatch with the
// contact's GJK simplex cache.
//
// Chain segments collide only on their right side and use their ghost vertices to reject normals
// that point into neighboring segments (reference src/manifold.c:1042-1641, after
// https://box2d.org/posts/2020/06/ghost-collisions/). Bit-identical to the reference: same
// expressions, same branch order.
#pragma once

#include "toi.h"

namespace b2g
{

B2G_DEV b2LocalManifold CollideChainSegmentAndCircle( const b2ChainSegment& segmentA, const b2Circle& circleB, b2Transform xf )
{
	b2LocalManifold manifold = EmptyManifold();

	b2Vec2 pB = b2TransformPoint( xf, circleB.center );
	b2Vec2 p1 = segmentA.segment.point1;
	b2Vec2 p2 = segmentA.segment.point2;
	b2Vec2 e = b2Sub( p2, p1 );

	// one-sided: the normal points to the right of the segment
	float offset = b2Dot( b2RightPerp( e ), b2Sub( pB, p1 ) );
	if ( offset < 0.0f )
	{
		return manifold;
	}

	float u = b2Dot( e, b2Sub( p2, pB ) );
	float v = b2Dot( e, b2Sub( pB, p1 ) );

	b2Vec2 pA;
	if ( v <= 0.0f )
	{
		// behind point1: the previous segment owns the Voronoi region unless pB is past it
		b2Vec2 prevEdge = b2Sub( p1, segmentA.ghost1 );
		float uPrev = b2Dot( prevEdge, b2Sub( pB, p1 ) );
		if ( uPrev <= 0.0f )
		{
			return manifold;
		}
		pA = p1;
	}
	else if ( u <= 0.0f )
	{
		b2Vec2 nextEdge = b2Sub( segmentA.ghost2, p2 );
		float vNext = b2Dot( nextEdge, b2Sub( pB, p2 ) );
		if ( vNext > 0.0f )
		{
			return manifold;
		}
		pA = p2;
	}
	else
	{
		float ee = b2Dot( e, e );
		pA = { u * p1.x + v * p2.x, u * p1.y + v * p2.y };
		pA = ee > 0.0f ? b2MulSV( 1.0f / ee, pA ) : p1;
	}

	float distance;
	b2Vec2 normal = b2GetLengthAndNormalize( &distance, b2Sub( pB, pA ) );

	float radius = circleB.radius;
	float separation = distance - radius;
	if ( separation > kSpeculativeDistance )
	{
		return manifold;
	}

	b2Vec2 cA = pA;
	b2Vec2 cB = b2MulAdd( pB, -radius, normal );

	manifold.normal = normal;
	manifold.points[0].point = b2Lerp( cA, cB, 0.5f );
	manifold.points[0].separation = separation;
	manifold.points[0].id = 0;
	manifold.pointCount = 1;
	return manifold;
}

// Clip incident segment (b1, b2) against reference segment (a1, a2) along normal.
B2G_DEV b2LocalManifold ClipSegments( b2Vec2 a1, b2Vec2 a2, b2Vec2 b1, b2Vec2 b2, b2Vec2 normal, float ra, float rb, uint16_t id1,
									  uint16_t id2 )
{
	b2LocalManifold manifold = EmptyManifold();

	b2Vec2 tangent = b2LeftPerp( normal );

	float lower1 = 0.0f;
	float upper1 = b2Dot( b2Sub( a2, a1 ), tangent );
	float upper2 = b2Dot( b2Sub( b1, a1 ), tangent );
	float lower2 = b2Dot( b2Sub( b2, a1 ), tangent );

	if ( upper2 < lower1 || upper1 < lower2 )
	{
		return manifold;
	}

	b2Vec2 vLower = b2;
	if ( lower2 < lower1 && upper2 - lower2 > FLT_EPSILON )
	{
		vLower = b2Lerp( b2, b1, ( lower1 - lower2 ) / ( upper2 - lower2 ) );
	}

	b2Vec2 vUpper = b1;
	if ( upper2 > upper1 && upper2 - lower2 > FLT_EPSILON )
	{
		vUpper = b2Lerp( b2, b1, ( upper1 - lower2 ) / ( upper2 - lower2 ) );
	}

	float separationLower = b2Dot( b2Sub( vLower, a1 ), normal );
	float separationUpper = b2Dot( b2Sub( vUpper, a1 ), normal );

	vLower = b2MulAdd( vLower, 0.5f * ( ra - rb - separationLower ), normal );
	vUpper = b2MulAdd( vUpper, 0.5f * ( ra - rb - separationUpper ), normal );

	float radius = ra + rb;

	manifold.normal = normal;
	manifold.points[0].point = vLower;
	manifold.points[0].separation = separationLower - radius;
	manifold.points[0].id = id1;
	manifold.points[1].point = vUpper;
	manifold.points[1].separation = separationUpper - radius;
	manifold.points[1].id = id2;
	manifold.pointCount = 2;
	return manifold;
}

enum NormalType
{
	kNormalSkip,  // non-smooth with respect to a convex neighbor vertex: ignore
	kNormalAdmit, // smooth: collide along it
	kNormalSnap,  // in the region of a concave vertex: snap to the segment normal
};

struct ChainSegmentParams
{
	b2Vec2 edge1;
	b2Vec2 normal0;
	b2Vec2 normal2;
	bool convex1;
	bool convex2;
};

// Gauss-map classification of a candidate normal against the segment's neighbors.
B2G_DEV NormalType ClassifyNormal( const ChainSegmentParams& params, b2Vec2 normal )
{
	const float sinTol = 0.01f;

	if ( b2Dot( normal, params.edge1 ) <= 0.0f )
	{
		if ( params.convex1 )
		{
			if ( b2Cross( normal, params.normal0 ) > sinTol )
			{
				return kNormalSkip;
			}
			return kNormalAdmit;
		}
		return kNormalSnap;
	}

	if ( params.convex2 )
	{
		if ( b2Cross( params.normal2, normal ) > sinTol )
		{
			return kNormalSkip;
		}
		return kNormalAdmit;
	}
	return kNormalSnap;
}

B2G_DEV b2LocalManifold CollideChainSegmentAndPolygon( const b2ChainSegment& segmentA, const b2Polygon& polygonB, b2Transform xf,
													   b2SimplexCache* cache )
{
	b2LocalManifold manifold = EmptyManifold();

	b2Vec2 centroidB = b2TransformPoint( xf, polygonB.centroid );
	float radiusB = polygonB.radius;

	b2Vec2 p1 = segmentA.segment.point1;
	b2Vec2 p2 = segmentA.segment.point2;

	b2Vec2 edge1 = b2Normalize( b2Sub( p2, p1 ) );

	ChainSegmentParams smooth;
	smooth.edge1 = edge1;

	const float convexTol = 0.01f;
	b2Vec2 edge0 = b2Normalize( b2Sub( p1, segmentA.ghost1 ) );
	smooth.normal0 = b2RightPerp( edge0 );
	smooth.convex1 = b2Cross( edge0, edge1 ) >= convexTol;

	b2Vec2 edge2 = b2Normalize( b2Sub( segmentA.ghost2, p2 ) );
	smooth.normal2 = b2RightPerp( edge2 );
	smooth.convex2 = b2Cross( edge1, edge2 ) >= convexTol;

	b2Vec2 normal1 = b2RightPerp( edge1 );
	bool behind1 = b2Dot( normal1, b2Sub( centroidB, p1 ) ) < 0.0f;
	bool behind0 = true;
	bool behind2 = true;
	if ( smooth.convex1 )
	{
		behind0 = b2Dot( smooth.normal0, b2Sub( centroidB, p1 ) ) < 0.0f;
	}
	if ( smooth.convex2 )
	{
		behind2 = b2Dot( smooth.normal2, b2Sub( centroidB, p2 ) ) < 0.0f;
	}

	if ( behind1 && behind0 && behind2 )
	{
		return manifold;
	}

	int count = polygonB.count;
	b2Vec2 vertices[B2_MAX_POLYGON_VERTICES];
	b2Vec2 normals[B2_MAX_POLYGON_VERTICES];
	for ( int i = 0; i < count; ++i )
	{
		vertices[i] = b2TransformPoint( xf, polygonB.vertices[i] );
		normals[i] = b2RotateVector( xf.q, polygonB.normals[i] );
	}

	b2DistanceInput input;
	input.proxyA = MakeProxy( &segmentA.segment.point1, 2, 0.0f );
	input.proxyB = MakeProxy( vertices, count, 0.0f );
	input.transform = { { 0.0f, 0.0f }, { 1.0f, 0.0f } };
	input.useRadii = false;

	b2DistanceOutput output = ShapeDistance( input, cache );

	if ( output.distance > radiusB + kSpeculativeDistance )
	{
		return manifold;
	}

	b2Vec2 n0 = smooth.convex1 ? smooth.normal0 : normal1;
	b2Vec2 n2 = smooth.convex2 ? smooth.normal2 : normal1;

	int incidentIndex = -1;
	int incidentNormal = -1;

	if ( behind1 == false && output.distance > 0.1f * kLinearSlop )
	{
		if ( cache->count == 1 )
		{
			// vertex-vertex
			b2Vec2 pA = output.pointA;
			b2Vec2 pB = output.pointB;
			b2Vec2 normal = b2Normalize( b2Sub( pB, pA ) );

			NormalType type = ClassifyNormal( smooth, normal );
			if ( type == kNormalSkip )
			{
				return manifold;
			}

			if ( type == kNormalAdmit )
			{
				manifold.normal = normal;
				manifold.points[0].point = pA;
				manifold.points[0].separation = output.distance - radiusB;
				manifold.points[0].id = B2G_MAKE_ID( cache->indexA[0], cache->indexB[0] );
				manifold.pointCount = 1;
				return manifold;
			}

			incidentIndex = cache->indexB[0];
		}
		else
		{
			// vertex-edge
			int ia1 = cache->indexA[0];
			int ia2 = cache->indexA[1];
			int ib1 = cache->indexB[0];
			int ib2 = cache->indexB[1];

			if ( ia1 == ia2 )
			{
				// one point on A, two on B: the polygon normal most aligned with the separation
				b2Vec2 normalB = b2Sub( output.pointA, output.pointB );
				float dot1 = b2Dot( normalB, normals[ib1] );
				float dot2 = b2Dot( normalB, normals[ib2] );
				int ib = dot1 > dot2 ? ib1 : ib2;

				normalB = normals[ib];

				NormalType type = ClassifyNormal( smooth, b2Neg( normalB ) );
				if ( type == kNormalSkip )
				{
					return manifold;
				}

				if ( type == kNormalAdmit )
				{
					ib1 = ib;
					ib2 = ib < count - 1 ? ib + 1 : 0;

					b2Vec2 b1 = vertices[ib1];
					b2Vec2 b2 = vertices[ib2];

					dot1 = b2Dot( normalB, b2Sub( p1, b1 ) );
					dot2 = b2Dot( normalB, b2Sub( p2, b1 ) );

					if ( dot1 < dot2 )
					{
						if ( b2Dot( n0, normalB ) < b2Dot( normal1, normalB ) )
						{
							return manifold;
						}
					}
					else
					{
						if ( b2Dot( n2, normalB ) < b2Dot( normal1, normalB ) )
						{
							return manifold;
						}
					}

					manifold = ClipSegments( b1, b2, p1, p2, normalB, radiusB, 0.0f, B2G_MAKE_ID( ib1, 1 ), B2G_MAKE_ID( ib2, 0 ) );
					if ( manifold.pointCount == 2 )
					{
						manifold.normal = b2Neg( normalB );
					}
					return manifold;
				}

				incidentNormal = ib;
			}
			else
			{
				float dot1 = b2Dot( normal1, b2Sub( vertices[ib1], p1 ) );
				float dot2 = b2Dot( normal1, b2Sub( vertices[ib2], p2 ) );
				incidentIndex = dot1 < dot2 ? ib1 : ib2;
			}
		}
	}
	else
	{
		// separating axis test: segment normal and the convex neighbors' normals
		float edgeSeparation = FLT_MAX;
		for ( int i = 0; i < count; ++i )
		{
			float s = b2Dot( normal1, b2Sub( vertices[i], p1 ) );
			if ( s < edgeSeparation )
			{
				edgeSeparation = s;
				incidentIndex = i;
			}
		}

		if ( smooth.convex1 )
		{
			float s0 = FLT_MAX;
			for ( int i = 0; i < count; ++i )
			{
				float s = b2Dot( smooth.normal0, b2Sub( vertices[i], p1 ) );
				if ( s < s0 )
				{
					s0 = s;
				}
			}
			if ( s0 > edgeSeparation )
			{
				edgeSeparation = s0;
				incidentIndex = -1;
			}
		}

		if ( smooth.convex2 )
		{
			float s2 = FLT_MAX;
			for ( int i = 0; i < count; ++i )
			{
				float s = b2Dot( smooth.normal2, b2Sub( vertices[i], p2 ) );
				if ( s < s2 )
				{
					s2 = s;
				}
			}
			if ( s2 > edgeSeparation )
			{
				edgeSeparation = s2;
				incidentIndex = -1;
			}
		}

		// polygon normals admitted by the Gauss map
		float polygonSeparation = -FLT_MAX;
		int referenceIndex = -1;
		for ( int i = 0; i < count; ++i )
		{
			b2Vec2 n = normals[i];
			if ( ClassifyNormal( smooth, b2Neg( n ) ) != kNormalAdmit )
			{
				continue;
			}
			b2Vec2 p = vertices[i];
			float s = b2MinFloat( b2Dot( n, b2Sub( p2, p ) ), b2Dot( n, b2Sub( p1, p ) ) );
			if ( s > polygonSeparation )
			{
				polygonSeparation = s;
				referenceIndex = i;
			}
		}

		if ( polygonSeparation > edgeSeparation )
		{
			int ia1 = referenceIndex;
			int ia2 = ia1 < count - 1 ? ia1 + 1 : 0;
			b2Vec2 a1 = vertices[ia1];
			b2Vec2 a2 = vertices[ia2];
			b2Vec2 n = normals[ia1];

			float dot1 = b2Dot( n, b2Sub( p1, a1 ) );
			float dot2 = b2Dot( n, b2Sub( p2, a1 ) );

			if ( dot1 < dot2 )
			{
				if ( b2Dot( n0, n ) < b2Dot( normal1, n ) )
				{
					return manifold;
				}
			}
			else
			{
				if ( b2Dot( n2, n ) < b2Dot( normal1, n ) )
				{
					return manifold;
				}
			}

			manifold = ClipSegments( a1, a2, p1, p2, normals[ia1], radiusB, 0.0f, B2G_MAKE_ID( ia1, 1 ), B2G_MAKE_ID( ia2, 0 ) );
			if ( manifold.pointCount == 2 )
			{
				manifold.normal = b2Neg( normals[ia1] );
			}
			return manifold;
		}

		if ( incidentIndex == -1 )
		{
			// a neighboring segment is the separating axis
			return manifold;
		}
	}

	// segment normal: incident polygon edge adjacent to the deepest vertex, most anti-parallel
	b2Vec2 b1, b2;
	int ib1, ib2;

	if ( incidentNormal != -1 )
	{
		ib1 = incidentNormal;
		ib2 = ib1 < count - 1 ? ib1 + 1 : 0;
		b1 = vertices[ib1];
		b2 = vertices[ib2];
	}
	else
	{
		int i2 = incidentIndex;
		int i1 = i2 > 0 ? i2 - 1 : count - 1;
		float d1 = b2Dot( normal1, normals[i1] );
		float d2 = b2Dot( normal1, normals[i2] );
		if ( d1 < d2 )
		{
			ib1 = i1;
			ib2 = i2;
		}
		else
		{
			ib1 = i2;
			ib2 = i2 < count - 1 ? i2 + 1 : 0;
		}
		b1 = vertices[ib1];
		b2 = vertices[ib2];
	}

	return ClipSegments( p1, p2, b1, b2, normal1, 0.0f, radiusB, B2G_MAKE_ID( 0, ib2 ), B2G_MAKE_ID( 1, ib1 ) );
}

B2G_DEV b2LocalManifold CollideChainSegmentAndCapsule( const b2ChainSegment& segmentA, const b2Capsule& capsuleB, b2Transform xf,
													   b2SimplexCache* cache )
{
	b2Polygon polyB;
	MakeCapsulePolygon( &polyB, capsuleB.center1, capsuleB.center2, capsuleB.radius );
	return CollideChainSegmentAndPolygon( segmentA, polyB, xf, cache );
}

// Manifold of a primary-ordered shape pair (reference registry src/contact.c:188-206); chain
// segment pairs read and update the contact's simplex cache.
B2G_DEV b2LocalManifold ComputeManifoldCached( const ShapeDev& shapeA, const ShapeDev& shapeB, b2Transform xf, b2SimplexCache* cache )
{
	if ( shapeA.type == b2_chainSegmentShape )
	{
		if ( shapeB.type == b2_circleShape )
		{
			return CollideChainSegmentAndCircle( shapeA.chainSegment, shapeB.circle, xf );
		}
		if ( shapeB.type == b2_capsuleShape )
		{
			return CollideChainSegmentAndCapsule( shapeA.chainSegment, shapeB.capsule, xf, cache );
		}
		if ( shapeB.type == b2_polygonShape )
		{
			return CollideChainSegmentAndPolygon( shapeA.chainSegment, shapeB.polygon, xf, cache );
		}
		return EmptyManifold();
	}
	return ComputeManifold( shapeA, shapeB, xf );
}

} // namespace b2g
