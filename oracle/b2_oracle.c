// CPU restatement of the Box2D v3.2 per-lane arithmetic. Test infrastructure only: see
// b2_oracle.h for who may use it and what pins each function.
//
// Build: gcc -std=gnu17 -O2 -ffp-contract=off (oracle/Makefile). Contraction must stay off: the
// reference is built that way for determinism (CMakeLists.txt:50-63) and every expression below
// keeps the reference's association order.

#include "b2_oracle.h"

#include <float.h>
#include <math.h>
#include <string.h>

#define O_PI 3.14159265359f
#define O_LINEAR_SLOP 0.005f
#define O_SPECULATIVE ( 4.0f * O_LINEAR_SLOP )
#define O_ID( a, b ) ( (uint16_t)( (uint8_t)( a ) << 8 | (uint8_t)( b ) ) )

// ---------------------------------------------------------------------------------------------
// vector helpers, include/box2d/math_functions.h:222-584 (same expressions)

static ov2 v( float x, float y )
{
	ov2 r = { x, y };
	return r;
}
static ov2 vsub( ov2 a, ov2 b )
{
	return v( a.x - b.x, a.y - b.y );
}
static ov2 vadd( ov2 a, ov2 b )
{
	return v( a.x + b.x, a.y + b.y );
}
static ov2 vneg( ov2 a )
{
	return v( -a.x, -a.y );
}
static float vdot( ov2 a, ov2 b )
{
	return a.x * b.x + a.y * b.y;
}
// a + s * b (b2MulAdd) and a - s * b (b2MulSub)
static ov2 vmadd( ov2 a, float s, ov2 b )
{
	return v( a.x + s * b.x, a.y + s * b.y );
}
static ov2 vmsub( ov2 a, float s, ov2 b )
{
	return v( a.x - s * b.x, a.y - s * b.y );
}
// (1 - t) * a + t * b (b2Lerp)
static ov2 vlerp( ov2 a, ov2 b, float t )
{
	return v( ( 1.0f - t ) * a.x + t * b.x, ( 1.0f - t ) * a.y + t * b.y );
}
static ov2 vleft( ov2 a )
{
	return v( -a.y, a.x );
}
static ov2 vright( ov2 a )
{
	return v( a.y, -a.x );
}
static float vdist2( ov2 a, ov2 b )
{
	ov2 c = v( b.x - a.x, b.y - a.y );
	return c.x * c.x + c.y * c.y;
}
// b2Normalize: zero vector below 1000 * FLT_MIN squared length
static ov2 vnorm( ov2 a )
{
	float l2 = a.x * a.x + a.y * a.y;
	if ( l2 > 1000.0f * FLT_MIN )
	{
		float s = 1.0f / sqrtf( l2 );
		return v( s * a.x, s * a.y );
	}
	return v( 0.0f, 0.0f );
}
// b2GetLengthAndNormalize
static ov2 vlen_norm( float* len, ov2 a )
{
	float l2 = a.x * a.x + a.y * a.y;
	if ( l2 > 1000.0f * FLT_MIN )
	{
		*len = sqrtf( l2 );
		float s = 1.0f / *len;
		return v( s * a.x, s * a.y );
	}
	*len = 0.0f;
	return v( 0.0f, 0.0f );
}
static ov2 rot_apply( orot q, ov2 a )
{
	return v( q.c * a.x - q.s * a.y, q.s * a.x + q.c * a.y );
}
// b2TransformPoint: rotation first, translation added last
static ov2 xf_apply( oxf t, ov2 p )
{
	float x = ( t.q.c * p.x - t.q.s * p.y ) + t.p.x;
	float y = ( t.q.s * p.x + t.q.c * p.y ) + t.p.y;
	return v( x, y );
}
static float clampf( float a, float lo, float hi )
{
	return a < lo ? lo : ( a > hi ? hi : a );
}

static omanifold empty_manifold( void )
{
	omanifold m;
	memset( &m, 0, sizeof( m ) );
	return m;
}

// ---------------------------------------------------------------------------------------------
// src/math_functions.c:96-136 (b2Atan2): minimax polynomial on [0, 1] mapped to the full circle

float o_atan2( float y, float x )
{
	if ( x == 0.0f && y == 0.0f )
	{
		return 0.0f;
	}
	float ax = fabsf( x );
	float ay = fabsf( y );
	float hi = ay > ax ? ay : ax;
	float lo = ay < ax ? ay : ax;
	float a = lo / hi;
	float a2 = a * a;
	float a3 = a2 * a;
	float a4 = a2 * a2;
	float p = 0.024840285f * a4 + 0.18681418f;
	float q = -0.094097948f * a4 - 0.33213072f;
	p = p * a2 + q;
	p = p * a3 + a;
	if ( ay > ax )
	{
		p = 1.57079637f - p;
	}
	if ( x < 0 )
	{
		p = 3.14159274f - p;
	}
	if ( y < 0 )
	{
		p = -p;
	}
	return p;
}

// include/box2d/math_functions.h:540-555 (b2UnwindAngle): round to the nearest turn in double
static float unwind( float radians )
{
	float x = clampf( radians, -1.0e6f, 1.0e6f );
	double turn = 2.0f * O_PI;
	double magic = 6755399441055744.0;
	double a = x;
	double k = ( a / turn + magic ) - magic;
	return (float)( a - k * turn );
}

// src/math_functions.c:138-183 (b2ComputeCosSin): Bhaskara rational approximations, normalized
void o_cos_sin( float radians, float* cOut, float* sOut )
{
	float x = unwind( radians );
	float pi2 = O_PI * O_PI;
	float c;
	if ( x < -0.5f * O_PI )
	{
		float y = x + O_PI;
		float y2 = y * y;
		c = -( pi2 - 4.0f * y2 ) / ( pi2 + y2 );
	}
	else if ( x > 0.5f * O_PI )
	{
		float y = x - O_PI;
		float y2 = y * y;
		c = -( pi2 - 4.0f * y2 ) / ( pi2 + y2 );
	}
	else
	{
		float y2 = x * x;
		c = ( pi2 - 4.0f * y2 ) / ( pi2 + y2 );
	}

	float s;
	if ( x < 0.0f )
	{
		float y = x + O_PI;
		s = -16.0f * y * ( O_PI - y ) / ( 5.0f * pi2 - 4.0f * y * ( O_PI - y ) );
	}
	else
	{
		s = 16.0f * x * ( O_PI - x ) / ( 5.0f * pi2 - 4.0f * x * ( O_PI - x ) );
	}

	float mag = sqrtf( s * s + c * c );
	float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
	*cOut = c * inv;
	*sOut = s * inv;
}

// src/solver.h:239-281 (b2MakeSoft)
osoft o_make_soft( float hertz, float zeta, float h )
{
	osoft r = { 0.0f, 0.0f, 0.0f };
	if ( hertz == 0.0f )
	{
		return r;
	}
	float omega = 2.0f * O_PI * hertz;
	float a1 = 2.0f * zeta + h * omega;
	float a2 = h * omega * a1;
	float a3 = 1.0f / ( 1.0f + a2 );
	r.biasRate = omega / a1;
	r.massScale = a2 * a3;
	r.impulseScale = a3;
	return r;
}

// ---------------------------------------------------------------------------------------------
// src/distance.c:33-106 (b2SegmentDistance), Ericson 5.1.9 with degenerate-segment handling

osegdist o_segment_distance( ov2 p1, ov2 q1, ov2 p2, ov2 q2 )
{
	osegdist r;
	memset( &r, 0, sizeof( r ) );
	ov2 d1 = vsub( q1, p1 );
	ov2 d2 = vsub( q2, p2 );
	ov2 w = vsub( p1, p2 );
	float dd1 = vdot( d1, d1 );
	float dd2 = vdot( d2, d2 );
	float rd1 = vdot( w, d1 );
	float rd2 = vdot( w, d2 );
	const float eps2 = FLT_EPSILON * FLT_EPSILON;

	if ( dd1 < eps2 || dd2 < eps2 )
	{
		if ( dd1 >= eps2 )
		{
			r.f1 = clampf( -rd1 / dd1, 0.0f, 1.0f );
			r.f2 = 0.0f;
		}
		else if ( dd2 >= eps2 )
		{
			r.f1 = 0.0f;
			r.f2 = clampf( rd2 / dd2, 0.0f, 1.0f );
		}
	}
	else
	{
		float d12 = vdot( d1, d2 );
		float den = dd1 * dd2 - d12 * d12;
		float f1 = 0.0f;
		if ( den != 0.0f )
		{
			f1 = clampf( ( d12 * rd2 - rd1 * dd2 ) / den, 0.0f, 1.0f );
		}
		float f2 = ( d12 * f1 + rd2 ) / dd2;
		if ( f2 < 0.0f )
		{
			f2 = 0.0f;
			f1 = clampf( -rd1 / dd1, 0.0f, 1.0f );
		}
		else if ( f2 > 1.0f )
		{
			f2 = 1.0f;
			f1 = clampf( ( d12 - rd1 ) / dd1, 0.0f, 1.0f );
		}
		r.f1 = f1;
		r.f2 = f2;
	}
	r.closest1 = vmadd( p1, r.f1, d1 );
	r.closest2 = vmadd( p2, r.f2, d2 );
	r.distSq = vdist2( r.closest1, r.closest2 );
	return r;
}

// ---------------------------------------------------------------------------------------------
// Manifolds, src/manifold.c

// Single point between two round features: point pA on A, pB on B (in A's frame).
static omanifold round_point( ov2 pA, ov2 pB, float rA, float rB )
{
	omanifold m = empty_manifold();
	float dist;
	ov2 n = vlen_norm( &dist, vsub( pB, pA ) );
	float sep = dist - rA - rB;
	if ( sep > O_SPECULATIVE )
	{
		return m;
	}
	ov2 cA = vmadd( pA, rA, n );
	ov2 cB = vmadd( pB, -rB, n );
	m.normal = n;
	m.pts[0].point = vlerp( cA, cB, 0.5f );
	m.pts[0].separation = sep;
	m.pts[0].id = 0;
	m.count = 1;
	return m;
}

// src/manifold.c:36-64 (b2CollideCircles)
omanifold o_collide_circles( const ocircle* a, const ocircle* b, oxf xf )
{
	return round_point( a->center, xf_apply( xf, b->center ), a->radius, b->radius );
}

// src/manifold.c:67-124 (b2CollideCapsuleAndCircle): closest point on the core segment
omanifold o_collide_capsule_circle( const ocapsule* a, const ocircle* b, oxf xf )
{
	ov2 pB = xf_apply( xf, b->center );
	ov2 e = vsub( a->c2, a->c1 );
	float s1 = vdot( vsub( pB, a->c1 ), e );
	float s2 = vdot( vsub( a->c2, pB ), e );
	ov2 pA;
	if ( s1 < 0.0f )
	{
		pA = a->c1;
	}
	else if ( s2 < 0.0f )
	{
		pA = a->c2;
	}
	else
	{
		float s = s1 / vdot( e, e );
		pA = vmadd( a->c1, s, e );
	}
	return round_point( pA, pB, a->radius, b->radius );
}

// src/manifold.c:127-234 (b2CollidePolygonAndCircle)
omanifold o_collide_polygon_circle( const opolygon* a, const ocircle* b, oxf xf )
{
	omanifold m = empty_manifold();
	ov2 c = xf_apply( xf, b->center );
	float rA = a->radius;
	float rB = b->radius;
	float radius = rA + rB;

	int best = 0;
	float sep = -FLT_MAX;
	for ( int i = 0; i < a->count; ++i )
	{
		float s = vdot( a->n[i], vsub( c, a->v[i] ) );
		if ( s > sep )
		{
			sep = s;
			best = i;
		}
	}
	if ( sep > radius + O_SPECULATIVE )
	{
		return m;
	}

	int i2 = best + 1 < a->count ? best + 1 : 0;
	ov2 v1 = a->v[best];
	ov2 v2 = a->v[i2];
	float u1 = vdot( vsub( c, v1 ), vsub( v2, v1 ) );
	float u2 = vdot( vsub( c, v2 ), vsub( v1, v2 ) );

	if ( ( u1 < 0.0f || u2 < 0.0f ) && sep > FLT_EPSILON )
	{
		// vertex region (v1 checked first)
		ov2 vert = u1 < 0.0f ? v1 : v2;
		ov2 n = vnorm( vsub( c, vert ) );
		float s = vdot( vsub( c, vert ), n );
		if ( s > radius + O_SPECULATIVE )
		{
			return m;
		}
		ov2 cA = vmadd( vert, rA, n );
		ov2 cB = vmsub( c, rB, n );
		m.normal = n;
		m.pts[0].point = vlerp( cA, cB, 0.5f );
		m.pts[0].separation = vdot( vsub( cB, cA ), n );
		m.pts[0].id = 0;
		m.count = 1;
		return m;
	}

	// face region; the center may be inside
	ov2 n = a->n[best];
	m.normal = n;
	ov2 cA = vmadd( c, rA - vdot( vsub( c, v1 ), n ), n );
	ov2 cB = vmsub( c, rB, n );
	m.pts[0].point = vlerp( cA, cB, 0.5f );
	m.pts[0].separation = sep - radius;
	m.pts[0].id = 0;
	m.count = 1;
	return m;
}

// Clip the B segment (cp, cq along the A segment's frame) for the capsule clip case.
// fp/fq are the projections of the incident end points on the reference axis.
static void clip_to_range( ov2 p, ov2 q, float fp, float fq, float len, ov2* cp, ov2* cq )
{
	*cp = p;
	*cq = q;
	if ( fp < 0.0f && fq > 0.0f )
	{
		*cp = vlerp( p, q, ( 0.0f - fp ) / ( fq - fp ) );
	}
	else if ( fq < 0.0f && fp > 0.0f )
	{
		*cq = vlerp( q, p, ( 0.0f - fq ) / ( fp - fq ) );
	}
	if ( fp > len && fq < len )
	{
		*cp = vlerp( p, q, ( fp - len ) / ( fp - fq ) );
	}
	else if ( fq > len && fp < len )
	{
		*cq = vlerp( q, p, ( fq - len ) / ( fq - fp ) );
	}
}

// SAT separation of segment (p, q) against the line through o with unit normal n, sign chosen
// so the separation is the larger of the two sides (src/manifold.c:366-401).
static float side_separation( ov2 o, ov2 axis, ov2 p, ov2 q, ov2* normalOut )
{
	ov2 n = vleft( axis );
	float a = vdot( vsub( p, o ), n );
	float b = vdot( vsub( q, o ), n );
	float pos = a < b ? a : b;
	float neg = -a < -b ? -a : -b;
	if ( pos > neg )
	{
		*normalOut = n;
		return pos;
	}
	*normalOut = vneg( n );
	return neg;
}

// src/manifold.c:237-495 (b2CollideCapsules)
omanifold o_collide_capsules( const ocapsule* a, const ocapsule* b, oxf xf )
{
	ov2 origin = a->c1;
	oxf xs = { vsub( xf.p, origin ), xf.q };
	ov2 p1 = v( 0.0f, 0.0f );
	ov2 q1 = vsub( a->c2, origin );
	ov2 p2 = xf_apply( xs, b->c1 );
	ov2 q2 = xf_apply( xs, b->c2 );

	// closest points of the two core segments (same steps as o_segment_distance's main branch)
	ov2 d1 = vsub( q1, p1 );
	ov2 d2 = vsub( q2, p2 );
	float dd1 = vdot( d1, d1 );
	float dd2 = vdot( d2, d2 );
	const float eps2 = FLT_EPSILON * FLT_EPSILON;
	ov2 w = vsub( p1, p2 );
	float rd1 = vdot( w, d1 );
	float rd2 = vdot( w, d2 );
	float d12 = vdot( d1, d2 );
	float den = dd1 * dd2 - d12 * d12;
	float f1 = 0.0f;
	if ( den != 0.0f )
	{
		f1 = clampf( ( d12 * rd2 - rd1 * dd2 ) / den, 0.0f, 1.0f );
	}
	float f2 = ( d12 * f1 + rd2 ) / dd2;
	if ( f2 < 0.0f )
	{
		f2 = 0.0f;
		f1 = clampf( -rd1 / dd1, 0.0f, 1.0f );
	}
	else if ( f2 > 1.0f )
	{
		f2 = 1.0f;
		f1 = clampf( ( d12 - rd1 ) / dd1, 0.0f, 1.0f );
	}
	ov2 k1 = vmadd( p1, f1, d1 );
	ov2 k2 = vmadd( p2, f2, d2 );
	float dist2 = vdist2( k1, k2 );

	omanifold m = empty_manifold();
	float rA = a->radius;
	float rB = b->radius;
	float radius = rA + rB;
	float maxDist = radius + O_SPECULATIVE;
	if ( dist2 > maxDist * maxDist )
	{
		return m;
	}
	float dist = sqrtf( dist2 );

	float len1, len2;
	ov2 u1 = vlen_norm( &len1, d1 );
	ov2 u2 = vlen_norm( &len2, d2 );

	float fp2 = vdot( vsub( p2, p1 ), u1 );
	float fq2 = vdot( vsub( q2, p1 ), u1 );
	int outsideA = ( fp2 <= 0.0f && fq2 <= 0.0f ) || ( fp2 >= len1 && fq2 >= len1 );
	float fp1 = vdot( vsub( p1, p2 ), u2 );
	float fq1 = vdot( vsub( q1, p2 ), u2 );
	int outsideB = ( fp1 <= 0.0f && fq1 <= 0.0f ) || ( fp1 >= len2 && fq1 >= len2 );

	if ( !outsideA && !outsideB )
	{
		ov2 nA, nB;
		float sA = side_separation( p1, u1, p2, q2, &nA );
		float sB = side_separation( p2, u2, p1, q1, &nB );

		if ( sA + 0.1f * O_LINEAR_SLOP >= sB )
		{
			m.normal = nA;
			ov2 cp, cq;
			clip_to_range( p2, q2, fp2, fq2, len1, &cp, &cq );
			float sp = vdot( vsub( cp, p1 ), nA );
			float sq = vdot( vsub( cq, p1 ), nA );
			if ( sp <= dist + O_LINEAR_SLOP || sq <= dist + O_LINEAR_SLOP )
			{
				m.pts[0].point = vmadd( cp, 0.5f * ( rA - rB - sp ), nA );
				m.pts[0].separation = sp - radius;
				m.pts[0].id = O_ID( 0, 0 );
				m.pts[1].point = vmadd( cq, 0.5f * ( rA - rB - sq ), nA );
				m.pts[1].separation = sq - radius;
				m.pts[1].id = O_ID( 0, 1 );
				m.count = 2;
			}
		}
		else
		{
			m.normal = vneg( nB );
			ov2 cp, cq;
			clip_to_range( p1, q1, fp1, fq1, len2, &cp, &cq );
			float sp = vdot( vsub( cp, p2 ), nB );
			float sq = vdot( vsub( cq, p2 ), nB );
			if ( sp <= dist + O_LINEAR_SLOP || sq <= dist + O_LINEAR_SLOP )
			{
				m.pts[0].point = vmadd( cp, 0.5f * ( rB - rA - sp ), nB );
				m.pts[0].separation = sp - radius;
				m.pts[0].id = O_ID( 0, 0 );
				m.pts[1].point = vmadd( cq, 0.5f * ( rB - rA - sq ), nB );
				m.pts[1].separation = sq - radius;
				m.pts[1].id = O_ID( 1, 0 );
				m.count = 2;
			}
		}
	}

	if ( m.count == 0 )
	{
		ov2 n = vsub( k2, k1 );
		n = vdot( n, n ) > eps2 ? vnorm( n ) : vleft( u1 );
		ov2 c1 = vmadd( k1, rA, n );
		ov2 c2 = vmadd( k2, -rB, n );
		m.normal = n;
		m.pts[0].point = vlerp( c1, c2, 0.5f );
		m.pts[0].separation = sqrtf( dist2 ) - radius;
		m.pts[0].id = O_ID( f1 == 0.0f ? 0 : 1, f2 == 0.0f ? 0 : 1 );
		m.count = 1;
	}

	for ( int i = 0; i < m.count; ++i )
	{
		m.pts[i].point = vadd( m.pts[i].point, origin );
	}
	return m;
}

// src/manifold.c:16-34 (b2MakeCapsule): a two-sided rounded segment as a 2-gon
static opolygon capsule_polygon( ov2 p1, ov2 p2, float radius )
{
	opolygon g;
	memset( &g, 0, sizeof( g ) );
	g.v[0] = p1;
	g.v[1] = p2;
	g.centroid = vlerp( p1, p2, 0.5f );
	ov2 axis = vnorm( vsub( p2, p1 ) );
	ov2 n = vright( axis );
	g.n[0] = n;
	g.n[1] = vneg( n );
	g.count = 2;
	g.radius = radius;
	return g;
}

// src/manifold.c:511-682 (b2ClipPolygons): clip the incident edge against the reference edge
static omanifold clip_polygons( const opolygon* A, const opolygon* B, int edgeA, int edgeB, int flip )
{
	omanifold m = empty_manifold();
	const opolygon* ref = flip ? B : A;
	const opolygon* inc = flip ? A : B;
	int r1 = flip ? edgeB : edgeA;
	int r2 = r1 + 1 < ref->count ? r1 + 1 : 0;
	int n1 = flip ? edgeA : edgeB;
	int n2 = n1 + 1 < inc->count ? n1 + 1 : 0;

	ov2 normal = ref->n[r1];
	ov2 v11 = ref->v[r1];
	ov2 v12 = ref->v[r2];
	ov2 v21 = inc->v[n1];
	ov2 v22 = inc->v[n2];
	ov2 tangent = v( -1.0f * normal.y, 1.0f * normal.x );

	float lower1 = 0.0f;
	float upper1 = vdot( vsub( v12, v11 ), tangent );
	float upper2 = vdot( vsub( v21, v11 ), tangent );
	float lower2 = vdot( vsub( v22, v11 ), tangent );
	if ( upper2 < lower1 || upper1 < lower2 )
	{
		return m;
	}

	ov2 vLower = v22;
	if ( lower2 < lower1 && upper2 - lower2 > FLT_EPSILON )
	{
		vLower = vlerp( v22, v21, ( lower1 - lower2 ) / ( upper2 - lower2 ) );
	}
	ov2 vUpper = v21;
	if ( upper2 > upper1 && upper2 - lower2 > FLT_EPSILON )
	{
		vUpper = vlerp( v22, v21, ( upper1 - lower2 ) / ( upper2 - lower2 ) );
	}

	float sLower = vdot( vsub( vLower, v11 ), normal );
	float sUpper = vdot( vsub( vUpper, v11 ), normal );
	float rr = ref->radius;
	float ri = inc->radius;
	vLower = vmadd( vLower, 0.5f * ( rr - ri - sLower ), normal );
	vUpper = vmadd( vUpper, 0.5f * ( rr - ri - sUpper ), normal );
	float radius = rr + ri;

	if ( !flip )
	{
		m.normal = normal;
		m.pts[0].point = vLower;
		m.pts[0].separation = sLower - radius;
		m.pts[0].id = O_ID( r1, n2 );
		m.pts[1].point = vUpper;
		m.pts[1].separation = sUpper - radius;
		m.pts[1].id = O_ID( r2, n1 );
	}
	else
	{
		m.normal = vneg( normal );
		m.pts[0].point = vUpper;
		m.pts[0].separation = sUpper - radius;
		m.pts[0].id = O_ID( n1, r2 );
		m.pts[1].point = vLower;
		m.pts[1].separation = sLower - radius;
		m.pts[1].id = O_ID( n2, r1 );
	}
	m.count = 2;
	return m;
}

// src/manifold.c:684-720 (b2FindMaxSeparation): best edge normal of p1 against p2's vertices
static float max_separation( int* edge, const opolygon* p1, const opolygon* p2 )
{
	int best = 0;
	float maxSep = -FLT_MAX;
	for ( int i = 0; i < p1->count; ++i )
	{
		float si = FLT_MAX;
		for ( int j = 0; j < p2->count; ++j )
		{
			float sij = vdot( p1->n[i], vsub( p2->v[j], p1->v[i] ) );
			if ( sij < si )
			{
				si = sij;
			}
		}
		if ( si > maxSep )
		{
			maxSep = si;
			best = i;
		}
	}
	*edge = best;
	return maxSep;
}

// Edge of p whose normal is most anti-parallel to dir (first minimum wins).
static int incident_edge( const opolygon* p, ov2 dir )
{
	int e = 0;
	float minDot = FLT_MAX;
	for ( int i = 0; i < p->count; ++i )
	{
		float d = vdot( dir, p->n[i] );
		if ( d < minDot )
		{
			minDot = d;
			e = i;
		}
	}
	return e;
}

// src/manifold.c:740-1018 (b2CollidePolygons)
omanifold o_collide_polygons( const opolygon* a, const opolygon* b, oxf xf )
{
	ov2 origin = a->v[0];
	oxf xs = { vsub( xf.p, origin ), xf.q };

	opolygon A, B;
	memset( &A, 0, sizeof( A ) );
	memset( &B, 0, sizeof( B ) );
	A.count = a->count;
	A.radius = a->radius;
	A.v[0] = v( 0.0f, 0.0f );
	A.n[0] = a->n[0];
	for ( int i = 1; i < A.count; ++i )
	{
		A.v[i] = vsub( a->v[i], origin );
		A.n[i] = a->n[i];
	}
	B.count = b->count;
	B.radius = b->radius;
	for ( int i = 0; i < B.count; ++i )
	{
		B.v[i] = xf_apply( xs, b->v[i] );
		B.n[i] = rot_apply( xs.q, b->n[i] );
	}

	int edgeA = 0, edgeB = 0;
	float sepA = max_separation( &edgeA, &A, &B );
	float sepB = max_separation( &edgeB, &B, &A );
	float radius = A.radius + B.radius;
	omanifold m = empty_manifold();
	if ( sepA > O_SPECULATIVE + radius || sepB > O_SPECULATIVE + radius )
	{
		return m;
	}

	int flip = !( sepA >= sepB );
	if ( !flip )
	{
		edgeB = incident_edge( &B, A.n[edgeA] );
	}
	else
	{
		edgeA = incident_edge( &A, B.n[edgeB] );
	}

	if ( sepA > 0.1f * O_LINEAR_SLOP || sepB > 0.1f * O_LINEAR_SLOP )
	{
		// separated: closest features of the two edges decide between clipping and a vertex pair
		int i11 = edgeA;
		int i12 = edgeA + 1 < A.count ? edgeA + 1 : 0;
		int i21 = edgeB;
		int i22 = edgeB + 1 < B.count ? edgeB + 1 : 0;
		ov2 v11 = A.v[i11], v12 = A.v[i12], v21 = B.v[i21], v22 = B.v[i22];
		osegdist sd = o_segment_distance( v11, v12, v21, v22 );
		float dist = sqrtf( sd.distSq );
		float sep = dist - radius;
		if ( dist - radius > O_SPECULATIVE )
		{
			return m;
		}

		m = clip_polygons( &A, &B, edgeA, edgeB, flip );
		float minSep = FLT_MAX;
		for ( int i = 0; i < m.count; ++i )
		{
			minSep = minSep < m.pts[i].separation ? minSep : m.pts[i].separation;
		}

		if ( sep + 0.1f * O_LINEAR_SLOP < minSep )
		{
			int endA = -1, endB = -1;
			if ( sd.f1 == 0.0f )
			{
				endA = 0;
			}
			else if ( sd.f1 == 1.0f )
			{
				endA = 1;
			}
			if ( sd.f2 == 0.0f )
			{
				endB = 0;
			}
			else if ( sd.f2 == 1.0f )
			{
				endB = 1;
			}
			if ( endA >= 0 && endB >= 0 )
			{
				// vertex-vertex contact
				int ia = endA == 0 ? i11 : i12;
				int ib = endB == 0 ? i21 : i22;
				ov2 va = endA == 0 ? v11 : v12;
				ov2 vb = endB == 0 ? v21 : v22;
				ov2 n = vsub( vb, va );
				float inv = 1.0f / dist;
				n.x *= inv;
				n.y *= inv;
				ov2 c1 = vmadd( va, A.radius, n );
				ov2 c2 = vmadd( vb, -B.radius, n );
				// the second clip point stays in the record, as in the reference
				m.normal = n;
				m.pts[0].point = vlerp( c1, c2, 0.5f );
				m.pts[0].separation = dist - radius;
				m.pts[0].id = O_ID( ia, ib );
				m.count = 1;
			}
		}
	}
	else
	{
		m = clip_polygons( &A, &B, edgeA, edgeB, flip );
	}

	for ( int i = 0; i < m.count; ++i )
	{
		m.pts[i].point = vadd( m.pts[i].point, origin );
	}
	return m;
}

// src/manifold.c:498-502 (b2CollidePolygonAndCapsule)
omanifold o_collide_polygon_capsule( const opolygon* a, const ocapsule* b, oxf xf )
{
	opolygon g = capsule_polygon( b->c1, b->c2, b->radius );
	return o_collide_polygons( a, &g, xf );
}

// src/manifold.c:1020-1040 (segment forms reuse the capsule and polygon routines)
omanifold o_collide_segment_circle( const osegment* a, const ocircle* b, oxf xf )
{
	ocapsule c = { a->p1, a->p2, 0.0f };
	return o_collide_capsule_circle( &c, b, xf );
}

omanifold o_collide_segment_capsule( const osegment* a, const ocapsule* b, oxf xf )
{
	ocapsule c = { a->p1, a->p2, 0.0f };
	return o_collide_capsules( &c, b, xf );
}

omanifold o_collide_segment_polygon( const osegment* a, const opolygon* b, oxf xf )
{
	opolygon g = capsule_polygon( a->p1, a->p2, 0.0f );
	return o_collide_polygons( &g, b, xf );
}

// ---------------------------------------------------------------------------------------------
// batch entry points (b2ShapeType: circle 0, capsule 1, segment 2, polygon 3)

int b2o_collide_batch( const oshape* sa, const oshape* sb, const oxf* xfs, int count, omanifold* out )
{
	for ( int i = 0; i < count; ++i )
	{
		const oshape* a = sa + i;
		const oshape* b = sb + i;
		int ta = a->type, tb = b->type;
		omanifold m;
		if ( ta == 0 && tb == 0 )
			m = o_collide_circles( &a->u.circle, &b->u.circle, xfs[i] );
		else if ( ta == 1 && tb == 0 )
			m = o_collide_capsule_circle( &a->u.capsule, &b->u.circle, xfs[i] );
		else if ( ta == 1 && tb == 1 )
			m = o_collide_capsules( &a->u.capsule, &b->u.capsule, xfs[i] );
		else if ( ta == 3 && tb == 0 )
			m = o_collide_polygon_circle( &a->u.polygon, &b->u.circle, xfs[i] );
		else if ( ta == 3 && tb == 1 )
			m = o_collide_polygon_capsule( &a->u.polygon, &b->u.capsule, xfs[i] );
		else if ( ta == 3 && tb == 3 )
			m = o_collide_polygons( &a->u.polygon, &b->u.polygon, xfs[i] );
		else if ( ta == 2 && tb == 0 )
			m = o_collide_segment_circle( &a->u.segment, &b->u.circle, xfs[i] );
		else if ( ta == 2 && tb == 1 )
			m = o_collide_segment_capsule( &a->u.segment, &b->u.capsule, xfs[i] );
		else if ( ta == 2 && tb == 3 )
			m = o_collide_segment_polygon( &a->u.segment, &b->u.polygon, xfs[i] );
		else
			return -1;
		out[i] = m;
	}
	return 0;
}

void b2o_atan2_batch( const float* y, const float* x, float* out, int count )
{
	for ( int i = 0; i < count; ++i )
	{
		out[i] = o_atan2( y[i], x[i] );
	}
}

void b2o_cos_sin_batch( const float* radians, float* c, float* s, int count )
{
	for ( int i = 0; i < count; ++i )
	{
		o_cos_sin( radians[i], c + i, s + i );
	}
}

void b2o_segment_distance_batch( const ov2* p1, const ov2* q1, const ov2* p2, const ov2* q2, osegdist* out, int count )
{
	for ( int i = 0; i < count; ++i )
	{
		out[i] = o_segment_distance( p1[i], q1[i], p2[i], q2[i] );
	}
}
