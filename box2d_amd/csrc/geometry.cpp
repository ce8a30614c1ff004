// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the hull, polygon, mass and bounding-box arithmetic
// below restates its src/hull.c and src/geometry.c in the same operation order)
// SPDX-License-Identifier: MIT
// Creation-time geometry and math exports: polygon builders, convex hull, mass properties,
// shape bounding boxes, deterministic trig, default definitions and the determinism hash.
//
// These run on the host when a scene is built (not inside b2World_Step). Their arithmetic keeps
// the reference's operation order because their outputs (vertices, normals, centroids, masses,
// inertias, initial AABBs) are inputs to the simulated step and must be bit-identical:
//   polygons        src/geometry.c:23-252
//   mass            src/geometry.c:254-400
//   AABBs           src/geometry.c:402-482 (built in double, narrowed to float once)
//   hull            src/hull.c:11-328
//   trig            src/math_functions.c:96-183
//   defaults        src/types.c:8-94, src/joint.c:23-94

#include "core.h"

#include <cfloat>
#include <chrono>
#include <cmath>
#include <thread>

static float s_lengthUnitsPerMeter = 1.0f;

extern "C" {

void b2SetLengthUnitsPerMeter( float lengthUnits )
{
	s_lengthUnitsPerMeter = lengthUnits;
}

float b2GetLengthUnitsPerMeter( void )
{
	return s_lengthUnitsPerMeter;
}

b2Version b2GetVersion( void )
{
	return b2Version{ 3, 2, 0 };
}

bool b2IsDoublePrecision( void )
{
	return false;
}

uint32_t b2Hash( uint32_t hash, const uint8_t* data, int count )
{
	uint32_t h = hash;
	for ( int i = 0; i < count; ++i )
	{
		h = ( h << 5 ) + h + data[i];
	}
	return h;
}

uint64_t b2GetTicks( void )
{
	return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
			   std::chrono::steady_clock::now().time_since_epoch() )
		.count();
}

float b2GetMilliseconds( uint64_t ticks )
{
	return (float)( (double)( b2GetTicks() - ticks ) * 1.0e-6 );
}

float b2GetMillisecondsAndReset( uint64_t* ticks )
{
	uint64_t now = b2GetTicks();
	float ms = (float)( (double)( now - *ticks ) * 1.0e-6 );
	*ticks = now;
	return ms;
}

void b2Yield( void )
{
	std::this_thread::yield();
}

bool b2IsValidFloat( float a )
{
	return std::isfinite( a );
}

bool b2IsValidVec2( b2Vec2 v )
{
	return std::isfinite( v.x ) && std::isfinite( v.y );
}

bool b2IsValidRotation( b2Rot q )
{
	if ( std::isfinite( q.s ) == false || std::isfinite( q.c ) == false )
	{
		return false;
	}
	return b2IsNormalizedRot( q );
}

bool b2IsValidTransform( b2Transform t )
{
	return b2IsValidVec2( t.p ) && b2IsValidRotation( t.q );
}

bool b2IsValidAABB( b2AABB a )
{
	b2Vec2 d = b2Sub( a.upperBound, a.lowerBound );
	bool valid = d.x >= 0.0f && d.y >= 0.0f;
	return valid && b2IsValidVec2( a.lowerBound ) && b2IsValidVec2( a.upperBound );
}

bool b2IsValidPlane( b2Plane a )
{
	return b2IsValidVec2( a.normal ) && b2IsNormalized( a.normal ) && std::isfinite( a.offset );
}

bool b2IsValidPosition( b2Pos p )
{
	return b2IsValidVec2( p );
}

bool b2IsValidWorldTransform( b2WorldTransform t )
{
	return b2IsValidTransform( t );
}

// Minimax polynomial arctangent, identical coefficients and order to the reference.
float b2Atan2( float y, float x )
{
	if ( x == 0.0f && y == 0.0f )
	{
		return 0.0f;
	}

	float ax = b2AbsFloat( x );
	float ay = b2AbsFloat( y );
	float mx = b2MaxFloat( ay, ax );
	float mn = b2MinFloat( ay, ax );
	float a = mn / mx;

	float s = a * a;
	float c = s * a;
	float q = s * s;
	float r = 0.024840285f * q + 0.18681418f;
	float t = -0.094097948f * q - 0.33213072f;
	r = r * s + t;
	r = r * c + a;

	if ( ay > ax )
	{
		r = 1.57079637f - r;
	}
	if ( x < 0 )
	{
		r = 3.14159274f - r;
	}
	if ( y < 0 )
	{
		r = -r;
	}
	return r;
}

// Bhaskara-style rational cosine and sine, renormalized.
b2CosSin b2ComputeCosSin( float radians )
{
	float x = b2UnwindAngle( radians );
	float pi2 = B2_PI * B2_PI;

	float c;
	if ( x < -0.5f * B2_PI )
	{
		float y = x + B2_PI;
		float y2 = y * y;
		c = -( pi2 - 4.0f * y2 ) / ( pi2 + y2 );
	}
	else if ( x > 0.5f * B2_PI )
	{
		float y = x - B2_PI;
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
		float y = x + B2_PI;
		s = -16.0f * y * ( B2_PI - y ) / ( 5.0f * pi2 - 4.0f * y * ( B2_PI - y ) );
	}
	else
	{
		s = 16.0f * x * ( B2_PI - x ) / ( 5.0f * pi2 - 4.0f * x * ( B2_PI - x ) );
	}

	float mag = sqrtf( s * s + c * c );
	float invMag = mag > 0.0f ? 1.0f / mag : 0.0f;
	b2CosSin cs = { c * invMag, s * invMag };
	return cs;
}

b2Rot b2ComputeRotationBetweenUnitVectors( b2Vec2 v1, b2Vec2 v2 )
{
	b2Rot rot;
	rot.c = b2Dot( v1, v2 );
	rot.s = b2Cross( v1, v2 );
	return b2NormalizeRot( rot );
}

// ---------------------------------------------------------------------------------------------
// Default definitions. internalValue carries the reference's definition cookie so definitions
// made by either library are accepted by both.

static const int kDefCookie = 1152023;

b2WorldDef b2DefaultWorldDef( void )
{
	float lengthUnits = b2GetLengthUnitsPerMeter();
	b2WorldDef def = {};
	def.gravity = b2Vec2{ 0.0f, -10.0f };
	def.hitEventThreshold = 1.0f * lengthUnits;
	def.restitutionThreshold = 1.0f * lengthUnits;
	def.contactSpeed = 3.0f * lengthUnits;
	def.contactHertz = 30.0;
	def.contactDampingRatio = 10.0f;
	def.maximumLinearSpeed = 400.0f * lengthUnits;
	def.enableSleep = true;
	def.enableContinuous = true;
	def.internalValue = kDefCookie;
	return def;
}

b2BodyDef b2DefaultBodyDef( void )
{
	b2BodyDef def = {};
	def.type = b2_staticBody;
	def.rotation = b2Rot{ 1.0f, 0.0f };
	def.safetyFactor = 0.5f;
	def.sleepThreshold = 0.05f * b2GetLengthUnitsPerMeter();
	def.gravityScale = 1.0f;
	def.enableSleep = true;
	def.enableContactRecycling = true;
	def.isAwake = true;
	def.isEnabled = true;
	def.internalValue = kDefCookie;
	return def;
}

b2Filter b2DefaultFilter( void )
{
	b2Filter filter = { B2_DEFAULT_CATEGORY_BITS, B2_DEFAULT_MASK_BITS, 0 };
	return filter;
}

b2QueryFilter b2DefaultQueryFilter( void )
{
	b2QueryFilter filter = { B2_DEFAULT_CATEGORY_BITS, B2_DEFAULT_MASK_BITS };
	return filter;
}

b2ShapeDef b2DefaultShapeDef( void )
{
	b2ShapeDef def = {};
	def.material.friction = 0.6f;
	def.density = 1.0f;
	def.filter = b2DefaultFilter();
	def.updateBodyMass = true;
	def.invokeContactCreation = true;
	def.internalValue = kDefCookie;
	return def;
}

b2SurfaceMaterial b2DefaultSurfaceMaterial( void )
{
	b2SurfaceMaterial material = {};
	material.friction = 0.6f;
	return material;
}

b2ChainDef b2DefaultChainDef( void )
{
	static b2SurfaceMaterial defaultMaterial = { 0.6f, 0.0f, 0.0f, 0.0f, 0, 0 };
	b2ChainDef def = {};
	def.materials = &defaultMaterial;
	def.materialCount = 1;
	def.filter = b2DefaultFilter();
	def.internalValue = kDefCookie;
	return def;
}

// Joint definition defaults (reference src/joint.c:23-113).
static b2JointDef DefaultJointDef()
{
	b2JointDef def = {};
	def.localFrameA.q = b2Rot{ 1.0f, 0.0f };
	def.localFrameB.q = b2Rot{ 1.0f, 0.0f };
	def.forceThreshold = FLT_MAX;
	def.torqueThreshold = FLT_MAX;
	def.constraintHertz = 60.0f;
	def.constraintDampingRatio = 2.0f;
	def.drawScale = b2GetLengthUnitsPerMeter();
	return def;
}

b2RevoluteJointDef b2DefaultRevoluteJointDef( void )
{
	b2RevoluteJointDef def = {};
	def.base = DefaultJointDef();
	def.internalValue = kDefCookie;
	return def;
}

b2DistanceJointDef b2DefaultDistanceJointDef( void )
{
	b2DistanceJointDef def = {};
	def.base = DefaultJointDef();
	def.lowerSpringForce = -FLT_MAX;
	def.upperSpringForce = FLT_MAX;
	def.length = 1.0f;
	def.maxLength = B2_HUGE;
	def.internalValue = kDefCookie;
	return def;
}

b2FilterJointDef b2DefaultFilterJointDef( void )
{
	b2FilterJointDef def = {};
	def.base = DefaultJointDef();
	def.internalValue = kDefCookie;
	return def;
}

b2MotorJointDef b2DefaultMotorJointDef( void )
{
	b2MotorJointDef def = {};
	def.base = DefaultJointDef();
	def.internalValue = kDefCookie;
	return def;
}

b2MoverJointDef b2DefaultMoverJointDef( void )
{
	b2MoverJointDef def = {};
	def.base = DefaultJointDef();
	def.internalValue = kDefCookie;
	return def;
}

b2PogoJointDef b2DefaultPogoJointDef( void )
{
	b2PogoJointDef def = {};
	def.base = DefaultJointDef();
	def.internalValue = kDefCookie;
	return def;
}

b2PrismaticJointDef b2DefaultPrismaticJointDef( void )
{
	b2PrismaticJointDef def = {};
	def.base = DefaultJointDef();
	def.internalValue = kDefCookie;
	return def;
}

b2WeldJointDef b2DefaultWeldJointDef( void )
{
	b2WeldJointDef def = {};
	def.base = DefaultJointDef();
	def.internalValue = kDefCookie;
	return def;
}

b2WheelJointDef b2DefaultWheelJointDef( void )
{
	b2WheelJointDef def = {};
	def.base = DefaultJointDef();
	def.enableSpring = true;
	def.hertz = 1.0f;
	def.dampingRatio = 0.7f;
	def.internalValue = kDefCookie;
	return def;
}

// ---------------------------------------------------------------------------------------------
// Polygons

static b2Vec2 PolygonCentroid( const b2Vec2* vertices, int count )
{
	b2Vec2 center = { 0.0f, 0.0f };
	float area = 0.0f;
	b2Vec2 origin = vertices[0];
	const float inv3 = 1.0f / 3.0f;

	for ( int i = 1; i < count - 1; ++i )
	{
		b2Vec2 e1 = b2Sub( vertices[i], origin );
		b2Vec2 e2 = b2Sub( vertices[i + 1], origin );
		float a = 0.5f * b2Cross( e1, e2 );
		center = b2MulAdd( center, a * inv3, b2Add( e1, e2 ) );
		area += a;
	}

	float invArea = 1.0f / area;
	center.x *= invArea;
	center.y *= invArea;
	return b2Add( origin, center );
}

static void FillNormals( b2Polygon* shape )
{
	for ( int i = 0; i < shape->count; ++i )
	{
		int i2 = i + 1 < shape->count ? i + 1 : 0;
		b2Vec2 edge = b2Sub( shape->vertices[i2], shape->vertices[i] );
		shape->normals[i] = b2Normalize( b2CrossVS( edge, 1.0f ) );
	}
}

b2Polygon b2MakeBox( float halfWidth, float halfHeight )
{
	b2Polygon shape = {};
	shape.count = 4;
	shape.vertices[0] = b2Vec2{ -halfWidth, -halfHeight };
	shape.vertices[1] = b2Vec2{ halfWidth, -halfHeight };
	shape.vertices[2] = b2Vec2{ halfWidth, halfHeight };
	shape.vertices[3] = b2Vec2{ -halfWidth, halfHeight };
	shape.normals[0] = b2Vec2{ 0.0f, -1.0f };
	shape.normals[1] = b2Vec2{ 1.0f, 0.0f };
	shape.normals[2] = b2Vec2{ 0.0f, 1.0f };
	shape.normals[3] = b2Vec2{ -1.0f, 0.0f };
	shape.radius = 0.0f;
	shape.centroid = b2Vec2{ 0.0f, 0.0f };
	return shape;
}

b2Polygon b2MakeSquare( float halfWidth )
{
	return b2MakeBox( halfWidth, halfWidth );
}

b2Polygon b2MakeRoundedBox( float halfWidth, float halfHeight, float radius )
{
	b2Polygon shape = b2MakeBox( halfWidth, halfHeight );
	shape.radius = radius;
	return shape;
}

b2Polygon b2MakeOffsetRoundedBox( float halfWidth, float halfHeight, b2Vec2 center, b2Rot rotation, float radius )
{
	b2Transform xf = { center, rotation };
	b2Polygon shape = {};
	shape.count = 4;
	shape.vertices[0] = b2TransformPoint( xf, b2Vec2{ -halfWidth, -halfHeight } );
	shape.vertices[1] = b2TransformPoint( xf, b2Vec2{ halfWidth, -halfHeight } );
	shape.vertices[2] = b2TransformPoint( xf, b2Vec2{ halfWidth, halfHeight } );
	shape.vertices[3] = b2TransformPoint( xf, b2Vec2{ -halfWidth, halfHeight } );
	shape.normals[0] = b2RotateVector( xf.q, b2Vec2{ 0.0f, -1.0f } );
	shape.normals[1] = b2RotateVector( xf.q, b2Vec2{ 1.0f, 0.0f } );
	shape.normals[2] = b2RotateVector( xf.q, b2Vec2{ 0.0f, 1.0f } );
	shape.normals[3] = b2RotateVector( xf.q, b2Vec2{ -1.0f, 0.0f } );
	shape.radius = radius;
	shape.centroid = xf.p;
	return shape;
}

b2Polygon b2MakeOffsetBox( float halfWidth, float halfHeight, b2Vec2 center, b2Rot rotation )
{
	return b2MakeOffsetRoundedBox( halfWidth, halfHeight, center, rotation, 0.0f );
}

b2Polygon b2MakeOffsetRoundedPolygon( const b2Hull* hull, b2Vec2 position, b2Rot rotation, float radius )
{
	if ( hull->count < 3 )
	{
		return b2MakeSquare( 0.5f );
	}

	b2Transform transform = { position, rotation };
	b2Polygon shape = {};
	shape.count = hull->count;
	shape.radius = radius;
	for ( int i = 0; i < shape.count; ++i )
	{
		shape.vertices[i] = b2TransformPoint( transform, hull->points[i] );
	}
	FillNormals( &shape );
	shape.centroid = PolygonCentroid( shape.vertices, shape.count );
	return shape;
}

b2Polygon b2MakeOffsetPolygon( const b2Hull* hull, b2Vec2 position, b2Rot rotation )
{
	return b2MakeOffsetRoundedPolygon( hull, position, rotation, 0.0f );
}

b2Polygon b2MakePolygon( const b2Hull* hull, float radius )
{
	if ( hull->count < 3 )
	{
		return b2MakeSquare( 0.5f );
	}

	b2Polygon shape = {};
	shape.count = hull->count;
	shape.radius = radius;
	for ( int i = 0; i < shape.count; ++i )
	{
		shape.vertices[i] = hull->points[i];
	}
	FillNormals( &shape );
	shape.centroid = PolygonCentroid( shape.vertices, shape.count );
	return shape;
}

b2Polygon b2TransformPolygon( b2Transform transform, const b2Polygon* polygon )
{
	b2Polygon p = *polygon;
	for ( int i = 0; i < p.count; ++i )
	{
		p.vertices[i] = b2TransformPoint( transform, p.vertices[i] );
		p.normals[i] = b2RotateVector( transform.q, p.normals[i] );
	}
	p.centroid = b2TransformPoint( transform, p.centroid );
	return p;
}

// ---------------------------------------------------------------------------------------------
// Mass properties

b2MassData b2ComputeCircleMass( const b2Circle* shape, float density )
{
	float rr = shape->radius * shape->radius;
	b2MassData massData;
	massData.mass = density * B2_PI * rr;
	massData.center = shape->center;
	massData.rotationalInertia = massData.mass * 0.5f * rr;
	return massData;
}

b2MassData b2ComputeCapsuleMass( const b2Capsule* shape, float density )
{
	float radius = shape->radius;
	float rr = radius * radius;
	b2Vec2 p1 = shape->center1;
	b2Vec2 p2 = shape->center2;
	float length = b2Length( b2Sub( p2, p1 ) );
	float ll = length * length;

	float circleMass = density * ( B2_PI * radius * radius );
	float boxMass = density * ( 2.0f * radius * length );

	b2MassData massData;
	massData.mass = circleMass + boxMass;
	massData.center.x = 0.5f * ( p1.x + p2.x );
	massData.center.y = 0.5f * ( p1.y + p2.y );

	// Half-disc centroid offset 4r/(3pi); parallel axis theorem applied twice.
	float lc = 4.0f * radius / ( 3.0f * B2_PI );
	float h = 0.5f * length;
	float circleInertia = circleMass * ( 0.5f * rr + h * h + 2.0f * h * lc );
	float boxInertia = boxMass * ( 4.0f * rr + ll ) / 12.0f;
	massData.rotationalInertia = circleInertia + boxInertia;
	return massData;
}

b2MassData b2ComputePolygonMass( const b2Polygon* shape, float density )
{
	if ( shape->count == 1 )
	{
		b2Circle circle;
		circle.center = shape->vertices[0];
		circle.radius = shape->radius;
		return b2ComputeCircleMass( &circle, density );
	}

	if ( shape->count == 2 )
	{
		b2Capsule capsule;
		capsule.center1 = shape->vertices[0];
		capsule.center2 = shape->vertices[1];
		capsule.radius = shape->radius;
		return b2ComputeCapsuleMass( &capsule, density );
	}

	b2Vec2 vertices[B2_MAX_POLYGON_VERTICES] = {};
	int count = shape->count;
	float radius = shape->radius;

	if ( radius > 0.0f )
	{
		// Rounded polygons: push each vertex out along the averaged corner normal.
		float sqrt2 = 1.412f;
		for ( int i = 0; i < count; ++i )
		{
			int j = i == 0 ? count - 1 : i - 1;
			b2Vec2 n1 = shape->normals[j];
			b2Vec2 n2 = shape->normals[i];
			b2Vec2 mid = b2Normalize( b2Add( n1, n2 ) );
			vertices[i] = b2MulAdd( shape->vertices[i], sqrt2 * radius, mid );
		}
	}
	else
	{
		for ( int i = 0; i < count; ++i )
		{
			vertices[i] = shape->vertices[i];
		}
	}

	b2Vec2 center = { 0.0f, 0.0f };
	float area = 0.0f;
	float rotationalInertia = 0.0f;
	b2Vec2 r = vertices[0];
	const float inv3 = 1.0f / 3.0f;

	for ( int i = 1; i < count - 1; ++i )
	{
		b2Vec2 e1 = b2Sub( vertices[i], r );
		b2Vec2 e2 = b2Sub( vertices[i + 1], r );
		float D = b2Cross( e1, e2 );
		float triangleArea = 0.5f * D;
		area += triangleArea;
		center = b2MulAdd( center, triangleArea * inv3, b2Add( e1, e2 ) );

		float ex1 = e1.x, ey1 = e1.y;
		float ex2 = e2.x, ey2 = e2.y;
		float intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
		float inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
		rotationalInertia += ( 0.25f * inv3 * D ) * ( intx2 + inty2 );
	}

	b2MassData massData;
	massData.mass = density * area;
	float invArea = 1.0f / area;
	center.x *= invArea;
	center.y *= invArea;
	massData.center = b2Add( r, center );
	massData.rotationalInertia = density * rotationalInertia;
	massData.rotationalInertia -= massData.mass * b2Dot( center, center );
	return massData;
}

// ---------------------------------------------------------------------------------------------
// Bounding boxes: world points in float, extremes and inflation in double, one narrowing cast.

b2AABB b2ComputeCircleAABB( const b2Circle* shape, b2WorldTransform xf )
{
	b2Pos c = b2TransformWorldPoint( xf, shape->center );
	double r = (double)shape->radius;
	b2AABB aabb = { { (float)( c.x - r ), (float)( c.y - r ) }, { (float)( c.x + r ), (float)( c.y + r ) } };
	return aabb;
}

b2AABB b2ComputeCapsuleAABB( const b2Capsule* shape, b2WorldTransform xf )
{
	b2Pos v1 = b2TransformWorldPoint( xf, shape->center1 );
	b2Pos v2 = b2TransformWorldPoint( xf, shape->center2 );
	double r = (double)shape->radius;
	b2AABB aabb = {
		{ (float)( ( v1.x < v2.x ? v1.x : v2.x ) - r ), (float)( ( v1.y < v2.y ? v1.y : v2.y ) - r ) },
		{ (float)( ( v1.x > v2.x ? v1.x : v2.x ) + r ), (float)( ( v1.y > v2.y ? v1.y : v2.y ) + r ) },
	};
	return aabb;
}

b2AABB b2ComputePolygonAABB( const b2Polygon* shape, b2WorldTransform xf )
{
	b2Pos v = b2TransformWorldPoint( xf, shape->vertices[0] );
	double lx = v.x, ly = v.y, ux = v.x, uy = v.y;
	for ( int i = 1; i < shape->count; ++i )
	{
		v = b2TransformWorldPoint( xf, shape->vertices[i] );
		lx = v.x < lx ? v.x : lx;
		ly = v.y < ly ? v.y : ly;
		ux = v.x > ux ? v.x : ux;
		uy = v.y > uy ? v.y : uy;
	}
	double r = (double)shape->radius;
	b2AABB aabb = { { (float)( lx - r ), (float)( ly - r ) }, { (float)( ux + r ), (float)( uy + r ) } };
	return aabb;
}

b2AABB b2ComputeSegmentAABB( const b2Segment* shape, b2WorldTransform xf )
{
	b2Pos v1 = b2TransformWorldPoint( xf, shape->point1 );
	b2Pos v2 = b2TransformWorldPoint( xf, shape->point2 );
	b2AABB aabb = {
		{ (float)( v1.x < v2.x ? v1.x : v2.x ), (float)( v1.y < v2.y ? v1.y : v2.y ) },
		{ (float)( v1.x > v2.x ? v1.x : v2.x ), (float)( v1.y > v2.y ? v1.y : v2.y ) },
	};
	return aabb;
}

// ---------------------------------------------------------------------------------------------
// Quickhull with welding and collinear removal.

static b2Hull RecurseHull( b2Vec2 p1, b2Vec2 p2, const b2Vec2* ps, int count )
{
	b2Hull hull;
	hull.count = 0;
	if ( count == 0 )
	{
		return hull;
	}

	b2Vec2 e = b2Normalize( b2Sub( p2, p1 ) );
	b2Vec2 rightPoints[B2_MAX_POLYGON_VERTICES];
	int rightCount = 0;

	int bestIndex = 0;
	float bestDistance = b2Cross( b2Sub( ps[0], p1 ), e );
	if ( bestDistance > 0.0f )
	{
		rightPoints[rightCount++] = ps[0];
	}

	for ( int i = 1; i < count; ++i )
	{
		float distance = b2Cross( b2Sub( ps[i], p1 ), e );
		if ( distance > bestDistance )
		{
			bestIndex = i;
			bestDistance = distance;
		}
		if ( distance > 0.0f )
		{
			rightPoints[rightCount++] = ps[i];
		}
	}

	if ( bestDistance < 2.0f * B2_LINEAR_SLOP )
	{
		return hull;
	}

	b2Vec2 bestPoint = ps[bestIndex];
	b2Hull hull1 = RecurseHull( p1, bestPoint, rightPoints, rightCount );
	b2Hull hull2 = RecurseHull( bestPoint, p2, rightPoints, rightCount );

	for ( int i = 0; i < hull1.count; ++i )
	{
		hull.points[hull.count++] = hull1.points[i];
	}
	hull.points[hull.count++] = bestPoint;
	for ( int i = 0; i < hull2.count; ++i )
	{
		hull.points[hull.count++] = hull2.points[i];
	}
	return hull;
}

b2Hull b2ComputeHull( const b2Vec2* points, int count )
{
	b2Hull hull = {};
	hull.count = 0;

	if ( count < 3 || count > B2_MAX_POLYGON_VERTICES )
	{
		return hull;
	}

	b2AABB aabb = { { FLT_MAX, FLT_MAX }, { -FLT_MAX, -FLT_MAX } };
	b2Vec2 ps[B2_MAX_POLYGON_VERTICES];
	int n = 0;
	const float linearSlop = B2_LINEAR_SLOP;
	const float tolSqr = 16.0f * linearSlop * linearSlop;
	for ( int i = 0; i < count; ++i )
	{
		aabb.lowerBound = b2Min( aabb.lowerBound, points[i] );
		aabb.upperBound = b2Max( aabb.upperBound, points[i] );
		b2Vec2 vi = points[i];
		bool unique = true;
		for ( int j = 0; j < i; ++j )
		{
			if ( b2DistanceSquared( vi, points[j] ) < tolSqr )
			{
				unique = false;
				break;
			}
		}
		if ( unique )
		{
			ps[n++] = vi;
		}
	}

	if ( n < 3 )
	{
		return hull;
	}

	b2Vec2 c = b2AABB_Center( aabb );
	int f1 = 0;
	float dsq1 = b2DistanceSquared( c, ps[f1] );
	for ( int i = 1; i < n; ++i )
	{
		float dsq = b2DistanceSquared( c, ps[i] );
		if ( dsq > dsq1 )
		{
			f1 = i;
			dsq1 = dsq;
		}
	}

	b2Vec2 p1 = ps[f1];
	ps[f1] = ps[n - 1];
	n = n - 1;

	int f2 = 0;
	float dsq2 = b2DistanceSquared( p1, ps[f2] );
	for ( int i = 1; i < n; ++i )
	{
		float dsq = b2DistanceSquared( p1, ps[i] );
		if ( dsq > dsq2 )
		{
			f2 = i;
			dsq2 = dsq;
		}
	}

	b2Vec2 p2 = ps[f2];
	ps[f2] = ps[n - 1];
	n = n - 1;

	b2Vec2 rightPoints[B2_MAX_POLYGON_VERTICES - 2];
	int rightCount = 0;
	b2Vec2 leftPoints[B2_MAX_POLYGON_VERTICES - 2];
	int leftCount = 0;

	b2Vec2 e = b2Normalize( b2Sub( p2, p1 ) );
	for ( int i = 0; i < n; ++i )
	{
		float d = b2Cross( b2Sub( ps[i], p1 ), e );
		if ( d >= 2.0f * linearSlop )
		{
			rightPoints[rightCount++] = ps[i];
		}
		else if ( d <= -2.0f * linearSlop )
		{
			leftPoints[leftCount++] = ps[i];
		}
	}

	b2Hull hull1 = RecurseHull( p1, p2, rightPoints, rightCount );
	b2Hull hull2 = RecurseHull( p2, p1, leftPoints, leftCount );

	if ( hull1.count == 0 && hull2.count == 0 )
	{
		return hull;
	}

	hull.points[hull.count++] = p1;
	for ( int i = 0; i < hull1.count; ++i )
	{
		hull.points[hull.count++] = hull1.points[i];
	}
	hull.points[hull.count++] = p2;
	for ( int i = 0; i < hull2.count; ++i )
	{
		hull.points[hull.count++] = hull2.points[i];
	}

	bool searching = true;
	while ( searching && hull.count > 2 )
	{
		searching = false;
		for ( int i = 0; i < hull.count; ++i )
		{
			int i1 = i;
			int i2 = ( i + 1 ) % hull.count;
			int i3 = ( i + 2 ) % hull.count;
			b2Vec2 s1 = hull.points[i1];
			b2Vec2 s2 = hull.points[i2];
			b2Vec2 s3 = hull.points[i3];
			b2Vec2 r = b2Normalize( b2Sub( s3, s1 ) );
			float distance = b2Cross( b2Sub( s2, s1 ), r );
			if ( distance <= 2.0f * linearSlop )
			{
				for ( int j = i2; j < hull.count - 1; ++j )
				{
					hull.points[j] = hull.points[j + 1];
				}
				hull.count -= 1;
				searching = true;
				break;
			}
		}
	}

	if ( hull.count < 3 )
	{
		hull.count = 0;
	}
	return hull;
}

bool b2ValidateHull( const b2Hull* hull )
{
	if ( hull->count < 3 || B2_MAX_POLYGON_VERTICES < hull->count )
	{
		return false;
	}

	for ( int i = 0; i < hull->count; ++i )
	{
		int i1 = i;
		int i2 = i < hull->count - 1 ? i1 + 1 : 0;
		b2Vec2 p = hull->points[i1];
		b2Vec2 e = b2Normalize( b2Sub( hull->points[i2], p ) );
		for ( int j = 0; j < hull->count; ++j )
		{
			if ( j == i1 || j == i2 )
			{
				continue;
			}
			if ( b2Cross( b2Sub( hull->points[j], p ), e ) >= 0.0f )
			{
				return false;
			}
		}
	}

	for ( int i = 0; i < hull->count; ++i )
	{
		b2Vec2 p1 = hull->points[i];
		b2Vec2 p2 = hull->points[( i + 1 ) % hull->count];
		b2Vec2 p3 = hull->points[( i + 2 ) % hull->count];
		b2Vec2 e = b2Normalize( b2Sub( p3, p1 ) );
		if ( b2Cross( b2Sub( p2, p1 ), e ) <= B2_LINEAR_SLOP )
		{
			return false;
		}
	}
	return true;
}

} // extern "C"
