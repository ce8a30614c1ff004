// Scene driver for the Box2D step path. Public Box2D C API only.
//
// The pyramid follows the reference benchmark scene (shared/benchmarks.c:88-131), the falling
// hinges follow the reference determinism scene (shared/determinism.c:11-134) and the hello
// scene follows test/test_world.c:19-98. The random generator is XorShift32 seeded with 12345,
// the generator the reference scene library uses (shared/utils.h:9-36).

#include "scenes.h"

#include "box2d/box2d.h"

#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define B2S_MAX_HANDLES 4096

#define B2S_EDIT_LINKS 5

typedef struct b2sScene
{
	b2WorldId worldId;
	b2BodyId* bodies;
	int bodyCount;
	int inUse;

	// B2S_EDITS: the edit script's state
	int kind;
	int stepIndex;
	int bodyCapacity;
	uint32_t probeHash;
	b2BodyId links[B2S_EDIT_LINKS];
	b2JointId hinges[B2S_EDIT_LINKS];
	b2BodyId kinematicId;

	// B2S_SETTERS: the chain links and the body with two shapes
	b2BodyId setterLinks[3];
	b2ShapeId extraShape;

	// B2S_JOINT_ZOO with event thresholds: the joints in ascending id
	b2JointId* zooJoints;
	int zooJointCount;

	// B2S_SENSORS: the sensor shapes, the body carrying one, a visitor destroyed mid-run
	b2ShapeId zoneShape;
	b2ShapeId stripShape;
	b2ShapeId sweepShape;
	b2ShapeId carriedShape;
	b2BodyId carrierId;
	b2ShapeId doomedShape;

	// B2S_PRESOLVE: the one-way platform and the callback's tally of the last step
	// (B2S_BULLETS: the pre-continuous callback's tally)
	b2ShapeId platformShape;
	uint32_t preSolveCalls;
	uint32_t preSolveSum;

	// B2S_BULLETS: the dynamic sensor and the ghost wall
	b2ShapeId ringShape;
	b2ShapeId ghostShape;
} b2sScene;

static b2sScene s_scenes[B2S_MAX_HANDLES];

static uint32_t s_rngState = 12345u;

static uint32_t b2sNextRandom( void )
{
	uint32_t x = s_rngState;
	x ^= x << 13;
	x ^= x >> 17;
	x ^= x << 5;
	s_rngState = x;
	return x;
}

// Uniform in [lo, hi] with 15-bit resolution.
static float b2sRandomRange( float lo, float hi )
{
	float r = (float)( b2sNextRandom() % 32768u ) / 32767.0f;
	return lo + r * ( hi - lo );
}

static void b2sPushBody( b2sScene* scene, int* capacity, b2BodyId id )
{
	if ( scene->bodyCount == *capacity )
	{
		*capacity = *capacity == 0 ? 64 : 2 * *capacity;
		scene->bodies = (b2BodyId*)realloc( scene->bodies, (size_t)*capacity * sizeof( b2BodyId ) );
	}
	scene->bodies[scene->bodyCount++] = id;
}

static void b2sBuildPyramid( b2sScene* scene, int rows, float groundHalfWidth, int enableSleep )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );

	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Pos ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( groundHalfWidth, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}

	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;

	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.density = 1.0f;

	float a = 0.5f;
	b2Polygon box = b2MakeSquare( a );
	float shift = 1.0f * a;
	int capacity = 0;

	for ( int i = 0; i < rows; ++i )
	{
		float y = ( 2.0f * i + 1.0f ) * shift;
		for ( int j = i; j < rows; ++j )
		{
			float x = ( i + 1.0f ) * shift + 2.0f * ( j - i ) * shift - a * rows;
			bodyDef.position = ( b2Pos ){ x, y };
			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			b2CreatePolygonShape( bodyId, &shapeDef, &box );
			b2sPushBody( scene, &capacity, bodyId );
		}
	}
}

static void b2sBuildFallingHinges( b2sScene* scene )
{
	b2WorldId worldId = scene->worldId;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Pos ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( 40.0f, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}

	int columnCount = 4;
	int rowCount = 20;
	float h = 0.25f;
	b2Polygon box = b2MakeSquare( h );
	b2ShapeDef shapeDef = b2DefaultShapeDef();

	b2RevoluteJointDef jointDef = b2DefaultRevoluteJointDef();
	jointDef.enableLimit = true;
	jointDef.lowerAngle = -0.1f * B2_PI;
	jointDef.upperAngle = 0.2f * B2_PI;
	jointDef.enableSpring = true;
	jointDef.hertz = 1.0f;
	jointDef.dampingRatio = 1.0f;
	jointDef.enableMotor = true;
	jointDef.maxMotorTorque = 0.25f;
	jointDef.base.localFrameA.p = ( b2Vec2 ){ -h, h };
	jointDef.base.localFrameB.p = ( b2Vec2 ){ -h, -h };
	jointDef.base.constraintHertz = 60.0f;
	jointDef.base.constraintDampingRatio = 0.0f;
	jointDef.base.drawScale = 0.5f;

	float offset = 0.4f * h;
	float dx = 10.0f * h;
	float xBase = -0.5f * dx * ( columnCount - 1.0f );
	int capacity = 0;

	for ( int j = 0; j < columnCount; ++j )
	{
		float x = xBase + j * dx;
		b2BodyId prevBodyId = b2_nullBodyId;
		for ( int i = 0; i < rowCount; ++i )
		{
			b2BodyDef bodyDef = b2DefaultBodyDef();
			bodyDef.type = b2_dynamicBody;
			bodyDef.position.x = x + offset * i;
			bodyDef.position.y = h + 2.0f * h * i;
			float angle = ( i & 1 ) == 0 ? -0.1f : 0.1f;
			bodyDef.rotation = b2MakeRot( angle );

			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			if ( ( i & 1 ) == 0 )
			{
				prevBodyId = bodyId;
			}
			else
			{
				jointDef.base.bodyIdA = prevBodyId;
				jointDef.base.bodyIdB = bodyId;
				b2CreateRevoluteJoint( worldId, &jointDef );
				prevBodyId = b2_nullBodyId;
			}

			b2CreatePolygonShape( bodyId, &shapeDef, &box );
			b2sPushBody( scene, &capacity, bodyId );
		}
	}
}

static void b2sBuildHello( b2sScene* scene )
{
	b2WorldId worldId = scene->worldId;
	b2BodyDef groundBodyDef = b2DefaultBodyDef();
	groundBodyDef.position = ( b2Pos ){ 0.0f, -10.0f };
	b2BodyId groundId = b2CreateBody( worldId, &groundBodyDef );
	b2Polygon groundBox = b2MakeBox( 50.0f, 10.0f );
	b2ShapeDef groundShapeDef = b2DefaultShapeDef();
	b2CreatePolygonShape( groundId, &groundShapeDef, &groundBox );

	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	bodyDef.position = ( b2Pos ){ 0.0f, 4.0f };
	b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
	b2Polygon dynamicBox = b2MakeBox( 1.0f, 1.0f );
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.density = 1.0f;
	shapeDef.material.friction = 0.3f;
	b2CreatePolygonShape( bodyId, &shapeDef, &dynamicBox );

	int capacity = 0;
	b2sPushBody( scene, &capacity, bodyId );
}

static void b2sBuildMixedPile( b2sScene* scene, int count, int enableSleep, uint32_t seed )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );
	s_rngState = seed != 0 ? seed : 12345u;

	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Pos ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( 40.0f, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}

	int columns = 20;
	int capacity = 0;
	for ( int k = 0; k < count; ++k )
	{
		int col = k % columns;
		int row = k / columns;
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ -19.0f + 2.0f * col + b2sRandomRange( -0.2f, 0.2f ), 1.5f + 2.0f * row };
		bodyDef.rotation = b2MakeRot( b2sRandomRange( -3.0f, 3.0f ) );
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );

		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.density = b2sRandomRange( 0.5f, 2.0f );

		int kind = k & 3;
		if ( kind == 0 )
		{
			b2Circle circle = { { 0.0f, 0.0f }, b2sRandomRange( 0.2f, 0.5f ) };
			b2CreateCircleShape( bodyId, &shapeDef, &circle );
		}
		else if ( kind == 1 )
		{
			float halfLength = b2sRandomRange( 0.2f, 0.6f );
			b2Capsule capsule = { { -halfLength, 0.0f }, { halfLength, 0.0f }, b2sRandomRange( 0.1f, 0.25f ) };
			b2CreateCapsuleShape( bodyId, &shapeDef, &capsule );
		}
		else if ( kind == 2 )
		{
			b2Vec2 points[8];
			int n = 3 + (int)( b2sNextRandom() % 6u );
			for ( int i = 0; i < n; ++i )
			{
				float angle = b2sRandomRange( -3.14f, 3.14f );
				float radius = b2sRandomRange( 0.3f, 0.9f );
				b2Rot q = b2MakeRot( angle );
				points[i] = ( b2Vec2 ){ radius * q.c, radius * q.s };
			}
			b2Hull hull = b2ComputeHull( points, n );
			b2Polygon polygon = hull.count >= 3 ? b2MakePolygon( &hull, 0.0f ) : b2MakeSquare( 0.4f );
			b2CreatePolygonShape( bodyId, &shapeDef, &polygon );
		}
		else
		{
			b2Polygon box = b2MakeRoundedBox( b2sRandomRange( 0.2f, 0.5f ), b2sRandomRange( 0.2f, 0.5f ), 0.05f );
			b2CreatePolygonShape( bodyId, &shapeDef, &box );
		}

		b2sPushBody( scene, &capacity, bodyId );
	}
}

static void b2sAddMixedShape( b2BodyId bodyId, int kind )
{
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.density = b2sRandomRange( 0.5f, 2.0f );

	if ( kind == 0 )
	{
		b2Circle circle = { { 0.0f, 0.0f }, b2sRandomRange( 0.2f, 0.5f ) };
		b2CreateCircleShape( bodyId, &shapeDef, &circle );
	}
	else if ( kind == 1 )
	{
		float halfLength = b2sRandomRange( 0.2f, 0.6f );
		b2Capsule capsule = { { -halfLength, 0.0f }, { halfLength, 0.0f }, b2sRandomRange( 0.1f, 0.25f ) };
		b2CreateCapsuleShape( bodyId, &shapeDef, &capsule );
	}
	else if ( kind == 2 )
	{
		b2Vec2 points[8];
		int n = 3 + (int)( b2sNextRandom() % 6u );
		for ( int i = 0; i < n; ++i )
		{
			float angle = b2sRandomRange( -3.14f, 3.14f );
			float radius = b2sRandomRange( 0.0f, 0.5f );
			b2Rot q = b2MakeRot( angle );
			points[i] = ( b2Vec2 ){ radius * q.c, radius * q.s };
		}
		b2Hull hull = b2ComputeHull( points, n );
		b2Polygon polygon = hull.count >= 3 ? b2MakePolygon( &hull, 0.0f ) : b2MakeSquare( 0.25f );
		b2CreatePolygonShape( bodyId, &shapeDef, &polygon );
	}
	else
	{
		b2Polygon box = b2MakeRoundedBox( b2sRandomRange( 0.2f, 0.5f ), b2sRandomRange( 0.2f, 0.5f ), 0.05f );
		b2CreatePolygonShape( bodyId, &shapeDef, &box );
	}
}

static void b2sBuildRain( b2sScene* scene, int count, int columns, int enableSleep, uint32_t seed, float width )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );
	s_rngState = seed != 0 ? seed : 12345u;

	// container: right wall down, floor right to left, left wall up (interior on the right side)
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );

		float halfWidth = 0.5f * width + 20.0f;
		float height = 100.0f;
		int floorSegments = (int)( 2.0f * halfWidth / 2.0f );
		int wallSegments = 10;
		int pointCount = 2 + ( wallSegments + 1 ) + ( floorSegments - 1 ) + ( wallSegments + 1 );
		b2Vec2* points = (b2Vec2*)malloc( (size_t)pointCount * sizeof( b2Vec2 ) );
		int n = 0;
		points[n++] = ( b2Vec2 ){ halfWidth, height + 1.0f };
		for ( int i = 0; i <= wallSegments; ++i )
		{
			points[n++] = ( b2Vec2 ){ halfWidth, height - height * i / wallSegments };
		}
		for ( int i = 1; i < floorSegments; ++i )
		{
			points[n++] = ( b2Vec2 ){ halfWidth - 2.0f * halfWidth * i / floorSegments, 0.0f };
		}
		for ( int i = 0; i <= wallSegments; ++i )
		{
			points[n++] = ( b2Vec2 ){ -halfWidth, height * i / wallSegments };
		}
		points[n++] = ( b2Vec2 ){ -halfWidth, height + 1.0f };

		b2SurfaceMaterial material = b2DefaultSurfaceMaterial();
		b2ChainDef chainDef = b2DefaultChainDef();
		chainDef.points = points;
		chainDef.count = n;
		chainDef.materials = &material;
		chainDef.materialCount = 1;
		chainDef.isLoop = false;
		b2CreateChain( groundId, &chainDef );
		free( points );
	}

	float spacing = width / (float)columns;
	int capacity = 0;
	for ( int k = 0; k < count; ++k )
	{
		int col = k % columns;
		int row = k / columns;
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ -0.5f * width + ( col + 0.5f ) * spacing, 110.0f + spacing * row };
		bodyDef.rotation = b2MakeRot( b2sRandomRange( -3.0f, 3.0f ) );
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2sAddMixedShape( bodyId, k & 3 );
		b2sPushBody( scene, &capacity, bodyId );
	}
}

static void b2sBuildBatchWorld( b2sScene* scene, int target, int kind, uint32_t seed )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, false );
	s_rngState = seed != 0 ? seed : 12345u;
	int capacity = 0;

	b2BodyId groundId;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Pos ){ 0.0f, -1.0f };
		groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( 40.0f, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}

	if ( kind == 0 )
	{
		int rows = 1;
		while ( ( rows + 1 ) * ( rows + 2 ) / 2 <= target )
		{
			rows += 1;
		}
		b2Polygon box = b2MakeSquare( 0.5f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.density = 1.0f;
		for ( int i = 0; i < rows; ++i )
		{
			float y = 0.5f + i * 1.0f;
			float x0 = -0.5f * ( rows - 1 - i );
			for ( int j = i; j < rows; ++j )
			{
				b2BodyDef bodyDef = b2DefaultBodyDef();
				bodyDef.type = b2_dynamicBody;
				bodyDef.position = ( b2Pos ){ x0 + ( j - i ) * 1.0f, y };
				b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
				b2CreatePolygonShape( bodyId, &shapeDef, &box );
				b2sPushBody( scene, &capacity, bodyId );
			}
		}
	}
	else if ( kind == 1 )
	{
		int columns = 1;
		while ( columns * columns < target )
		{
			columns += 1;
		}
		for ( int k = 0; k < target; ++k )
		{
			b2BodyDef bodyDef = b2DefaultBodyDef();
			bodyDef.type = b2_dynamicBody;
			bodyDef.position = ( b2Pos ){ -0.6f * columns + 1.2f * ( k % columns ) + b2sRandomRange( -0.1f, 0.1f ),
										  1.0f + 1.2f * ( k / columns ) };
			bodyDef.rotation = b2MakeRot( b2sRandomRange( -3.0f, 3.0f ) );
			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			b2ShapeDef shapeDef = b2DefaultShapeDef();
			shapeDef.density = b2sRandomRange( 0.5f, 2.0f );
			if ( b2sNextRandom() & 1u )
			{
				b2Polygon box = b2MakeBox( b2sRandomRange( 0.2f, 0.5f ), b2sRandomRange( 0.2f, 0.5f ) );
				b2CreatePolygonShape( bodyId, &shapeDef, &box );
			}
			else
			{
				b2Circle circle = { { 0.0f, 0.0f }, b2sRandomRange( 0.2f, 0.5f ) };
				b2CreateCircleShape( bodyId, &shapeDef, &circle );
			}
			b2sPushBody( scene, &capacity, bodyId );
		}
	}
	else
	{
		int chainCount = 1 + (int)( b2sNextRandom() % 8u );
		int links = target / chainCount;
		links = links < 5 ? 5 : ( links > 20 ? 20 : links );
		float halfLength = 0.25f;
		b2Capsule capsule = { { -halfLength, 0.0f }, { halfLength, 0.0f }, 0.125f };
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.density = 20.0f;
		b2RevoluteJointDef jointDef = b2DefaultRevoluteJointDef();
		for ( int c = 0; c < chainCount; ++c )
		{
			float x = -2.0f * chainCount + 4.0f * c;
			float y = 2.0f * halfLength * links + 2.0f;
			// bodies are unrotated, so a local point is the world point minus the body position
			b2BodyId prevId = groundId;
			b2Vec2 prevPosition = { 0.0f, -1.0f };
			for ( int i = 0; i < links; ++i )
			{
				b2BodyDef bodyDef = b2DefaultBodyDef();
				bodyDef.type = b2_dynamicBody;
				bodyDef.position = ( b2Pos ){ x + ( 1.0f + 2.0f * i ) * halfLength, y };
				b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
				b2CreateCapsuleShape( bodyId, &shapeDef, &capsule );

				b2Vec2 pivot = { x + ( 2.0f * i ) * halfLength, y };
				jointDef.base.bodyIdA = prevId;
				jointDef.base.bodyIdB = bodyId;
				jointDef.base.localFrameA.p = ( b2Vec2 ){ pivot.x - prevPosition.x, pivot.y - prevPosition.y };
				jointDef.base.localFrameB.p = ( b2Vec2 ){ pivot.x - bodyDef.position.x, pivot.y - bodyDef.position.y };
				b2CreateRevoluteJoint( worldId, &jointDef );
				prevId = bodyId;
				prevPosition = ( b2Vec2 ){ bodyDef.position.x, bodyDef.position.y };
				b2sPushBody( scene, &capacity, bodyId );
			}
		}
	}
}

// Pyramid of `rows` rows of 1 m boxes with a 1 m pusher box on the ground to its left, 0.5 m
// from the pyramid's base, launched with the given velocity and spin. The pusher's friction is a
// parameter; every other shape keeps the default friction 0.6. Bodies in creation order: the
// pyramid row by row (the top box last), then the pusher.
static void b2sBuildPush( b2sScene* scene, int rows, float vx, float vy, float omega, float friction )
{
	b2sBuildPyramid( scene, rows, 40.0f, 0 );

	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	bodyDef.position = ( b2Pos ){ -0.5f * rows - 1.0f, 0.5f };
	bodyDef.linearVelocity = ( b2Vec2 ){ vx, vy };
	bodyDef.angularVelocity = omega;
	b2BodyId pusherId = b2CreateBody( scene->worldId, &bodyDef );

	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.density = 1.0f;
	shapeDef.material.friction = friction;
	b2Polygon box = b2MakeSquare( 0.5f );
	b2CreatePolygonShape( pusherId, &shapeDef, &box );

	int capacity = scene->bodyCount;
	b2sPushBody( scene, &capacity, pusherId );
}

static b2BodyId b2sDynamicBox( b2sScene* scene, int* capacity, b2Vec2 position, float angle, float hx, float hy, float friction )
{
	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	bodyDef.position = ( b2Pos ){ position.x, position.y };
	bodyDef.rotation = b2MakeRot( angle );
	b2BodyId bodyId = b2CreateBody( scene->worldId, &bodyDef );
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.material.friction = friction;
	b2Polygon box = b2MakeBox( hx, hy );
	b2CreatePolygonShape( bodyId, &shapeDef, &box );
	b2sPushBody( scene, capacity, bodyId );
	return bodyId;
}

static int b2sCompareJointIds( const void* a, const void* b )
{
	int ia = ( (const b2JointId*)a )->index1;
	int ib = ( (const b2JointId*)b )->index1;
	return ( ia > ib ) - ( ia < ib );
}

// Event thresholds of joint k in round r (r = 0 at creation, 1 after the mid-run change): every
// awake step, a force level, a torque level, or never.
static void b2sSetZooThresholds( b2JointId jointId, int k, int round )
{
	static const float force[4] = { 0.0f, 40.0f, FLT_MAX, FLT_MAX };
	static const float torque[4] = { 0.0f, FLT_MAX, 2.0f, FLT_MAX };
	int v = ( k + round ) & 3;
	b2Joint_SetForceThreshold( jointId, force[v] );
	b2Joint_SetTorqueThreshold( jointId, torque[v] );
}

// Collect the zoo's joints through the bodies' joint lists, give each user data k + 1 and
// thresholds by its rank k in ascending joint id.
static void b2sZooThresholds( b2sScene* scene )
{
	int capacity = 0;
	for ( int i = 0; i < scene->bodyCount; ++i )
	{
		capacity += b2Body_GetJointCount( scene->bodies[i] );
	}
	scene->zooJoints = (b2JointId*)malloc( ( capacity > 0 ? capacity : 1 ) * sizeof( b2JointId ) );
	int n = 0;
	for ( int i = 0; i < scene->bodyCount; ++i )
	{
		n += b2Body_GetJoints( scene->bodies[i], scene->zooJoints + n, capacity - n );
	}
	qsort( scene->zooJoints, n, sizeof( b2JointId ), b2sCompareJointIds );
	int unique = 0;
	for ( int i = 0; i < n; ++i )
	{
		if ( unique == 0 || scene->zooJoints[unique - 1].index1 != scene->zooJoints[i].index1 )
		{
			scene->zooJoints[unique++] = scene->zooJoints[i];
		}
	}
	scene->zooJointCount = unique;
	for ( int k = 0; k < unique; ++k )
	{
		b2Joint_SetUserData( scene->zooJoints[k], (void*)(intptr_t)( k + 1 ) );
		b2sSetZooThresholds( scene->zooJoints[k], k, 0 );
	}
}

static void b2sBuildJointZoo( b2sScene* scene, int enableSleep )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );
	int capacity = 0;

	b2BodyId groundId;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Pos ){ 0.0f, -1.0f };
		groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( 80.0f, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}

	// weld cantilevers from a static post: rigid, soft angular, soft linear and angular
	for ( int variant = 0; variant < 3; ++variant )
	{
		float x0 = -70.0f + 8.0f * variant;
		b2BodyDef postDef = b2DefaultBodyDef();
		postDef.position = ( b2Pos ){ x0, 6.0f };
		b2BodyId postId = b2CreateBody( worldId, &postDef );
		b2Polygon postBox = b2MakeBox( 0.25f, 0.25f );
		b2ShapeDef postShape = b2DefaultShapeDef();
		b2CreatePolygonShape( postId, &postShape, &postBox );

		b2BodyId prev = postId;
		for ( int i = 0; i < 6; ++i )
		{
			b2BodyId link = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0 + 0.75f + 1.0f * i, 6.0f }, 0.0f, 0.5f, 0.125f, 0.6f );
			b2WeldJointDef wd = b2DefaultWeldJointDef();
			wd.base.bodyIdA = prev;
			wd.base.bodyIdB = link;
			wd.base.localFrameA.p = ( b2Vec2 ){ i == 0 ? 0.25f : 0.5f, 0.0f };
			wd.base.localFrameB.p = ( b2Vec2 ){ -0.5f, 0.0f };
			wd.linearHertz = variant == 2 ? 5.0f : 0.0f;
			wd.linearDampingRatio = 0.7f;
			wd.angularHertz = variant >= 1 ? 3.0f : 0.0f;
			wd.angularDampingRatio = 0.5f;
			b2CreateWeldJoint( worldId, &wd );
			prev = link;
		}
	}

	// prismatic sliders on the ground: limit only, motor only, spring and limit, all three, tilted
	for ( int variant = 0; variant < 4; ++variant )
	{
		float x0 = -40.0f + 6.0f * variant;
		b2BodyId slider = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0, 3.0f }, 0.0f, 0.5f, 0.5f, 0.6f );
		b2PrismaticJointDef pd = b2DefaultPrismaticJointDef();
		pd.base.bodyIdA = groundId;
		pd.base.bodyIdB = slider;
		pd.base.localFrameA.p = ( b2Vec2 ){ x0, 4.0f };
		pd.base.localFrameA.q = b2MakeRot( 0.25f * B2_PI + 0.2f * variant );
		pd.base.collideConnected = variant == 3;
		pd.enableLimit = variant != 1;
		pd.lowerTranslation = -2.0f;
		pd.upperTranslation = 1.5f;
		pd.enableMotor = variant == 1 || variant == 3;
		pd.motorSpeed = variant == 3 ? -1.0f : 2.0f;
		pd.maxMotorForce = 40.0f;
		pd.enableSpring = variant >= 2;
		pd.hertz = 2.0f;
		pd.dampingRatio = 0.4f;
		pd.targetTranslation = 0.5f;
		b2CreatePrismaticJoint( worldId, &pd );

		// a box dropped on the slider
		b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0 + 0.2f, 6.0f + variant }, 0.3f, 0.3f, 0.3f, 0.6f );
	}

	// wheel cars driving on the ground
	for ( int variant = 0; variant < 3; ++variant )
	{
		float x0 = -10.0f + 10.0f * variant;
		b2BodyId chassis = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0, 1.4f }, 0.0f, 1.5f, 0.3f, 0.6f );
		for ( int side = 0; side < 2; ++side )
		{
			b2BodyDef wheelDef = b2DefaultBodyDef();
			wheelDef.type = b2_dynamicBody;
			wheelDef.position = ( b2Pos ){ x0 + ( side == 0 ? -1.0f : 1.0f ), 0.45f };
			b2BodyId wheel = b2CreateBody( worldId, &wheelDef );
			b2ShapeDef wheelShape = b2DefaultShapeDef();
			wheelShape.material.friction = 0.9f;
			b2Circle circle = { { 0.0f, 0.0f }, 0.4f };
			b2CreateCircleShape( wheel, &wheelShape, &circle );
			b2sPushBody( scene, &capacity, wheel );

			b2WheelJointDef jd = b2DefaultWheelJointDef();
			jd.base.bodyIdA = chassis;
			jd.base.bodyIdB = wheel;
			jd.base.localFrameA.p = ( b2Vec2 ){ side == 0 ? -1.0f : 1.0f, -0.95f };
			jd.base.localFrameA.q = b2MakeRot( 0.5f * B2_PI );
			jd.enableSpring = variant != 1;
			jd.hertz = 4.0f;
			jd.dampingRatio = 0.7f;
			jd.enableLimit = variant >= 1;
			jd.lowerTranslation = -0.25f;
			jd.upperTranslation = 0.25f;
			jd.enableMotor = side == 0;
			jd.motorSpeed = variant == 2 ? 4.0f : -3.0f;
			jd.maxMotorTorque = 15.0f;
			b2CreateWheelJoint( worldId, &jd );
		}
	}

	// distance ropes from static anchors: rigid, spring, spring with limits, spring with motor
	for ( int variant = 0; variant < 4; ++variant )
	{
		float x0 = 25.0f + 6.0f * variant;
		b2BodyId prev = groundId;
		b2Vec2 prevAnchor = { x0, 9.0f };
		for ( int i = 0; i < 5; ++i )
		{
			b2BodyId bob = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0 + 0.9f * ( i + 1 ), 9.0f - 0.3f * ( i + 1 ) }, 0.1f * i, 0.2f, 0.2f, 0.6f );
			b2DistanceJointDef dd = b2DefaultDistanceJointDef();
			dd.base.bodyIdA = prev;
			dd.base.bodyIdB = bob;
			dd.base.localFrameA.p = prevAnchor;
			dd.base.localFrameB.p = ( b2Vec2 ){ -0.15f, 0.0f };
			dd.length = 1.0f;
			dd.enableSpring = variant >= 1;
			dd.hertz = 3.0f;
			dd.dampingRatio = 0.3f;
			dd.enableLimit = variant == 2;
			dd.minLength = 0.6f;
			dd.maxLength = 1.4f;
			dd.enableMotor = variant == 3;
			dd.maxMotorForce = 10.0f;
			dd.motorSpeed = 0.5f;
			if ( variant == 3 )
			{
				dd.lowerSpringForce = -20.0f;
				dd.upperSpringForce = 30.0f;
			}
			b2CreateDistanceJoint( worldId, &dd );
			prev = bob;
			prevAnchor = ( b2Vec2 ){ 0.15f, 0.0f };
		}
	}

	// motor-joint followers relative to the ground: velocity targets, springs, both
	for ( int variant = 0; variant < 3; ++variant )
	{
		float x0 = 55.0f + 5.0f * variant;
		b2BodyId body = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0, 2.0f + variant }, 0.2f * variant, 0.4f, 0.4f, 0.6f );
		b2MotorJointDef md = b2DefaultMotorJointDef();
		md.base.bodyIdA = groundId;
		md.base.bodyIdB = body;
		md.base.localFrameA.p = ( b2Vec2 ){ x0, 4.0f };
		md.base.collideConnected = true;
		if ( variant != 1 )
		{
			md.linearVelocity = ( b2Vec2 ){ 0.5f, 0.2f };
			md.maxVelocityForce = 20.0f;
			md.angularVelocity = 1.0f;
			md.maxVelocityTorque = 5.0f;
		}
		if ( variant != 0 )
		{
			md.linearHertz = 2.0f;
			md.linearDampingRatio = 0.5f;
			md.maxSpringForce = 50.0f;
			md.angularHertz = 1.0f;
			md.angularDampingRatio = 0.3f;
			md.maxSpringTorque = 10.0f;
		}
		b2CreateMotorJoint( worldId, &md );
	}

	// filter joints: overlapping box pairs that fall without colliding with each other
	for ( int variant = 0; variant < 3; ++variant )
	{
		float x0 = 72.0f + 2.0f * variant;
		b2BodyId a = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0, 2.0f + variant }, 0.0f, 0.5f, 0.5f, 0.6f );
		b2BodyId b = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0 + 0.3f, 2.2f + variant }, 0.4f, 0.5f, 0.5f, 0.6f );
		b2FilterJointDef fd = b2DefaultFilterJointDef();
		fd.base.bodyIdA = a;
		fd.base.bodyIdB = b;
		b2CreateFilterJoint( worldId, &fd );
	}

	// mover joints relative to the ground: driven, and with zero force limits (impulse cleared)
	for ( int variant = 0; variant < 2; ++variant )
	{
		float x0 = -78.0f + 2.5f * variant;
		b2BodyId body = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0, 1.0f + variant }, 0.3f, 0.4f, 0.4f, 0.6f );
		b2MoverJointDef md = b2DefaultMoverJointDef();
		md.base.bodyIdA = groundId;
		md.base.bodyIdB = body;
		md.base.collideConnected = true;
		md.linearVelocity = ( b2Vec2 ){ 0.8f, 0.6f };
		md.maxVelocityForce = variant == 0 ? ( b2Vec2 ){ 30.0f, 15.0f } : ( b2Vec2 ){ 0.0f, 0.0f };
		b2CreateMoverJoint( worldId, &md );
	}

	// pogo joints holding boxes above the ground: sprung with tension and compression limits,
	// and spring-less with a nonzero initial impulse (cleared by the solver)
	for ( int variant = 0; variant < 2; ++variant )
	{
		float x0 = -73.0f + 2.5f * variant;
		b2BodyId body = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ x0, 2.0f }, 0.0f, 0.3f, 0.5f, 0.6f );
		b2PogoJointDef pd = b2DefaultPogoJointDef();
		pd.base.bodyIdA = groundId;
		pd.base.bodyIdB = body;
		pd.base.localFrameA.p = ( b2Vec2 ){ x0, 0.0f };
		pd.base.collideConnected = true;
		pd.normal = ( b2Vec2 ){ 0.0f, 1.0f };
		pd.restLength = 1.5f;
		pd.hertz = variant == 0 ? 4.0f : 0.0f;
		pd.dampingRatio = 0.6f;
		pd.maxTensionForce = 5.0f;
		pd.maxCompressionForce = 60.0f;
		pd.impulse = variant == 1 ? 0.25f : 0.0f;
		pd.velocity = 0.1f;
		b2CreatePogoJoint( worldId, &pd );
	}

	// revolute hinges with spring, limit and motor hanging from a static bar
	{
		b2BodyId prev = groundId;
		b2Vec2 anchor = { 0.0f, 12.0f };
		for ( int i = 0; i < 6; ++i )
		{
			b2BodyId link = b2sDynamicBox( scene, &capacity, ( b2Vec2 ){ 0.5f + 1.0f * i, 12.0f }, 0.0f, 0.5f, 0.1f, 0.6f );
			b2RevoluteJointDef rd = b2DefaultRevoluteJointDef();
			rd.base.bodyIdA = prev;
			rd.base.bodyIdB = link;
			rd.base.localFrameA.p = anchor;
			rd.base.localFrameB.p = ( b2Vec2 ){ -0.5f, 0.0f };
			rd.enableSpring = ( i & 1 ) == 0;
			rd.hertz = 2.0f;
			rd.dampingRatio = 0.5f;
			rd.enableLimit = i >= 2;
			rd.lowerAngle = -0.25f * B2_PI;
			rd.upperAngle = 0.3f * B2_PI;
			rd.enableMotor = i == 3;
			rd.motorSpeed = 1.0f;
			rd.maxMotorTorque = 2.0f;
			b2CreateRevoluteJoint( worldId, &rd );
			prev = link;
			anchor = ( b2Vec2 ){ 0.5f, 0.0f };
		}
	}
}

static void b2sBuildEvents( b2sScene* scene, int count, uint32_t seed )
{
	b2WorldId worldId = scene->worldId;
	s_rngState = seed;
	int capacity = 0;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Pos ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( 40.0f, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.enableContactEvents = true;
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}
	int columns = 12;
	for ( int i = 0; i < count; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ -18.0f + 3.0f * ( i % columns ), 2.0f + 2.5f * ( i / columns ) };
		bodyDef.rotation = b2MakeRot( b2sRandomRange( -1.0f, 1.0f ) );
		bodyDef.linearVelocity = ( b2Vec2 ){ b2sRandomRange( -3.0f, 3.0f ), b2sRandomRange( -8.0f, 0.0f ) };
		bodyDef.angularVelocity = b2sRandomRange( -2.0f, 2.0f );
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );

		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.enableContactEvents = ( i % 3 ) != 2;
		shapeDef.enableHitEvents = ( i % 2 ) == 0;
		shapeDef.material.restitution = ( i % 4 ) == 1 ? 0.3f : 0.0f;
		switch ( i % 3 )
		{
			case 0:
			{
				b2Circle circle = { { 0.0f, 0.0f }, b2sRandomRange( 0.3f, 0.6f ) };
				b2CreateCircleShape( bodyId, &shapeDef, &circle );
			}
			break;
			case 1:
			{
				b2Polygon box = b2MakeBox( b2sRandomRange( 0.3f, 0.8f ), b2sRandomRange( 0.3f, 0.8f ) );
				b2CreatePolygonShape( bodyId, &shapeDef, &box );
			}
			break;
			default:
			{
				b2Capsule capsule = { { -0.4f, 0.0f }, { 0.4f, 0.0f }, b2sRandomRange( 0.2f, 0.4f ) };
				b2CreateCapsuleShape( bodyId, &shapeDef, &capsule );
			}
			break;
		}
		b2sPushBody( scene, &capacity, bodyId );
	}
}

// Bodies, a hinge chain and a kinematic paddle that the edit script changes between steps.
static void b2sBuildEdits( b2sScene* scene, int enableSleep )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );
	scene->bodyCapacity = 0;

	b2BodyDef groundDef = b2DefaultBodyDef();
	b2BodyId groundId = b2CreateBody( worldId, &groundDef );
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	b2Polygon groundBox = b2MakeOffsetBox( 30.0f, 1.0f, ( b2Vec2 ){ 0.0f, -1.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &shapeDef, &groundBox );

	// 24 dynamic bodies: boxes, circles and capsules in turn
	for ( int i = 0; i < 24; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ -7.0f + 2.0f * (float)( i % 6 ), 1.0f + 1.5f * (float)( i / 6 ) };
		bodyDef.rotation = b2MakeRot( 0.1f * (float)i );
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.material.friction = 0.4f + 0.02f * (float)i;
		if ( i % 3 == 0 )
		{
			b2Polygon box = b2MakeBox( 0.4f, 0.3f );
			b2CreatePolygonShape( bodyId, &def, &box );
		}
		else if ( i % 3 == 1 )
		{
			b2Circle circle = { { 0.0f, 0.0f }, 0.4f };
			b2CreateCircleShape( bodyId, &def, &circle );
		}
		else
		{
			b2Capsule capsule = { { -0.3f, 0.0f }, { 0.3f, 0.0f }, 0.25f };
			b2CreateCapsuleShape( bodyId, &def, &capsule );
		}
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}

	// a hinge chain of capsule links hanging from the ground body
	b2BodyId prev = groundId;
	b2Vec2 anchor = { 10.0f, 8.0f };
	for ( int i = 0; i < B2S_EDIT_LINKS; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ anchor.x + 0.5f + 1.0f * (float)i, anchor.y };
		b2BodyId link = b2CreateBody( worldId, &bodyDef );
		b2Capsule capsule = { { -0.5f, 0.0f }, { 0.5f, 0.0f }, 0.15f };
		b2CreateCapsuleShape( link, &shapeDef, &capsule );

		b2RevoluteJointDef jd = b2DefaultRevoluteJointDef();
		jd.base.bodyIdA = prev;
		jd.base.bodyIdB = link;
		b2Pos pivot = { anchor.x + 1.0f * (float)i, anchor.y };
		jd.base.localFrameA.p = b2Body_GetLocalPoint( prev, pivot );
		jd.base.localFrameB.p = b2Body_GetLocalPoint( link, pivot );
		scene->hinges[i] = b2CreateRevoluteJoint( worldId, &jd );
		scene->links[i] = link;
		b2sPushBody( scene, &scene->bodyCapacity, link );
		prev = link;
	}

	// a kinematic paddle
	b2BodyDef kinDef = b2DefaultBodyDef();
	kinDef.type = b2_kinematicBody;
	kinDef.position = ( b2Pos ){ -10.0f, 3.0f };
	kinDef.angularVelocity = 1.0f;
	scene->kinematicId = b2CreateBody( worldId, &kinDef );
	b2Polygon paddle = b2MakeBox( 1.5f, 0.2f );
	b2CreatePolygonShape( scene->kinematicId, &shapeDef, &paddle );
	b2sPushBody( scene, &scene->bodyCapacity, scene->kinematicId );
}

static uint32_t b2sFnvWord( uint32_t hash, uint32_t word )
{
	for ( int b = 0; b < 4; ++b )
	{
		hash ^= ( word >> ( 8 * b ) ) & 0xFFu;
		hash *= 0x01000193u;
	}
	return hash;
}

static uint32_t b2sFnvFloat( uint32_t hash, float x )
{
	uint32_t u;
	memcpy( &u, &x, sizeof( u ) );
	return b2sFnvWord( hash, u );
}

static void b2sRemoveBody( b2sScene* scene, b2BodyId id )
{
	for ( int i = 0; i < scene->bodyCount; ++i )
	{
		if ( scene->bodies[i].index1 == id.index1 && scene->bodies[i].generation == id.generation )
		{
			memmove( scene->bodies + i, scene->bodies + i + 1, (size_t)( scene->bodyCount - i - 1 ) * sizeof( b2BodyId ) );
			scene->bodyCount -= 1;
			return;
		}
	}
}

// The edit script, run before world step k (0-based): body, shape and joint calls made between
// steps, as a user program makes them.
static void b2sEditsBeforeStep( b2sScene* scene, int k, float timeStep )
{
	b2WorldId worldId = scene->worldId;
	b2BodyId* b = scene->bodies;
	if ( k == 10 )
	{
		b2Body_SetTransform( b[3], ( b2Pos ){ 2.0f, 6.0f }, b2MakeRot( 0.5f ) );
	}
	if ( k == 12 )
	{
		b2Body_ApplyForce( b[5], ( b2Vec2 ){ 30.0f, 200.0f }, b2Body_GetWorldPoint( b[5], ( b2Vec2 ){ 0.2f, 0.1f } ), true );
		b2Body_ApplyTorque( b[6], 20.0f, true );
	}
	if ( k >= 15 && k < 40 )
	{
		b2Body_ApplyForceToCenter( b[4], ( b2Vec2 ){ 0.0f, 15.0f }, true );
	}
	if ( k == 20 )
	{
		b2Body_ApplyLinearImpulse( b[7], ( b2Vec2 ){ 2.0f, 5.0f }, b2Body_GetWorldPoint( b[7], ( b2Vec2 ){ 0.1f, 0.2f } ), true );
		b2Body_ApplyLinearImpulseToCenter( b[8], ( b2Vec2 ){ -3.0f, 2.0f }, true );
		b2Body_ApplyAngularImpulse( b[9], 1.5f, true );
	}
	if ( k == 25 )
	{
		for ( int i = 10; i < 13; ++i )
		{
			b2ShapeId shapeId;
			if ( b2Body_GetShapes( b[i], &shapeId, 1 ) != 1 )
			{
				continue;
			}
			b2ShapeType type = b2Shape_GetType( shapeId );
			if ( type == b2_polygonShape )
			{
				b2Polygon box = b2MakeBox( 0.6f, 0.3f );
				b2Shape_SetPolygon( shapeId, &box );
			}
			else if ( type == b2_circleShape )
			{
				b2Circle circle = { { 0.1f, 0.0f }, 0.6f };
				b2Shape_SetCircle( shapeId, &circle );
			}
			else if ( type == b2_capsuleShape )
			{
				b2Capsule capsule = { { -0.5f, 0.0f }, { 0.5f, 0.1f }, 0.2f };
				b2Shape_SetCapsule( shapeId, &capsule );
			}
			b2Body_ApplyMassFromShapes( b[i] );
		}
	}
	if ( k == 30 )
	{
		b2RevoluteJoint_EnableMotor( scene->hinges[1], true );
		b2RevoluteJoint_SetMotorSpeed( scene->hinges[1], 2.0f );
		b2RevoluteJoint_SetMaxMotorTorque( scene->hinges[1], 50.0f );
		b2RevoluteJoint_EnableSpring( scene->hinges[2], true );
		b2RevoluteJoint_SetSpringHertz( scene->hinges[2], 2.0f );
		b2RevoluteJoint_SetSpringDampingRatio( scene->hinges[2], 0.5f );
		b2RevoluteJoint_SetTargetAngle( scene->hinges[2], 0.3f );
		b2RevoluteJoint_EnableLimit( scene->hinges[3], true );
		b2RevoluteJoint_SetLimits( scene->hinges[3], 0.4f, -0.2f );
	}
	if ( k == 40 )
	{
		b2RevoluteJoint_EnableMotor( scene->hinges[1], false );
		b2RevoluteJoint_EnableLimit( scene->hinges[3], false );
	}
	if ( k == 45 )
	{
		b2DestroyJoint( scene->hinges[4] );
	}
	if ( k == 50 )
	{
		b2BodyId gone = b[13];
		b2DestroyBody( gone );
		b2sRemoveBody( scene, gone );
	}
	if ( k == 55 )
	{
		b2Transform frame = { { 0.5f, 0.0f }, b2MakeRot( 0.1f ) };
		b2Joint_SetLocalFrameA( scene->hinges[0], frame );
	}
	if ( k >= 60 )
	{
		b2WorldTransform target = { { -10.0f + 0.05f * (float)( k - 60 ), 3.0f }, b2MakeRot( 0.02f * (float)k ) };
		b2Body_SetTargetTransform( scene->kinematicId, target, timeStep, true );
	}
	if ( k == 70 )
	{
		b2BodyId gone = scene->links[2];
		b2DestroyBody( gone );
		b2sRemoveBody( scene, gone );
	}
	if ( k == 80 )
	{
		for ( int i = 0; i < 3; ++i )
		{
			b2BodyDef bodyDef = b2DefaultBodyDef();
			bodyDef.type = b2_dynamicBody;
			bodyDef.position = ( b2Pos ){ -4.0f + 3.0f * (float)i, 9.0f };
			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			b2ShapeDef def = b2DefaultShapeDef();
			b2Polygon box = b2MakeBox( 0.5f, 0.5f );
			b2CreatePolygonShape( bodyId, &def, &box );
			b2sPushBody( scene, &scene->bodyCapacity, bodyId );
		}
	}
	if ( k == 150 )
	{
		// a body that may be asleep: no wake, then a transform that wakes nothing by itself
		b2Body_ApplyForceToCenter( b[0], ( b2Vec2 ){ 0.0f, 500.0f }, false );
		b2Body_SetTransform( b[1], ( b2Pos ){ -3.0f, 5.0f }, b2Rot_identity );
		b2Body_ApplyLinearImpulseToCenter( b[2], ( b2Vec2 ){ 0.0f, 4.0f }, true );
	}
}

// Getter results after a step, folded into the scene's probe hash.
static void b2sEditsProbe( b2sScene* scene )
{
	uint32_t h = scene->probeHash;
	b2BodyId* b = scene->bodies;
	b2Vec2 p = b2Body_GetLocalPoint( b[0], ( b2Pos ){ 1.0f, 1.0f } );
	h = b2sFnvFloat( b2sFnvFloat( h, p.x ), p.y );
	b2Pos w = b2Body_GetWorldPoint( b[2], ( b2Vec2 ){ 0.5f, 0.5f } );
	h = b2sFnvFloat( b2sFnvFloat( h, w.x ), w.y );
	b2Vec2 v = b2Body_GetWorldVector( b[1], ( b2Vec2 ){ 1.0f, 0.0f } );
	h = b2sFnvFloat( b2sFnvFloat( h, v.x ), v.y );
	v = b2Body_GetLocalVector( b[2], ( b2Vec2 ){ 0.0f, 1.0f } );
	h = b2sFnvFloat( b2sFnvFloat( h, v.x ), v.y );
	for ( int i = 0; i < B2S_EDIT_LINKS; ++i )
	{
		b2JointId j = scene->hinges[i];
		if ( b2Joint_IsValid( j ) == false )
		{
			h = b2sFnvWord( h, 0xDEADu );
			continue;
		}
		h = b2sFnvFloat( h, b2RevoluteJoint_GetAngle( j ) );
		h = b2sFnvFloat( h, b2RevoluteJoint_GetMotorTorque( j ) );
		h = b2sFnvWord( h, (uint32_t)b2RevoluteJoint_IsMotorEnabled( j ) | (uint32_t)b2RevoluteJoint_IsSpringEnabled( j ) << 1 |
							   (uint32_t)b2RevoluteJoint_IsLimitEnabled( j ) << 2 );
		h = b2sFnvFloat( h, b2RevoluteJoint_GetLowerLimit( j ) );
		h = b2sFnvFloat( h, b2RevoluteJoint_GetUpperLimit( j ) );
		b2Transform frame = b2Joint_GetLocalFrameA( j );
		h = b2sFnvFloat( b2sFnvFloat( h, frame.p.x ), frame.q.s );
	}
	for ( int i = 0; i < B2S_EDIT_LINKS; ++i )
	{
		if ( b2Body_IsValid( scene->links[i] ) )
		{
			b2JointId joints[4];
			int n = b2Body_GetJoints( scene->links[i], joints, 4 );
			h = b2sFnvWord( h, (uint32_t)n | (uint32_t)b2Body_GetJointCount( scene->links[i] ) << 8 );
		}
	}
	for ( int i = 10; i < 13; ++i )
	{
		b2ShapeId shapeId;
		if ( b2Body_GetShapes( b[i], &shapeId, 1 ) != 1 )
		{
			continue;
		}
		switch ( b2Shape_GetType( shapeId ) )
		{
			case b2_polygonShape:
			{
				b2Polygon poly = b2Shape_GetPolygon( shapeId );
				h = b2sFnvFloat( b2sFnvFloat( h, poly.vertices[0].x ), poly.centroid.y );
			}
			break;
			case b2_circleShape:
				h = b2sFnvFloat( h, b2Shape_GetCircle( shapeId ).radius );
				break;
			case b2_capsuleShape:
				h = b2sFnvFloat( h, b2Shape_GetCapsule( shapeId ).center2.y );
				break;
			default:
				break;
		}
		h = b2sFnvFloat( h, b2Body_GetMass( b[i] ) );
		h = b2sFnvFloat( h, b2Body_GetRotationalInertia( b[i] ) );
	}
	h = b2sFnvWord( h, (uint32_t)b2World_GetAwakeBodyCount( scene->worldId ) );
	scene->probeHash = h;
}


// ---------------------------------------------------------------------------------------------
// Setters and callbacks scene

static float b2sFrictionMin( float frictionA, uint64_t materialA, float frictionB, uint64_t materialB )
{
	if ( materialA == 7 || materialB == 7 )
	{
		return 0.05f;
	}
	return frictionA < frictionB ? frictionA : frictionB;
}

static float b2sFrictionMax( float frictionA, uint64_t materialA, float frictionB, uint64_t materialB )
{
	(void)materialA;
	(void)materialB;
	return 1.5f * ( frictionA > frictionB ? frictionA : frictionB );
}

static float b2sRestitution( float restitutionA, uint64_t materialA, float restitutionB, uint64_t materialB )
{
	if ( materialA == 3 || materialB == 3 )
	{
		return 0.5f;
	}
	return restitutionA > restitutionB ? restitutionA : restitutionB;
}

// Rejects pairs whose bodies both have odd user data.
static bool b2sOddFilter( b2ShapeId shapeIdA, b2ShapeId shapeIdB, void* context )
{
	(void)context;
	intptr_t a = (intptr_t)b2Body_GetUserData( b2Shape_GetBody( shapeIdA ) );
	intptr_t b = (intptr_t)b2Body_GetUserData( b2Shape_GetBody( shapeIdB ) );
	return ( a & 1 ) == 0 || ( b & 1 ) == 0;
}

static b2ShapeId b2sFirstShape( b2BodyId bodyId )
{
	b2ShapeId shapeId = b2_nullShapeId;
	b2Body_GetShapes( bodyId, &shapeId, 1 );
	return shapeId;
}

static uint32_t b2sMix32( uint32_t x )
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

// One-way platform: a contact with the platform keeps its points only when the other shape is
// above it (normal within 60 degrees of up, pointing from the platform to the other shape); some
// other pairs lose their second point. The tally is a count and an order-independent sum, since
// the reference may call this from several threads in any order.
static void b2sPreSolve( b2ShapeId shapeIdA, b2ShapeId shapeIdB, b2Manifold* manifold, void* context )
{
	b2sScene* scene = (b2sScene*)context;
	scene->preSolveCalls += 1;
	scene->preSolveSum += b2sMix32( (uint32_t)shapeIdA.index1 * 65599u + (uint32_t)shapeIdB.index1 ) ^ (uint32_t)manifold->pointCount;
	bool platformA = B2_ID_EQUALS( shapeIdA, scene->platformShape );
	bool platformB = B2_ID_EQUALS( shapeIdB, scene->platformShape );
	if ( platformA || platformB )
	{
		float up = platformA ? manifold->normal.y : -manifold->normal.y;
		if ( up < 0.5f )
		{
			manifold->pointCount = 0;
		}
	}
	else if ( manifold->pointCount == 2 && ( shapeIdA.index1 + shapeIdB.index1 ) % 5 == 0 )
	{
		manifold->pointCount = 1;
	}
}

static void b2sBuildPreSolve( b2sScene* scene, int enableSleep )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );
	b2World_SetPreSolveCallback( worldId, b2sPreSolve, NULL, scene );
	scene->bodyCapacity = 0;

	b2BodyDef groundDef = b2DefaultBodyDef();
	b2BodyId groundId = b2CreateBody( worldId, &groundDef );
	b2ShapeDef groundShape = b2DefaultShapeDef();
	b2Polygon groundBox = b2MakeOffsetBox( 30.0f, 1.0f, ( b2Vec2 ){ 0.0f, -1.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &groundShape, &groundBox );

	b2BodyDef platformDef = b2DefaultBodyDef();
	platformDef.position = ( b2Pos ){ 0.0f, 4.0f };
	b2BodyId platformId = b2CreateBody( worldId, &platformDef );
	b2ShapeDef platformShape = b2DefaultShapeDef();
	platformShape.enablePreSolveEvents = true;
	b2Polygon platformBox = b2MakeBox( 8.0f, 0.25f );
	scene->platformShape = b2CreatePolygonShape( platformId, &platformShape, &platformBox );

	// boxes launched up through the platform from below, then shapes dropped onto it from above
	for ( int i = 0; i < 24; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bool launched = i < 12;
		float x = -6.5f + 1.2f * (float)( i % 12 );
		bodyDef.position = launched ? ( b2Pos ){ x, 1.0f + 0.3f * (float)( i % 3 ) } : ( b2Pos ){ x, 8.0f + 0.8f * (float)( i % 4 ) };
		bodyDef.rotation = b2MakeRot( 0.11f * (float)i );
		if ( launched )
		{
			bodyDef.linearVelocity = ( b2Vec2 ){ 0.3f * (float)( ( i % 5 ) - 2 ), 11.0f + 0.25f * (float)( i % 4 ) };
		}
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.material.friction = 0.5f;
		def.enablePreSolveEvents = ( i % 3 ) == 0;
		if ( launched || i % 2 == 0 )
		{
			b2Polygon box = b2MakeBox( 0.4f, 0.3f );
			b2CreatePolygonShape( bodyId, &def, &box );
		}
		else if ( i % 4 == 1 )
		{
			b2Circle circle = { { 0.0f, 0.0f }, 0.35f };
			b2CreateCircleShape( bodyId, &def, &circle );
		}
		else
		{
			b2Capsule capsule = { { -0.3f, 0.0f }, { 0.3f, 0.0f }, 0.2f };
			b2CreateCapsuleShape( bodyId, &def, &capsule );
		}
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}
}

// Before step k (0-based) of B2S_PRESOLVE: pre-solve off on the platform (existing contacts keep
// the flag), bodies launched again, the flag back on, the callback cleared and restored.
static void b2sPreSolveBeforeStep( b2sScene* scene, int k )
{
	b2BodyId* b = scene->bodies;
	scene->preSolveCalls = 0;
	scene->preSolveSum = 0;
	if ( k == 60 )
	{
		b2Shape_EnablePreSolveEvents( scene->platformShape, false );
	}
	if ( k == 60 || k == 110 )
	{
		for ( int i = 0; i < 4; ++i )
		{
			b2Body_SetLinearVelocity( b[3 * i + 1], ( b2Vec2 ){ 0.0f, 12.0f } );
			b2Body_SetAwake( b[3 * i + 1], true );
		}
	}
	if ( k == 90 )
	{
		b2Shape_EnablePreSolveEvents( scene->platformShape, true );
	}
	if ( k == 130 )
	{
		b2World_SetPreSolveCallback( scene->worldId, NULL, NULL, NULL );
	}
	if ( k == 150 )
	{
		b2World_SetPreSolveCallback( scene->worldId, b2sPreSolve, NULL, scene );
		for ( int i = 0; i < 4; ++i )
		{
			b2Body_SetLinearVelocity( b[3 * i + 2], ( b2Vec2 ){ 0.5f, 13.0f } );
			b2Body_SetAwake( b[3 * i + 2], true );
		}
	}
}

static void b2sPreSolveProbe( b2sScene* scene )
{
	uint32_t h = scene->probeHash;
	h = b2sFnvWord( h, scene->preSolveCalls );
	h = b2sFnvWord( h, scene->preSolveSum );
	h = b2sFnvWord( h, (uint32_t)b2Shape_ArePreSolveEventsEnabled( scene->platformShape ) );
	b2ContactData data[8];
	int n = b2Shape_GetContactData( scene->platformShape, data, 8 );
	h = b2sFnvWord( h, (uint32_t)n );
	scene->probeHash = h;
}

// Sensor filter: visitors with odd user data and custom filtering are not sensed.
static bool b2sSensorFilter( b2ShapeId shapeIdA, b2ShapeId shapeIdB, void* context )
{
	(void)context;
	intptr_t a = (intptr_t)b2Body_GetUserData( b2Shape_GetBody( shapeIdA ) );
	intptr_t b = (intptr_t)b2Body_GetUserData( b2Shape_GetBody( shapeIdB ) );
	return ( ( a | b ) & 1 ) == 0;
}

static void b2sBuildSensors( b2sScene* scene, int enableSleep )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );
	b2World_SetCustomFilterCallback( worldId, b2sSensorFilter, NULL );
	scene->bodyCapacity = 0;

	b2BodyDef groundDef = b2DefaultBodyDef();
	groundDef.userData = (void*)(intptr_t)1000;
	b2BodyId groundId = b2CreateBody( worldId, &groundDef );
	b2ShapeDef groundShape = b2DefaultShapeDef();
	groundShape.enableSensorEvents = true;
	b2Polygon groundBox = b2MakeOffsetBox( 40.0f, 1.0f, ( b2Vec2 ){ 0.0f, -1.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &groundShape, &groundBox );
	// a wall that stops the projectiles
	b2Polygon wall = b2MakeOffsetBox( 0.5f, 4.0f, ( b2Vec2 ){ 32.0f, 4.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &groundShape, &wall );

	// a static zone and a thin static strip, both sensors
	b2ShapeDef sensorDef = b2DefaultShapeDef();
	sensorDef.isSensor = true;
	sensorDef.enableSensorEvents = true;
	sensorDef.filter.maskBits = ~(uint64_t)0x4;
	b2Polygon zone = b2MakeOffsetBox( 4.0f, 2.5f, ( b2Vec2 ){ -10.0f, 3.0f }, b2Rot_identity );
	scene->zoneShape = b2CreatePolygonShape( groundId, &sensorDef, &zone );
	sensorDef.filter.maskBits = B2_DEFAULT_MASK_BITS;
	b2Polygon strip = b2MakeOffsetBox( 0.05f, 3.0f, ( b2Vec2 ){ 20.0f, 3.0f }, b2Rot_identity );
	scene->stripShape = b2CreatePolygonShape( groundId, &sensorDef, &strip );

	// a kinematic sensor sweeping across the zone
	b2BodyDef sweepDef = b2DefaultBodyDef();
	sweepDef.type = b2_kinematicBody;
	sweepDef.position = ( b2Pos ){ -25.0f, 1.5f };
	sweepDef.linearVelocity = ( b2Vec2 ){ 3.0f, 0.0f };
	sweepDef.userData = (void*)(intptr_t)2000;
	b2BodyId sweepId = b2CreateBody( worldId, &sweepDef );
	b2Circle sweepCircle = { { 0.0f, 0.0f }, 1.5f };
	scene->sweepShape = b2CreateCircleShape( sweepId, &sensorDef, &sweepCircle );
	b2sPushBody( scene, &scene->bodyCapacity, sweepId );

	// a dynamic carrier with a solid box and a sensor circle around it
	b2BodyDef carrierDef = b2DefaultBodyDef();
	carrierDef.type = b2_dynamicBody;
	carrierDef.position = ( b2Pos ){ 5.0f, 3.0f };
	carrierDef.userData = (void*)(intptr_t)3000;
	scene->carrierId = b2CreateBody( worldId, &carrierDef );
	b2ShapeDef solid = b2DefaultShapeDef();
	solid.enableSensorEvents = true;
	b2Polygon carrierBox = b2MakeBox( 0.5f, 0.5f );
	b2CreatePolygonShape( scene->carrierId, &solid, &carrierBox );
	b2Circle carried = { { 0.0f, 0.0f }, 2.0f };
	scene->carriedShape = b2CreateCircleShape( scene->carrierId, &sensorDef, &carried );
	b2sPushBody( scene, &scene->bodyCapacity, scene->carrierId );

	// visitors dropped over the zone and the carrier: two in three sense, every fifth is outside
	// the zone's mask, every seventh uses the custom filter
	for ( int i = 0; i < 30; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ ( i < 20 ? -13.0f : 3.0f ) + 0.9f * (float)( i % 10 ), 7.0f + 1.1f * (float)( i / 10 ) };
		bodyDef.rotation = b2MakeRot( 0.13f * (float)i );
		bodyDef.userData = (void*)(intptr_t)i;
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.enableSensorEvents = ( i % 3 ) != 2;
		def.enableCustomFiltering = ( i % 7 ) == 0;
		if ( i % 5 == 4 )
		{
			def.filter.categoryBits = 0x4;
		}
		b2ShapeId shapeId;
		if ( i % 3 == 0 )
		{
			b2Polygon box = b2MakeBox( 0.35f, 0.3f );
			shapeId = b2CreatePolygonShape( bodyId, &def, &box );
		}
		else if ( i % 3 == 1 )
		{
			b2Circle circle = { { 0.0f, 0.0f }, 0.3f };
			shapeId = b2CreateCircleShape( bodyId, &def, &circle );
		}
		else
		{
			b2Capsule capsule = { { -0.25f, 0.0f }, { 0.25f, 0.0f }, 0.2f };
			shapeId = b2CreateCapsuleShape( bodyId, &def, &capsule );
		}
		if ( i == 12 )
		{
			scene->doomedShape = shapeId;
		}
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}

	// fast projectiles that cross the thin strip within one step
	for ( int i = 0; i < 3; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ 8.0f, 1.0f + 1.5f * (float)i };
		bodyDef.linearVelocity = ( b2Vec2 ){ 110.0f + 7.0f * (float)i, 1.0f };
		bodyDef.userData = (void*)(intptr_t)( 100 + 2 * i );
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.enableSensorEvents = i != 2;
		b2Circle circle = { { 0.0f, 0.0f }, 0.1f };
		b2CreateCircleShape( bodyId, &def, &circle );
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}
}

// Before step k (0-based) of B2S_SENSORS.
static void b2sSensorsBeforeStep( b2sScene* scene, int k )
{
	b2BodyId* b = scene->bodies;
	if ( k == 40 )
	{
		b2Shape_EnableSensorEvents( scene->zoneShape, false );
	}
	if ( k == 60 )
	{
		b2Shape_EnableSensorEvents( scene->zoneShape, true );
	}
	if ( k == 70 )
	{
		b2DestroyShape( scene->doomedShape, true );
	}
	if ( k == 80 )
	{
		b2Body_Disable( scene->carrierId );
	}
	if ( k == 100 )
	{
		b2Body_Enable( scene->carrierId );
	}
	if ( k == 110 )
	{
		b2DestroyShape( scene->carriedShape, false );
	}
	if ( k == 120 )
	{
		// a visitor stops sensing, another changes its category into the zone's mask
		b2Shape_EnableSensorEvents( b2sFirstShape( b[2 + 3] ), false );
		b2Filter f = b2DefaultFilter();
		b2Shape_SetFilter( b2sFirstShape( b[2 + 4] ), f );
	}
	if ( k == 130 )
	{
		// projectiles again
		for ( int i = 0; i < 3; ++i )
		{
			b2BodyId p = b[32 + i];
			b2Body_SetTransform( p, ( b2Pos ){ 8.0f, 1.0f + 1.5f * (float)i }, b2Rot_identity );
			b2Body_SetLinearVelocity( p, ( b2Vec2 ){ 115.0f, 0.5f } );
		}
	}
}

static uint32_t b2sFnvShapeId( uint32_t h, b2ShapeId id )
{
	return b2sFnvWord( b2sFnvWord( h, (uint32_t)id.index1 ), (uint32_t)id.generation );
}

static void b2sSensorsProbe( b2sScene* scene )
{
	uint32_t h = scene->probeHash;
	b2SensorEvents ev = b2World_GetSensorEvents( scene->worldId );
	h = b2sFnvWord( h, (uint32_t)ev.beginCount | (uint32_t)ev.endCount << 16 );
	for ( int i = 0; i < ev.beginCount; ++i )
	{
		h = b2sFnvShapeId( b2sFnvShapeId( h, ev.beginEvents[i].sensorShapeId ), ev.beginEvents[i].visitorShapeId );
	}
	for ( int i = 0; i < ev.endCount; ++i )
	{
		h = b2sFnvShapeId( b2sFnvShapeId( h, ev.endEvents[i].sensorShapeId ), ev.endEvents[i].visitorShapeId );
	}
	b2ShapeId sensors[3] = { scene->zoneShape, scene->stripShape, scene->sweepShape };
	for ( int s = 0; s < 3; ++s )
	{
		b2ShapeId visitors[64];
		int n = b2Shape_GetSensorData( sensors[s], visitors, 64 );
		h = b2sFnvWord( h, (uint32_t)n | (uint32_t)b2Shape_GetSensorCapacity( sensors[s] ) << 16 );
		for ( int i = 0; i < n; ++i )
		{
			h = b2sFnvShapeId( h, visitors[i] );
		}
	}
	scene->probeHash = h;
}

// Pre-continuous callback of B2S_BULLETS: a hit on the ghost wall above y = 3 is ignored, every
// other hit stands. The tally is a count and an order-independent sum of the arguments.
static bool b2sPreContinuous( b2ShapeId shapeIdA, b2ShapeId shapeIdB, b2Pos point, b2Vec2 normal, void* context )
{
	b2sScene* scene = (b2sScene*)context;
	uint32_t px, py, nx, ny;
	float fx = (float)point.x, fy = (float)point.y;
	memcpy( &px, &fx, 4 );
	memcpy( &py, &fy, 4 );
	memcpy( &nx, &normal.x, 4 );
	memcpy( &ny, &normal.y, 4 );
	scene->preSolveCalls += 1;
	scene->preSolveSum += b2sMix32( (uint32_t)shapeIdA.index1 * 65599u + (uint32_t)shapeIdB.index1 ) ^ b2sMix32( px ^ b2sMix32( py ) ) ^
						  b2sMix32( nx + 3u * ny );
	bool ghost = B2_ID_EQUALS( shapeIdA, scene->ghostShape ) || B2_ID_EQUALS( shapeIdB, scene->ghostShape );
	return ghost == false || fy <= 3.0f;
}

static void b2sFire( b2BodyId id, float x, float y, float vx, float vy )
{
	b2Body_SetTransform( id, ( b2Pos ){ x, y }, b2Rot_identity );
	b2Body_SetLinearVelocity( id, ( b2Vec2 ){ vx, vy } );
	b2Body_SetAngularVelocity( id, 0.0f );
}

static void b2sBuildBullets( b2sScene* scene, int enableSleep, int preContinuous )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );
	if ( preContinuous )
	{
		b2World_SetPreSolveCallback( worldId, NULL, b2sPreContinuous, scene );
	}
	scene->bodyCapacity = 0;

	b2BodyDef groundDef = b2DefaultBodyDef();
	b2BodyId groundId = b2CreateBody( worldId, &groundDef );
	b2ShapeDef groundShape = b2DefaultShapeDef();
	groundShape.enableSensorEvents = true;
	b2Polygon groundBox = b2MakeOffsetBox( 40.0f, 1.0f, ( b2Vec2 ){ 0.0f, -1.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &groundShape, &groundBox );
	b2Polygon left = b2MakeOffsetBox( 0.1f, 6.0f, ( b2Vec2 ){ -35.0f, 6.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &groundShape, &left );
	b2Polygon right = b2MakeOffsetBox( 0.1f, 6.0f, ( b2Vec2 ){ 35.0f, 6.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &groundShape, &right );
	b2Polygon roof = b2MakeOffsetBox( 35.0f, 0.1f, ( b2Vec2 ){ 0.0f, 12.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &groundShape, &roof );

	// the ghost wall: pre-solve events on, the callback lets hits above y = 3 through
	b2ShapeDef ghostDef = b2DefaultShapeDef();
	ghostDef.enablePreSolveEvents = true;
	b2Polygon ghost = b2MakeOffsetBox( 0.1f, 4.0f, ( b2Vec2 ){ -22.0f, 4.0f }, b2Rot_identity );
	scene->ghostShape = b2CreatePolygonShape( groundId, &ghostDef, &ghost );

	// bullets: a circle, a box, a capsule, a circle with sensor and pre-solve events, a hexagon
	for ( int i = 0; i < 5; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.isBullet = true;
		bodyDef.position = ( b2Pos ){ 30.0f, 0.6f + 1.0f * (float)i };
		bodyDef.linearVelocity = ( b2Vec2 ){ -240.0f - 10.0f * (float)i, 0.0f };
		bodyDef.gravityScale = ( i % 2 ) == 0 ? 0.0f : 1.0f;
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.enableSensorEvents = i == 0 || i == 3;
		def.enablePreSolveEvents = i == 3;
		if ( i == 1 )
		{
			b2Polygon box = b2MakeBox( 0.1f, 0.1f );
			b2CreatePolygonShape( bodyId, &def, &box );
		}
		else if ( i == 2 )
		{
			b2Capsule capsule = { { -0.15f, 0.0f }, { 0.15f, 0.0f }, 0.08f };
			b2CreateCapsuleShape( bodyId, &def, &capsule );
		}
		else if ( i == 4 )
		{
			b2Vec2 points[6] = { { 0.12f, 0.0f }, { 0.06f, 0.1f }, { -0.06f, 0.1f }, { -0.12f, 0.0f }, { -0.06f, -0.1f }, { 0.06f, -0.1f } };
			b2Hull hull = b2ComputeHull( points, 6 );
			b2Polygon hex = b2MakePolygon( &hull, 0.0f );
			b2CreatePolygonShape( bodyId, &def, &hex );
		}
		else
		{
			b2Circle circle = { { 0.0f, 0.0f }, i == 3 ? 0.15f : 0.1f };
			b2CreateCircleShape( bodyId, &def, &circle );
		}
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}

	// fast bodies that are not bullets: they sweep against static shapes only
	for ( int i = 0; i < 2; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ 30.0f, 1.2f + 2.0f * (float)i };
		bodyDef.linearVelocity = ( b2Vec2 ){ -230.0f, 0.0f };
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.enableSensorEvents = true;
		b2Circle circle = { { 0.0f, 0.0f }, 0.1f };
		b2CreateCircleShape( bodyId, &def, &circle );
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}

	// targets: a stack of boxes, a thin standing plank, a rotating kinematic paddle and a dynamic
	// body carrying a sensor ring
	for ( int i = 0; i < 12; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ 10.0f + 0.8f * (float)( i % 2 ), 0.4f + 0.8f * (float)( i / 2 ) };
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.enableSensorEvents = true;
		b2Polygon box = b2MakeBox( 0.4f, 0.4f );
		b2CreatePolygonShape( bodyId, &def, &box );
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ 0.0f, 2.0f };
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		b2Polygon plank = b2MakeBox( 0.05f, 2.0f );
		b2CreatePolygonShape( bodyId, &def, &plank );
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_kinematicBody;
		bodyDef.position = ( b2Pos ){ -10.0f, 3.0f };
		bodyDef.angularVelocity = 1.5f;
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		b2Polygon paddle = b2MakeBox( 0.1f, 2.5f );
		b2CreatePolygonShape( bodyId, &def, &paddle );
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ -15.0f, 0.5f };
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.enableSensorEvents = true;
		b2Polygon box = b2MakeBox( 0.5f, 0.5f );
		b2CreatePolygonShape( bodyId, &def, &box );
		b2ShapeDef ringDef = b2DefaultShapeDef();
		ringDef.isSensor = true;
		ringDef.enableSensorEvents = true;
		b2Circle ring = { { 0.0f, 1.5f }, 1.2f };
		scene->ringShape = b2CreateCircleShape( bodyId, &ringDef, &ring );
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}
}

// Before step k (0-based) of B2S_BULLETS: the bullets and the fast bodies are fired again from
// either end and at angles, and bullet flags are switched.
static void b2sBulletsBeforeStep( b2sScene* scene, int k )
{
	b2BodyId* b = scene->bodies;
	if ( k == 30 || k == 140 )
	{
		for ( int i = 0; i < 7; ++i )
		{
			b2sFire( b[i], 30.0f, 0.7f + 0.8f * (float)i, -260.0f, 5.0f - 1.5f * (float)i );
		}
	}
	if ( k == 50 )
	{
		b2Body_SetBullet( b[5], true );
		b2Body_SetBullet( b[4], false );
		for ( int i = 0; i < 7; ++i )
		{
			b2sFire( b[i], -30.0f, 0.8f + 0.75f * (float)i, 250.0f, 2.0f );
		}
	}
	if ( k == 80 )
	{
		for ( int i = 0; i < 7; ++i )
		{
			b2sFire( b[i], 30.0f, 6.0f + 0.5f * (float)i, -200.0f, -25.0f );
		}
	}
	if ( k == 110 )
	{
		b2Body_SetBullet( b[4], true );
		for ( int i = 0; i < 7; ++i )
		{
			b2sFire( b[i], -30.0f, 1.0f + 0.6f * (float)i, 220.0f, 30.0f );
		}
	}
}

static void b2sBulletsProbe( b2sScene* scene )
{
	uint32_t h = scene->probeHash;
	b2SensorEvents ev = b2World_GetSensorEvents( scene->worldId );
	h = b2sFnvWord( h, (uint32_t)ev.beginCount | (uint32_t)ev.endCount << 16 );
	for ( int i = 0; i < ev.beginCount; ++i )
	{
		h = b2sFnvShapeId( b2sFnvShapeId( h, ev.beginEvents[i].sensorShapeId ), ev.beginEvents[i].visitorShapeId );
	}
	for ( int i = 0; i < ev.endCount; ++i )
	{
		h = b2sFnvShapeId( b2sFnvShapeId( h, ev.endEvents[i].sensorShapeId ), ev.endEvents[i].visitorShapeId );
	}
	b2ShapeId visitors[64];
	int n = b2Shape_GetSensorData( scene->ringShape, visitors, 64 );
	h = b2sFnvWord( h, (uint32_t)n );
	for ( int i = 0; i < n; ++i )
	{
		h = b2sFnvShapeId( h, visitors[i] );
	}
	uint32_t bits = 0;
	for ( int i = 0; i < 7; ++i )
	{
		bits |= b2Body_IsBullet( scene->bodies[i] ) ? 1u << i : 0u;
	}
	h = b2sFnvWord( h, bits );
	h = b2sFnvWord( h, scene->preSolveCalls );
	h = b2sFnvWord( h, scene->preSolveSum );
	scene->preSolveCalls = 0;
	scene->preSolveSum = 0;
	scene->probeHash = h;
}

static void b2sBuildSetters( b2sScene* scene, int enableSleep )
{
	b2WorldId worldId = scene->worldId;
	b2World_EnableSleeping( worldId, enableSleep != 0 );
	b2World_SetFrictionCallback( worldId, b2sFrictionMin );
	b2World_SetRestitutionCallback( worldId, b2sRestitution );
	b2World_SetCustomFilterCallback( worldId, b2sOddFilter, scene );
	scene->bodyCapacity = 0;

	b2BodyDef groundDef = b2DefaultBodyDef();
	groundDef.userData = (void*)(intptr_t)1001;
	b2BodyId groundId = b2CreateBody( worldId, &groundDef );
	b2ShapeDef groundShape = b2DefaultShapeDef();
	b2Polygon groundBox = b2MakeOffsetBox( 30.0f, 1.0f, ( b2Vec2 ){ 0.0f, -1.0f }, b2Rot_identity );
	b2CreatePolygonShape( groundId, &groundShape, &groundBox );

	for ( int i = 0; i < 30; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ -9.0f + 2.5f * (float)( i % 6 ), 1.0f + 1.6f * (float)( i / 6 ) };
		bodyDef.rotation = b2MakeRot( 0.07f * (float)i );
		bodyDef.userData = (void*)(intptr_t)i;
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef def = b2DefaultShapeDef();
		def.material.friction = 0.3f + 0.02f * (float)i;
		def.material.restitution = ( i % 5 == 0 ) ? 0.2f : 0.0f;
		def.material.userMaterialId = (uint64_t)( i % 8 );
		def.enableCustomFiltering = ( i % 3 ) == 0;
		if ( i % 3 == 0 )
		{
			b2Polygon box = b2MakeBox( 0.45f, 0.35f );
			b2CreatePolygonShape( bodyId, &def, &box );
		}
		else if ( i % 3 == 1 )
		{
			b2Circle circle = { { 0.0f, 0.0f }, 0.42f };
			b2CreateCircleShape( bodyId, &def, &circle );
		}
		else
		{
			b2Capsule capsule = { { -0.35f, 0.0f }, { 0.35f, 0.0f }, 0.25f };
			b2CreateCapsuleShape( bodyId, &def, &capsule );
		}
		if ( i == 11 )
		{
			b2Circle extra = { { 0.6f, 0.0f }, 0.2f };
			scene->extraShape = b2CreateCircleShape( bodyId, &def, &extra );
		}
		b2sPushBody( scene, &scene->bodyCapacity, bodyId );
	}

	b2ShapeDef linkShape = b2DefaultShapeDef();
	b2BodyId prev = groundId;
	b2Vec2 anchor = { 12.0f, 7.0f };
	for ( int i = 0; i < 3; ++i )
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Pos ){ anchor.x + 0.5f + 1.0f * (float)i, anchor.y };
		bodyDef.userData = (void*)(intptr_t)( 100 + i );
		b2BodyId link = b2CreateBody( worldId, &bodyDef );
		b2Capsule capsule = { { -0.5f, 0.0f }, { 0.5f, 0.0f }, 0.15f };
		b2CreateCapsuleShape( link, &linkShape, &capsule );
		b2RevoluteJointDef jd = b2DefaultRevoluteJointDef();
		jd.base.bodyIdA = prev;
		jd.base.bodyIdB = link;
		b2Pos pivot = { anchor.x + 1.0f * (float)i, anchor.y };
		jd.base.localFrameA.p = b2Body_GetLocalPoint( prev, pivot );
		jd.base.localFrameB.p = b2Body_GetLocalPoint( link, pivot );
		b2CreateRevoluteJoint( worldId, &jd );
		scene->setterLinks[i] = link;
		b2sPushBody( scene, &scene->bodyCapacity, link );
		prev = link;
	}

	b2BodyDef kinDef = b2DefaultBodyDef();
	kinDef.type = b2_kinematicBody;
	kinDef.position = ( b2Pos ){ -12.0f, 2.5f };
	kinDef.angularVelocity = 0.8f;
	kinDef.userData = (void*)(intptr_t)200;
	scene->kinematicId = b2CreateBody( worldId, &kinDef );
	b2Polygon paddle = b2MakeBox( 1.5f, 0.2f );
	b2CreatePolygonShape( scene->kinematicId, &linkShape, &paddle );
	b2sPushBody( scene, &scene->bodyCapacity, scene->kinematicId );
}

static void b2sSettersBeforeStep( b2sScene* scene, int k )
{
	b2WorldId worldId = scene->worldId;
	b2BodyId* b = scene->bodies;
	if ( k == 20 )
	{
		b2Body_SetLinearDamping( b[0], 0.8f );
		b2Body_SetAngularDamping( b[1], 0.5f );
		b2Body_SetGravityScale( b[2], -0.3f );
	}
	if ( k == 30 )
	{
		b2Shape_SetFriction( b2sFirstShape( b[3] ), 0.05f );
		b2Shape_SetRestitution( b2sFirstShape( b[4] ), 0.6f );
		b2Shape_SetUserMaterial( b2sFirstShape( b[5] ), 7 );
	}
	if ( k == 40 )
	{
		b2World_SetFrictionCallback( worldId, NULL );
	}
	if ( k == 45 )
	{
		b2Body_SetType( b[6], b2_staticBody );
	}
	if ( k == 50 )
	{
		b2World_SetRestitutionThreshold( worldId, 0.3f );
		b2World_SetHitEventThreshold( worldId, 0.5f );
	}
	if ( k == 55 )
	{
		b2Body_SetType( b[6], b2_dynamicBody );
		b2Body_SetType( b[7], b2_kinematicBody );
		b2Body_SetLinearVelocity( b[7], ( b2Vec2 ){ 0.5f, 1.0f } );
	}
	if ( k == 60 )
	{
		b2World_SetFrictionCallback( worldId, b2sFrictionMax );
	}
	if ( k == 65 )
	{
		b2Body_Disable( b[8] );
		b2Body_Disable( scene->setterLinks[1] );
	}
	if ( k == 75 )
	{
		b2Body_Enable( b[8] );
	}
	if ( k == 80 )
	{
		b2Filter filter = b2DefaultFilter();
		filter.categoryBits = 2;
		filter.maskBits = ~(uint64_t)4;
		b2Shape_SetFilter( b2sFirstShape( b[9] ), filter );
		b2Filter groupFilter = b2DefaultFilter();
		groupFilter.groupIndex = -3;
		b2Shape_SetFilter( b2sFirstShape( b[15] ), groupFilter );
		b2Shape_SetFilter( b2sFirstShape( b[21] ), groupFilter );
		b2Shape_SetDensity( b2sFirstShape( b[10] ), 3.0f, true );
	}
	if ( k == 85 )
	{
		b2DestroyShape( scene->extraShape, true );
	}
	if ( k == 90 )
	{
		b2Body_Enable( scene->setterLinks[1] );
		b2World_SetCustomFilterCallback( worldId, NULL, NULL );
	}
	if ( k == 100 )
	{
		b2Body_SetAwake( b[12], false );
		b2Body_EnableSleep( b[13], false );
		b2Body_SetSleepThreshold( b[14], 0.01f );
	}
	if ( k == 110 )
	{
		b2Body_SetType( scene->setterLinks[0], b2_staticBody );
	}
	if ( k == 120 )
	{
		b2Body_SetAwake( b[12], true );
		b2Body_WakeTouching( b[0] );
		b2Body_EnableSleep( b[13], true );
	}
	if ( k == 130 )
	{
		b2World_SetRestitutionCallback( worldId, NULL );
		b2SurfaceMaterial m = b2DefaultSurfaceMaterial();
		m.friction = 0.9f;
		m.restitution = 0.3f;
		m.userMaterialId = 3;
		b2Shape_SetSurfaceMaterial( b2sFirstShape( b[16] ), &m );
	}
	if ( k == 140 )
	{
		b2Body_SetType( b[7], b2_dynamicBody );
		b2Body_SetType( scene->setterLinks[0], b2_dynamicBody );
	}
	if ( k == 150 )
	{
		b2Body_Disable( scene->kinematicId );
	}
	if ( k == 160 )
	{
		b2Body_Enable( scene->kinematicId );
		b2World_SetRestitutionCallback( worldId, b2sRestitution );
	}
}

static uint32_t b2sFnvManifold( uint32_t h, const b2Manifold* m )
{
	h = b2sFnvFloat( b2sFnvFloat( h, m->normal.x ), m->normal.y );
	h = b2sFnvWord( h, (uint32_t)m->pointCount );
	for ( int p = 0; p < m->pointCount; ++p )
	{
		const b2ManifoldPoint* mp = m->points + p;
		h = b2sFnvFloat( b2sFnvFloat( h, mp->anchorA.x ), mp->anchorB.y );
		h = b2sFnvFloat( b2sFnvFloat( h, mp->separation ), mp->normalImpulse );
		h = b2sFnvFloat( b2sFnvFloat( h, mp->tangentImpulse ), mp->totalNormalImpulse );
		h = b2sFnvWord( h, (uint32_t)mp->id | ( mp->persisted ? 0x10000u : 0u ) );
	}
	return h;
}

static void b2sSettersProbe( b2sScene* scene )
{
	uint32_t h = scene->probeHash;
	b2BodyId* b = scene->bodies;
	b2WorldId worldId = scene->worldId;
	h = b2sFnvFloat( h, b2World_GetRestitutionThreshold( worldId ) );
	h = b2sFnvFloat( h, b2World_GetHitEventThreshold( worldId ) );
	for ( int i = 0; i < 16; ++i )
	{
		b2BodyId id = b[i];
		h = b2sFnvFloat( h, b2Body_GetLinearDamping( id ) );
		h = b2sFnvFloat( h, b2Body_GetAngularDamping( id ) );
		h = b2sFnvFloat( h, b2Body_GetGravityScale( id ) );
		h = b2sFnvFloat( h, b2Body_GetSleepThreshold( id ) );
		h = b2sFnvWord( h, (uint32_t)b2Body_GetType( id ) | (uint32_t)b2Body_IsEnabled( id ) << 4 |
							   (uint32_t)b2Body_IsSleepEnabled( id ) << 5 | (uint32_t)b2Body_IsAwake( id ) << 6 );
		h = b2sFnvFloat( h, b2Body_GetMass( id ) );
		b2ShapeId shapeId = b2sFirstShape( id );
		h = b2sFnvFloat( h, b2Shape_GetFriction( shapeId ) );
		h = b2sFnvFloat( h, b2Shape_GetRestitution( shapeId ) );
		h = b2sFnvFloat( h, b2Shape_GetDensity( shapeId ) );
		b2Filter f = b2Shape_GetFilter( shapeId );
		h = b2sFnvWord( h, (uint32_t)f.categoryBits ^ (uint32_t)( f.maskBits >> 3 ) ^ (uint32_t)f.groupIndex );
		h = b2sFnvWord( h, (uint32_t)b2Shape_GetUserMaterial( shapeId ) | (uint32_t)b2Body_GetShapeCount( id ) << 8 );
	}
	for ( int i = 0; i < 3; ++i )
	{
		h = b2sFnvWord( h, (uint32_t)b2Body_GetType( scene->setterLinks[i] ) | (uint32_t)b2Body_IsEnabled( scene->setterLinks[i] ) << 4 );
	}
	h = b2sFnvWord( h, (uint32_t)b2Body_IsEnabled( scene->kinematicId ) );

	// contact data of a body, of a shape and of one contact
	b2ContactData data[16];
	int n = b2Body_GetContactData( b[0], data, 16 );
	h = b2sFnvWord( h, (uint32_t)n | (uint32_t)b2Body_GetContactCapacity( b[0] ) << 8 );
	for ( int i = 0; i < n; ++i )
	{
		h = b2sFnvWord( h, (uint32_t)data[i].shapeIdA.index1 | (uint32_t)data[i].shapeIdB.index1 << 16 );
		h = b2sFnvManifold( h, &data[i].manifold );
	}
	if ( n > 0 && b2Contact_IsValid( data[0].contactId ) )
	{
		b2ContactData one = b2Contact_GetData( data[0].contactId );
		h = b2sFnvManifold( h, &one.manifold );
	}
	n = b2Shape_GetContactData( b2sFirstShape( b[6] ), data, 16 );
	h = b2sFnvWord( h, (uint32_t)n | (uint32_t)b2Shape_GetContactCapacity( b2sFirstShape( b[6] ) ) << 8 );
	for ( int i = 0; i < n; ++i )
	{
		h = b2sFnvManifold( h, &data[i].manifold );
	}
	h = b2sFnvWord( h, (uint32_t)b2World_GetAwakeBodyCount( worldId ) );
	scene->probeHash = h;
}

int b2s_create( int kind, int workerCount, const int* iparams, int iparamCount, const float* fparams, int fparamCount )
{
	int handle = -1;
	for ( int i = 0; i < B2S_MAX_HANDLES; ++i )
	{
		if ( s_scenes[i].inUse == 0 )
		{
			handle = i;
			break;
		}
	}

	if ( handle < 0 )
	{
		return -1;
	}

	b2WorldDef worldDef = b2DefaultWorldDef();
	worldDef.workerCount = workerCount;
	b2WorldId worldId = b2CreateWorld( &worldDef );
	if ( worldId.index1 == 0 )
	{
		return -1;
	}

	b2sScene* scene = s_scenes + handle;
	memset( scene, 0, sizeof( *scene ) );
	scene->worldId = worldId;
	scene->inUse = 1;
	scene->kind = kind;

	switch ( kind )
	{
		case B2S_PYRAMID:
		{
			int rows = iparamCount > 0 ? iparams[0] : 20;
			int sleep = iparamCount > 1 ? iparams[1] : 0;
			float groundHalfWidth = fparamCount > 0 ? fparams[0] : 40.0f;
			b2sBuildPyramid( scene, rows, groundHalfWidth, sleep );
		}
		break;

		case B2S_FALLING_HINGES:
			b2sBuildFallingHinges( scene );
			break;

		case B2S_HELLO:
			b2sBuildHello( scene );
			break;

		case B2S_MIXED_PILE:
		{
			int count = iparamCount > 0 ? iparams[0] : 100;
			int sleep = iparamCount > 1 ? iparams[1] : 0;
			uint32_t seed = iparamCount > 2 ? (uint32_t)iparams[2] : 12345u;
			b2sBuildMixedPile( scene, count, sleep, seed );
		}
		break;

		case B2S_BATCH_WORLD:
		{
			int target = iparamCount > 0 ? iparams[0] : 64;
			int which = iparamCount > 1 ? iparams[1] : 0;
			uint32_t seed = iparamCount > 2 ? (uint32_t)iparams[2] : 12345u;
			b2sBuildBatchWorld( scene, target, which, seed );
		}
		break;

		case B2S_PUSH:
		{
			int rows = iparamCount > 0 ? iparams[0] : 10;
			float vx = fparamCount > 0 ? fparams[0] : 4.0f;
			float vy = fparamCount > 1 ? fparams[1] : 0.0f;
			float omega = fparamCount > 2 ? fparams[2] : 0.0f;
			float friction = fparamCount > 3 ? fparams[3] : 0.6f;
			b2sBuildPush( scene, rows, vx, vy, omega, friction );
		}
		break;

		case B2S_EVENTS:
			b2sBuildEvents( scene, iparamCount > 0 ? iparams[0] : 60, iparamCount > 1 ? (uint32_t)iparams[1] : 12345u );
			break;

		case B2S_EDITS:
			b2sBuildEdits( scene, iparamCount > 0 ? iparams[0] : 0 );
			break;

		case B2S_JOINT_ZOO:
			b2sBuildJointZoo( scene, iparamCount > 0 ? iparams[0] : 0 );
			if ( iparamCount > 1 && iparams[1] != 0 )
			{
				b2sZooThresholds( scene );
			}
			break;

		case B2S_SETTERS:
			b2sBuildSetters( scene, iparamCount > 0 ? iparams[0] : 0 );
			break;

		case B2S_PRESOLVE:
			b2sBuildPreSolve( scene, iparamCount > 0 ? iparams[0] : 0 );
			break;

		case B2S_SENSORS:
			b2sBuildSensors( scene, iparamCount > 0 ? iparams[0] : 0 );
			break;

		case B2S_BULLETS:
			b2sBuildBullets( scene, iparamCount > 0 ? iparams[0] : 0, iparamCount > 1 ? iparams[1] : 0 );
			break;

		case B2S_RAIN:
		{
			int count = iparamCount > 0 ? iparams[0] : 20000;
			int columns = iparamCount > 1 ? iparams[1] : 100;
			int sleep = iparamCount > 2 ? iparams[2] : 0;
			uint32_t seed = iparamCount > 3 ? (uint32_t)iparams[3] : 12345u;
			float width = fparamCount > 0 ? fparams[0] : 120.0f;
			b2sBuildRain( scene, count, columns, sleep, seed, width );
		}
		break;

		default:
			b2DestroyWorld( worldId );
			scene->inUse = 0;
			return -1;
	}

	return handle;
}

static b2sScene* b2sGet( int handle )
{
	if ( handle < 0 || handle >= B2S_MAX_HANDLES || s_scenes[handle].inUse == 0 )
	{
		return NULL;
	}
	return s_scenes + handle;
}

uint32_t b2s_world_id( int handle )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	return (uint32_t)scene->worldId.index1 | ( (uint32_t)scene->worldId.generation << 16 );
}

void b2s_destroy( int handle )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return;
	}
	b2DestroyWorld( scene->worldId );
	free( scene->bodies );
	free( scene->zooJoints );
	memset( scene, 0, sizeof( *scene ) );
}

void b2s_step( int handle, float timeStep, int subStepCount, int count )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return;
	}
	for ( int i = 0; i < count; ++i )
	{
		if ( scene->kind == B2S_EDITS )
		{
			b2sEditsBeforeStep( scene, scene->stepIndex, timeStep );
		}
		else if ( scene->kind == B2S_SETTERS )
		{
			b2sSettersBeforeStep( scene, scene->stepIndex );
		}
		else if ( scene->kind == B2S_PRESOLVE )
		{
			b2sPreSolveBeforeStep( scene, scene->stepIndex );
		}
		else if ( scene->kind == B2S_SENSORS )
		{
			b2sSensorsBeforeStep( scene, scene->stepIndex );
		}
		else if ( scene->kind == B2S_BULLETS )
		{
			b2sBulletsBeforeStep( scene, scene->stepIndex );
		}
		else if ( scene->kind == B2S_JOINT_ZOO && scene->zooJointCount > 0 && scene->stepIndex == 150 )
		{
			for ( int k = 0; k < scene->zooJointCount; ++k )
			{
				b2sSetZooThresholds( scene->zooJoints[k], k, 1 );
			}
		}
		b2World_Step( scene->worldId, timeStep, subStepCount );
		scene->stepIndex += 1;
		if ( scene->kind == B2S_EDITS )
		{
			b2sEditsProbe( scene );
		}
		else if ( scene->kind == B2S_SETTERS )
		{
			b2sSettersProbe( scene );
		}
		else if ( scene->kind == B2S_PRESOLVE )
		{
			b2sPreSolveProbe( scene );
		}
		else if ( scene->kind == B2S_SENSORS )
		{
			b2sSensorsProbe( scene );
		}
		else if ( scene->kind == B2S_BULLETS )
		{
			b2sBulletsProbe( scene );
		}
	}
}

uint32_t b2s_probe_hash( int handle )
{
	b2sScene* scene = b2sGet( handle );
	return scene != NULL ? scene->probeHash : 0;
}

uint64_t b2s_state_hash( int handle )
{
	b2sScene* scene = b2sGet( handle );
	return scene != NULL ? b2World_GetStateHash( scene->worldId ) : 0;
}

static double b2sNowMs( void )
{
	struct timespec ts;
	clock_gettime( CLOCK_MONOTONIC, &ts );
	return 1000.0 * (double)ts.tv_sec + 1.0e-6 * (double)ts.tv_nsec;
}

double b2s_step_timed( int handle, float timeStep, int subStepCount, int count )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return -1.0;
	}
	double start = b2sNowMs();
	for ( int i = 0; i < count; ++i )
	{
		b2World_Step( scene->worldId, timeStep, subStepCount );
	}
	return b2sNowMs() - start;
}

double b2s_step_profiled( int handle, float timeStep, int subStepCount, int count, double* profileSums )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return -1.0;
	}
	double total = 0.0;
	for ( int i = 0; i < count; ++i )
	{
		double start = b2sNowMs();
		b2World_Step( scene->worldId, timeStep, subStepCount );
		total += b2sNowMs() - start;
		b2Profile p = b2World_GetProfile( scene->worldId );
		const float* f = &p.step;
		for ( int k = 0; k < (int)( sizeof( b2Profile ) / sizeof( float ) ); ++k )
		{
			profileSums[k] += f[k];
		}
	}
	return total;
}

int b2s_body_ids( int handle, int* out, int capacity )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	int n = scene->bodyCount < capacity ? scene->bodyCount : capacity;
	for ( int i = 0; i < n; ++i )
	{
		out[i] = scene->bodies[i].index1 - 1;
	}
	return n;
}

int b2s_body_count( int handle )
{
	b2sScene* scene = b2sGet( handle );
	return scene == NULL ? 0 : scene->bodyCount;
}

uint32_t b2s_hash( int handle )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	uint32_t hash = B2_HASH_INIT;
	for ( int i = 0; i < scene->bodyCount; ++i )
	{
		b2WorldTransform transform = b2Body_GetTransform( scene->bodies[i] );
		hash = b2Hash( hash, (uint8_t*)( &transform ), sizeof( b2WorldTransform ) );
	}
	return hash;
}

int b2s_get_transforms( int handle, float* out, int capacity )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	int n = scene->bodyCount < capacity ? scene->bodyCount : capacity;
	for ( int i = 0; i < n; ++i )
	{
		b2WorldTransform t = b2Body_GetTransform( scene->bodies[i] );
		out[4 * i + 0] = (float)t.p.x;
		out[4 * i + 1] = (float)t.p.y;
		out[4 * i + 2] = t.q.c;
		out[4 * i + 3] = t.q.s;
	}
	return n;
}

int b2s_get_velocities( int handle, float* out, int capacity )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	int n = scene->bodyCount < capacity ? scene->bodyCount : capacity;
	for ( int i = 0; i < n; ++i )
	{
		b2Vec2 v = b2Body_GetLinearVelocity( scene->bodies[i] );
		out[3 * i + 0] = v.x;
		out[3 * i + 1] = v.y;
		out[3 * i + 2] = b2Body_GetAngularVelocity( scene->bodies[i] );
	}
	return n;
}

int b2s_move_count( int handle )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	b2BodyEvents events = b2World_GetBodyEvents( scene->worldId );
	return events.moveCount;
}

static uint32_t b2sFloatBits( float x )
{
	uint32_t u;
	memcpy( &u, &x, sizeof( u ) );
	return u;
}

int b2s_sensor_events( int handle, uint32_t* out, int capacity )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	b2SensorEvents events = b2World_GetSensorEvents( scene->worldId );
	int n = 0;
	for ( int k = 0; k < events.beginCount + events.endCount && 5 * ( n + 1 ) <= capacity; ++k, ++n )
	{
		bool begin = k < events.beginCount;
		b2ShapeId sensorId = begin ? events.beginEvents[k].sensorShapeId : events.endEvents[k - events.beginCount].sensorShapeId;
		b2ShapeId visitorId = begin ? events.beginEvents[k].visitorShapeId : events.endEvents[k - events.beginCount].visitorShapeId;
		uint32_t* r = out + 5 * n;
		r[0] = begin ? 0u : 1u;
		r[1] = (uint32_t)sensorId.index1;
		r[2] = sensorId.generation;
		r[3] = (uint32_t)visitorId.index1;
		r[4] = visitorId.generation;
	}
	return n;
}

int b2s_joint_events( int handle, uint32_t* out, int capacity )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	b2JointEvents events = b2World_GetJointEvents( scene->worldId );
	int n = 0;
	for ( int k = 0; k < events.count && 3 * ( n + 1 ) <= capacity; ++k, ++n )
	{
		const b2JointEvent* e = events.jointEvents + k;
		out[3 * n + 0] = (uint32_t)e->jointId.index1;
		out[3 * n + 1] = e->jointId.generation;
		out[3 * n + 2] = (uint32_t)(intptr_t)e->userData;
	}
	return n;
}

int b2s_contact_events( int handle, uint32_t* out, int capacity )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL )
	{
		return 0;
	}
	b2ContactEvents events = b2World_GetContactEvents( scene->worldId );
	int total = events.beginCount + events.endCount + events.hitCount;
	int n = 0;
	for ( int k = 0; k < total && 10 * ( n + 1 ) <= capacity; ++k )
	{
		uint32_t* r = out + 10 * n;
		memset( r, 0, 10 * sizeof( uint32_t ) );
		if ( k < events.beginCount )
		{
			const b2ContactBeginTouchEvent* e = events.beginEvents + k;
			r[0] = 0;
			r[1] = (uint32_t)e->shapeIdA.index1;
			r[2] = (uint32_t)e->shapeIdB.index1;
			r[3] = (uint32_t)e->contactId.index1;
			r[4] = e->contactId.generation;
		}
		else if ( k < events.beginCount + events.endCount )
		{
			const b2ContactEndTouchEvent* e = events.endEvents + ( k - events.beginCount );
			r[0] = 1;
			r[1] = (uint32_t)e->shapeIdA.index1;
			r[2] = (uint32_t)e->shapeIdB.index1;
			r[3] = (uint32_t)e->contactId.index1;
			r[4] = e->contactId.generation;
		}
		else
		{
			const b2ContactHitEvent* e = events.hitEvents + ( k - events.beginCount - events.endCount );
			r[0] = 2;
			r[1] = (uint32_t)e->shapeIdA.index1;
			r[2] = (uint32_t)e->shapeIdB.index1;
			r[3] = (uint32_t)e->contactId.index1;
			r[4] = e->contactId.generation;
			r[5] = b2sFloatBits( (float)e->point.x );
			r[6] = b2sFloatBits( (float)e->point.y );
			r[7] = b2sFloatBits( e->normal.x );
			r[8] = b2sFloatBits( e->normal.y );
			r[9] = b2sFloatBits( e->approachSpeed );
		}
		n += 1;
	}
	return n;
}

int b2s_awake_body_count( int handle )
{
	b2sScene* scene = b2sGet( handle );
	return scene == NULL ? 0 : b2World_GetAwakeBodyCount( scene->worldId );
}

int b2s_counters( int handle, int* out, int capacity )
{
	b2sScene* scene = b2sGet( handle );
	if ( scene == NULL || capacity < 30 )
	{
		return 0;
	}
	b2Counters c = b2World_GetCounters( scene->worldId );
	out[0] = c.bodyCount;
	out[1] = c.shapeCount;
	out[2] = c.contactCount;
	out[3] = c.jointCount;
	out[4] = c.islandCount;
	out[5] = c.awakeContactCount;
	for ( int i = 0; i < 24; ++i )
	{
		out[6 + i] = c.colorCounts[i];
	}
	return 30;
}

int b2s_falling_hinges( int workerCount, int stepLimit, int* sleepStep, uint32_t* hash )
{
	int handle = b2s_create( B2S_FALLING_HINGES, workerCount, NULL, 0, NULL, 0 );
	if ( handle < 0 )
	{
		return 0;
	}

	b2sScene* scene = s_scenes + handle;
	uint32_t h = 0;
	int step = 0;
	int slept = 0;
	for ( int i = 0; i < stepLimit; ++i )
	{
		b2World_Step( scene->worldId, 1.0f / 60.0f, 4 );
		b2BodyEvents events = b2World_GetBodyEvents( scene->worldId );
		if ( events.moveCount == 0 )
		{
			h = b2s_hash( handle );
			slept = 1;
			break;
		}
		step += 1;
	}

	*sleepStep = slept ? step : -1;
	*hash = h;
	b2s_destroy( handle );
	return slept;
}
