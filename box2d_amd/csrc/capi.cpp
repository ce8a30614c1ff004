// The Box2D C API entry points (declared in include/box2d/box2d.h) on top of the host world and
// its GPU step. Each function keeps the reference's meaning; the reference implementation is at
// the file:line given in the header next to the declaration.

#include "box2d_amd.h"
#include "device.h"
#include "world.h"

#include <cstring>

namespace b2g
{
World* CreateWorldInternal( const b2WorldDef* def, b2WorldId* outId );
void DestroyWorldInternal( World* world );
void SetLogFcn( b2LogFcn* fcn );
void SetAssertFcn( b2AssertFcn* fcn );
int CollideBatch( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
				  b2LocalManifold* manifolds );
void DeviceSynchronize();
int TimeOfImpactBatch( const b2TOIInput* inputs, int count, b2TOIOutput* outputs );
int ShapeDistanceBatch( const b2DistanceInput* inputs, b2SimplexCache* caches, int count, b2DistanceOutput* outputs );
} // namespace b2g

using namespace b2g;

static World* WorldFromBody( b2BodyId id, Body** bodyOut )
{
	World* world = GetWorldFromIndex( id.world0 );
	if ( world == nullptr )
	{
		return nullptr;
	}
	Body* body = world->GetBody( id );
	if ( body == nullptr )
	{
		return nullptr;
	}
	*bodyOut = body;
	return world;
}

static World* WorldFromShape( b2ShapeId id, Shape** shapeOut )
{
	World* world = GetWorldFromIndex( id.world0 );
	if ( world == nullptr )
	{
		return nullptr;
	}
	int index = id.index1 - 1;
	if ( index < 0 || index >= (int)world->shapes.size() )
	{
		return nullptr;
	}
	Shape* shape = &world->shapes[index];
	if ( shape->id != index || shape->generation != id.generation )
	{
		return nullptr;
	}
	*shapeOut = shape;
	return world;
}

static World* WorldFromJoint( b2JointId id, Joint** jointOut )
{
	World* world = GetWorldFromIndex( id.world0 );
	if ( world == nullptr )
	{
		return nullptr;
	}
	int index = id.index1 - 1;
	if ( index < 0 || index >= (int)world->joints.size() )
	{
		return nullptr;
	}
	Joint* joint = &world->joints[index];
	if ( joint->jointId != index || joint->generation != id.generation )
	{
		return nullptr;
	}
	*jointOut = joint;
	return world;
}

extern "C" {

static int64_t s_byteCount = 0;

void b2SetAllocator( b2AllocFcn* allocFcn, b2FreeFcn* freeFcn )
{
	(void)allocFcn;
	(void)freeFcn;
}

int64_t b2GetByteCount( void )
{
	return s_byteCount;
}

void b2SetAssertFcn( b2AssertFcn* assertFcn )
{
	SetAssertFcn( assertFcn );
}

void b2SetLogFcn( b2LogFcn* logFcn )
{
	SetLogFcn( logFcn );
}

int b2InternalAssert( const char* condition, const char* fileName, int lineNumber )
{
	Fatal( "assertion failed: %s (%s:%d)", condition, fileName, lineNumber );
}

b2WorldId b2CreateWorld( const b2WorldDef* def )
{
	b2WorldId id;
	CreateWorldInternal( def, &id );
	return id;
}

void b2DestroyWorld( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		DestroyWorldInternal( world );
	}
}

bool b2World_IsValid( b2WorldId id )
{
	return GetWorld( id ) != nullptr;
}

void b2World_Step( b2WorldId worldId, float timeStep, int subStepCount )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return;
	}
	world->Step( timeStep, subStepCount );
}

b2BodyEvents b2World_GetBodyEvents( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || world->locked )
	{
		return b2BodyEvents{ nullptr, 0 };
	}
	return world->GetBodyEvents();
}

b2ContactEvents b2World_GetContactEvents( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2ContactEvents{};
	}
	return world->GetContactEvents();
}

void b2World_EnableSleeping( b2WorldId worldId, bool flag )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->EnableSleeping( flag );
	}
}

bool b2World_IsSleepingEnabled( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr && world->enableSleep;
}

void b2World_EnableContinuous( b2WorldId worldId, bool flag )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->enableContinuous = flag;
	}
}

bool b2World_IsContinuousEnabled( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr && world->enableContinuous;
}

void b2World_EnableWarmStarting( b2WorldId worldId, bool flag )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->enableWarmStarting = flag;
	}
}

bool b2World_IsWarmStartingEnabled( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr && world->enableWarmStarting;
}

void b2World_SetGravity( b2WorldId worldId, b2Vec2 gravity )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->gravity = gravity;
	}
}

b2Vec2 b2World_GetGravity( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->gravity : b2Vec2{ 0.0f, 0.0f };
}

void b2World_SetContactTuning( b2WorldId worldId, float hertz, float dampingRatio, float pushSpeed )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr && world->locked == false )
	{
		world->contactHertz = b2ClampFloat( hertz, 0.0f, FLT_MAX );
		world->contactDampingRatio = b2ClampFloat( dampingRatio, 0.0f, FLT_MAX );
		world->contactSpeed = b2ClampFloat( pushSpeed, 0.0f, FLT_MAX );
	}
}

void b2World_SetMaximumLinearSpeed( b2WorldId worldId, float maximumLinearSpeed )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->maxLinearSpeed = maximumLinearSpeed;
	}
}

float b2World_GetMaximumLinearSpeed( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->maxLinearSpeed : 0.0f;
}

int b2World_GetAwakeBodyCount( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? (int)world->solverSets[kAwakeSet].bodyIds.size() : 0;
}

b2Profile b2World_GetProfile( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->profile : b2Profile{};
}

b2Counters b2World_GetCounters( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->GetCounters() : b2Counters{};
}

void b2World_SetUserData( b2WorldId worldId, void* userData )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->userData = userData;
	}
}

void* b2World_GetUserData( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->userData : nullptr;
}

// Bodies ---------------------------------------------------------------------------------------

b2BodyId b2CreateBody( b2WorldId worldId, const b2BodyDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullBodyId;
	}
	return world->CreateBody( def );
}

bool b2Body_IsValid( b2BodyId id )
{
	Body* body;
	return WorldFromBody( id, &body ) != nullptr;
}

b2BodyType b2Body_GetType( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? body->type : b2_staticBody;
}

b2WorldTransform b2Body_GetTransform( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr )
	{
		return b2Transform_identity;
	}
	world->PullBodies();
	return world->bodySims[body->id].transform;
}

b2Pos b2Body_GetPosition( b2BodyId bodyId )
{
	return b2Body_GetTransform( bodyId ).p;
}

b2Rot b2Body_GetRotation( b2BodyId bodyId )
{
	return b2Body_GetTransform( bodyId ).q;
}

b2Pos b2Body_GetWorldCenterOfMass( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr )
	{
		return b2Pos_zero;
	}
	world->PullBodies();
	return world->bodySims[body->id].center;
}

b2Vec2 b2Body_GetLocalCenterOfMass( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr )
	{
		return b2Vec2_zero;
	}
	world->PullBodies();
	return world->bodySims[body->id].localCenter;
}

b2Vec2 b2Body_GetLinearVelocity( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr || body->setIndex != kAwakeSet )
	{
		return b2Vec2_zero;
	}
	world->PullBodies();
	return world->bodyStates[body->id].linearVelocity;
}

float b2Body_GetAngularVelocity( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr || body->setIndex != kAwakeSet )
	{
		return 0.0f;
	}
	world->PullBodies();
	return world->bodyStates[body->id].angularVelocity;
}

void b2Body_SetLinearVelocity( b2BodyId bodyId, b2Vec2 linearVelocity )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr || body->type == b2_staticBody )
	{
		return;
	}
	world->PullBodies();
	if ( b2LengthSquared( linearVelocity ) > 0.0f && body->setIndex >= kFirstSleepingSet )
	{
		world->WakeBodyPublic( body );
	}
	if ( body->setIndex != kAwakeSet )
	{
		return;
	}
	world->bodyStates[body->id].linearVelocity = linearVelocity;
	world->bodiesDirty = true;
}

void b2Body_SetAngularVelocity( b2BodyId bodyId, float angularVelocity )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr || body->type == b2_staticBody )
	{
		return;
	}
	world->PullBodies();
	if ( body->setIndex != kAwakeSet )
	{
		return;
	}
	world->bodyStates[body->id].angularVelocity = angularVelocity;
	world->bodiesDirty = true;
}

float b2Body_GetMass( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? body->mass : 0.0f;
}

float b2Body_GetRotationalInertia( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? body->inertia : 0.0f;
}

bool b2Body_IsAwake( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr && body->setIndex == kAwakeSet;
}

int b2Body_GetShapeCount( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? body->shapeCount : 0;
}

void b2Body_SetUserData( b2BodyId bodyId, void* userData )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world != nullptr )
	{
		body->userData = userData;
	}
}

void* b2Body_GetUserData( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? body->userData : nullptr;
}

b2WorldId b2Body_GetWorld( b2BodyId bodyId )
{
	World* world = GetWorldFromIndex( bodyId.world0 );
	if ( world == nullptr )
	{
		return b2_nullWorldId;
	}
	return b2WorldId{ (uint16_t)( bodyId.world0 + 1 ), world->generation };
}

// Shapes ---------------------------------------------------------------------------------------

static b2ShapeId CreateShapeEntry( b2BodyId bodyId, const b2ShapeDef* def, const void* geometry, b2ShapeType type )
{
	World* world = GetWorldFromIndex( bodyId.world0 );
	if ( world == nullptr )
	{
		return b2_nullShapeId;
	}
	return world->CreateShape( bodyId, def, geometry, type );
}

b2ShapeId b2CreateCircleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Circle* circle )
{
	return CreateShapeEntry( bodyId, def, circle, b2_circleShape );
}

b2ShapeId b2CreateCapsuleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Capsule* capsule )
{
	float lengthSqr = b2DistanceSquared( capsule->center1, capsule->center2 );
	if ( lengthSqr <= B2_LINEAR_SLOP * B2_LINEAR_SLOP )
	{
		return b2_nullShapeId;
	}
	return CreateShapeEntry( bodyId, def, capsule, b2_capsuleShape );
}

b2ShapeId b2CreatePolygonShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Polygon* polygon )
{
	return CreateShapeEntry( bodyId, def, polygon, b2_polygonShape );
}

b2ShapeId b2CreateSegmentShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Segment* segment )
{
	float lengthSqr = b2DistanceSquared( segment->point1, segment->point2 );
	if ( lengthSqr <= B2_LINEAR_SLOP * B2_LINEAR_SLOP )
	{
		return b2_nullShapeId;
	}
	return CreateShapeEntry( bodyId, def, segment, b2_segmentShape );
}

bool b2Shape_IsValid( b2ShapeId id )
{
	Shape* shape;
	return WorldFromShape( id, &shape ) != nullptr;
}

b2ShapeType b2Shape_GetType( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? shape->type : b2_circleShape;
}

b2BodyId b2Shape_GetBody( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world == nullptr )
	{
		return b2_nullBodyId;
	}
	Body& body = world->bodies[shape->bodyId];
	return b2BodyId{ body.id + 1, shapeId.world0, body.generation };
}

b2AABB b2Shape_GetAABB( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world == nullptr )
	{
		return b2AABB{};
	}
	world->device->DownloadShapes( *world );
	return world->shapeData[shape->id].aabb;
}

b2ChainId b2CreateChain( b2BodyId bodyId, const b2ChainDef* def )
{
	World* world = GetWorldFromIndex( bodyId.world0 );
	if ( world == nullptr )
	{
		return b2_nullChainId;
	}
	return world->CreateChain( bodyId, def );
}

// Joints ---------------------------------------------------------------------------------------

b2JointId b2CreateRevoluteJoint( b2WorldId worldId, const b2RevoluteJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreateRevoluteJoint( def );
}

b2JointId b2CreateDistanceJoint( b2WorldId worldId, const b2DistanceJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreateDistanceJoint( def );
}

b2JointId b2CreateFilterJoint( b2WorldId worldId, const b2FilterJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreateFilterJoint( def );
}

b2JointId b2CreateMotorJoint( b2WorldId worldId, const b2MotorJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreateMotorJoint( def );
}

b2JointId b2CreateMoverJoint( b2WorldId worldId, const b2MoverJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreateMoverJoint( def );
}

b2JointId b2CreatePogoJoint( b2WorldId worldId, const b2PogoJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreatePogoJoint( def );
}

b2JointId b2CreatePrismaticJoint( b2WorldId worldId, const b2PrismaticJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreatePrismaticJoint( def );
}

b2JointId b2CreateWeldJoint( b2WorldId worldId, const b2WeldJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreateWeldJoint( def );
}

b2JointId b2CreateWheelJoint( b2WorldId worldId, const b2WheelJointDef* def )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return b2_nullJointId;
	}
	return world->CreateWheelJoint( def );
}

bool b2Joint_IsValid( b2JointId id )
{
	Joint* joint;
	return WorldFromJoint( id, &joint ) != nullptr;
}

b2JointType b2Joint_GetType( b2JointId jointId )
{
	Joint* joint;
	World* world = WorldFromJoint( jointId, &joint );
	return world != nullptr ? joint->type : b2_distanceJoint;
}

b2BodyId b2Joint_GetBodyA( b2JointId jointId )
{
	Joint* joint;
	World* world = WorldFromJoint( jointId, &joint );
	if ( world == nullptr )
	{
		return b2_nullBodyId;
	}
	Body& body = world->bodies[joint->edges[0].bodyId];
	return b2BodyId{ body.id + 1, jointId.world0, body.generation };
}

b2BodyId b2Joint_GetBodyB( b2JointId jointId )
{
	Joint* joint;
	World* world = WorldFromJoint( jointId, &joint );
	if ( world == nullptr )
	{
		return b2_nullBodyId;
	}
	Body& body = world->bodies[joint->edges[1].bodyId];
	return b2BodyId{ body.id + 1, jointId.world0, body.generation };
}

// Extensions (include/box2d_amd.h) -----------------------------------------------------------

void b2amd_World_EnableKernelTiming( b2WorldId worldId, bool flag )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->device->EnableKernelTiming( flag );
	}
}

int b2amd_World_GetKernelStats( b2WorldId worldId, b2amd_KernelStats* stats, int capacity )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || stats == nullptr )
	{
		return 0;
	}
	return world->device->GetKernelStats( stats, capacity );
}

void b2amd_World_ResetKernelStats( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->device->ResetKernelStats();
	}
}

void b2amd_World_SetSolverPath( b2WorldId worldId, b2amd_SolverPath path )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->device->SetSolverPath( (int)path );
	}
}

void b2amd_World_SetPairQuery( b2WorldId worldId, b2amd_PairQuery mode )
{
	World* world = GetWorld( worldId );
	if ( world != nullptr )
	{
		world->pairQuery = (int)mode;
	}
}

int b2amd_World_GetLastPairQuery( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->lastPairQueryPath : 0;
}

int b2amd_TimeOfImpactBatch( const b2TOIInput* inputs, int count, b2TOIOutput* outputs )
{
	return TimeOfImpactBatch( inputs, count, outputs );
}

int b2amd_ShapeDistanceBatch( const b2DistanceInput* inputs, b2SimplexCache* caches, int count, b2DistanceOutput* outputs )
{
	return ShapeDistanceBatch( inputs, caches, count, outputs );
}

int b2amd_StepWorlds( const b2WorldId* worldIds, int count, float timeStep, int subStepCount, b2amd_BatchStats* stats )
{
	std::vector<World*> worlds;
	worlds.reserve( count );
	for ( int i = 0; i < count; ++i )
	{
		World* world = GetWorld( worldIds[i] );
		if ( world != nullptr )
		{
			worlds.push_back( world );
		}
	}
	BatchStats bs = {};
	int n = StepWorlds( worlds.data(), (int)worlds.size(), timeStep, subStepCount, &bs );
	if ( stats != nullptr )
	{
		stats->solveKernelMs = bs.solveKernelMs;
		stats->solveThreads = bs.solveThreads;
		stats->solveGrid = bs.solveGrid;
		stats->worlds = bs.worlds;
		for ( int k = 0; k < 5; ++k )
		{
			stats->phaseMs[k] = bs.phaseMs[k];
		}
		stats->solveAlgorithmicBytes = bs.solveBytes;
	}
	return n;
}

void b2amd_Synchronize( void )
{
	DeviceSynchronize();
}

int b2amd_CollideBatch( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
						b2LocalManifold* manifolds )
{
	return CollideBatch( shapesA, shapesB, transforms, count, manifolds );
}

} // extern "C"

// Layout checks: definitions filled by reference-compiled code must read field for field.
static_assert( sizeof( b2WorldDef ) == 112, "b2WorldDef layout" );
static_assert( offsetof( b2WorldDef, enableSleep ) == 48, "b2WorldDef layout" );
static_assert( offsetof( b2WorldDef, workerCount ) == 52, "b2WorldDef layout" );
static_assert( offsetof( b2WorldDef, internalValue ) == 108, "b2WorldDef layout" );
static_assert( sizeof( b2BodyDef ) == 88, "b2BodyDef layout" );
static_assert( offsetof( b2BodyDef, name ) == 56, "b2BodyDef layout" );
static_assert( offsetof( b2BodyDef, motionLocks ) == 72, "b2BodyDef layout" );
static_assert( sizeof( b2ShapeDef ) == 88, "b2ShapeDef layout" );
static_assert( offsetof( b2ShapeDef, density ) == 40, "b2ShapeDef layout" );
static_assert( offsetof( b2ShapeDef, filter ) == 48, "b2ShapeDef layout" );
static_assert( sizeof( b2SurfaceMaterial ) == 32, "b2SurfaceMaterial layout" );
static_assert( sizeof( b2Filter ) == 24, "b2Filter layout" );
static_assert( sizeof( b2ChainDef ) == 72, "b2ChainDef layout" );
static_assert( sizeof( b2Polygon ) == 144, "b2Polygon layout" );
static_assert( sizeof( b2Hull ) == 68, "b2Hull layout" );
static_assert( sizeof( b2RevoluteJointDef ) == 128, "b2RevoluteJointDef layout" );
static_assert( offsetof( b2RevoluteJointDef, targetAngle ) == 80, "b2RevoluteJointDef layout" );
static_assert( sizeof( b2DistanceJointDef ) == 136 && offsetof( b2DistanceJointDef, motorSpeed ) == 124, "b2DistanceJointDef layout" );
static_assert( sizeof( b2FilterJointDef ) == 88, "b2FilterJointDef layout" );
static_assert( sizeof( b2MotorJointDef ) == 128 && offsetof( b2MotorJointDef, maxSpringTorque ) == 120, "b2MotorJointDef layout" );
static_assert( sizeof( b2PrismaticJointDef ) == 128 && offsetof( b2PrismaticJointDef, motorSpeed ) == 116, "b2PrismaticJointDef layout" );
static_assert( sizeof( b2WeldJointDef ) == 104 && offsetof( b2WeldJointDef, angularDampingRatio ) == 92, "b2WeldJointDef layout" );
static_assert( sizeof( b2MoverJointDef ) == 104 && offsetof( b2MoverJointDef, maxVelocityForce ) == 88, "b2MoverJointDef layout" );
static_assert( sizeof( b2PogoJointDef ) == 120 && offsetof( b2PogoJointDef, velocity ) == 112, "b2PogoJointDef layout" );
static_assert( sizeof( b2WheelJointDef ) == 120 && offsetof( b2WheelJointDef, motorSpeed ) == 112, "b2WheelJointDef layout" );
static_assert( sizeof( b2JointDef ) == 80, "b2JointDef layout" );
static_assert( offsetof( b2JointDef, localFrameA ) == 24, "b2JointDef layout" );
static_assert( sizeof( b2Profile ) == 92, "b2Profile layout" );
static_assert( sizeof( b2Counters ) == 152, "b2Counters layout" );
static_assert( sizeof( b2BodyMoveEvent ) == 40, "b2BodyMoveEvent layout" );
static_assert( sizeof( b2Manifold ) == 104, "b2Manifold layout" );
static_assert( sizeof( b2ManifoldPoint ) == 44, "b2ManifoldPoint layout" );
static_assert( sizeof( b2ContactHitEvent ) == 48, "b2ContactHitEvent layout" );
static_assert( sizeof( BodyState ) == 32, "body state must stay 32 bytes for two 16-byte loads" );
