// The Box2D C API entry points (declared in include/box2d/box2d.h) on top of the host world and
// its GPU step. Each function keeps the reference's meaning; the reference implementation is at
// the file:line given in the header next to the declaration.

#include "box2d_amd.h"
#include "device.h"
#include "world.h"

#include <cfloat>
#include <cstring>

namespace b2g
{
World* CreateWorldInternal( const b2WorldDef* def, b2WorldId* outId );
void DestroyWorldInternal( World* world );
void SetLogFcn( b2LogFcn* fcn );
void SetAssertFcn( b2AssertFcn* fcn );
int CollideBatch( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
				  b2LocalManifold* manifolds );
int CollideBatchNoAlloc( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
						 b2LocalManifold* manifolds );
void DeviceSynchronize();
bool SpinWaitActive();
bool SetBatchSolverThreads( int threads );
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

// Snapshots (reference src/world_snapshot.c:477-590, 899-970; this library's own image layout).
int b2World_Snapshot( b2WorldId worldId, uint8_t* image, int capacity )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->Snapshot( image, capacity ) : 0;
}

bool b2World_Restore( b2WorldId worldId, const uint8_t* image, int size )
{
	World* world = GetWorld( worldId );
	return world != nullptr && world->Restore( image, size );
}

b2WorldId b2CreateWorldFromSnapshot( const uint8_t* image, int size, int workerCount )
{
	b2WorldDef def = b2DefaultWorldDef();
	def.workerCount = workerCount;
	b2WorldId id = b2CreateWorld( &def );
	World* world = GetWorld( id );
	if ( world == nullptr || world->Restore( image, size ) == false )
	{
		if ( world != nullptr )
		{
			DestroyWorldInternal( world );
		}
		return b2_nullWorldId;
	}
	return id;
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

b2SensorEvents b2World_GetSensorEvents( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->GetSensorEvents() : b2SensorEvents{};
}

b2JointEvents b2World_GetJointEvents( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || world->locked )
	{
		return b2JointEvents{};
	}
	return b2JointEvents{ world->jointEvents.data(), (int)world->jointEvents.size() };
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

uint64_t b2World_GetStateHash( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || world->locked )
	{
		return 0;
	}
	return world->StateHash();
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

// reference src/physics_world.c:1789-1827
void b2World_SetRestitutionThreshold( b2WorldId worldId, float value )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || world->locked )
	{
		return;
	}
	world->restitutionThreshold = b2ClampFloat( value, 0.0f, FLT_MAX );
}

float b2World_GetRestitutionThreshold( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->restitutionThreshold : 0.0f;
}

void b2World_SetHitEventThreshold( b2WorldId worldId, float value )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || world->locked )
	{
		return;
	}
	world->hitEventThreshold = b2ClampFloat( value, 0.0f, FLT_MAX );
}

float b2World_GetHitEventThreshold( b2WorldId worldId )
{
	World* world = GetWorld( worldId );
	return world != nullptr ? world->hitEventThreshold : 0.0f;
}

// reference src/physics_world.c:2972-2982
void b2World_SetCustomFilterCallback( b2WorldId worldId, b2CustomFilterFcn* fcn, void* context )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr )
	{
		return;
	}
	world->customFilterFcn = fcn;
	world->customFilterContext = context;
}

// reference src/physics_world.c:1950-1984. A change of callback takes effect for each contact at
// its next full narrowphase update, as in the reference, where the callback runs in that update.
void b2World_SetPreSolveCallback( b2WorldId worldId, b2PreSolveFcn* preSolveFcn, b2PreContinuousFcn* preContinuousFcn,
								  void* context )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || world->locked )
	{
		return;
	}
	world->preSolveFcn = preSolveFcn;
	world->preContinuousFcn = preContinuousFcn;
	world->preSolveContext = context;
}

void b2World_SetFrictionCallback( b2WorldId worldId, b2FrictionCallback* callback )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || world->locked )
	{
		return;
	}
	if ( callback != world->frictionCallback )
	{
		world->frictionCallback = callback;
		world->mixingStale = true;
	}
}

void b2World_SetRestitutionCallback( b2WorldId worldId, b2RestitutionCallback* callback )
{
	World* world = GetWorld( worldId );
	if ( world == nullptr || world->locked )
	{
		return;
	}
	if ( callback != world->restitutionCallback )
	{
		world->restitutionCallback = callback;
		world->mixingStale = true;
	}
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

void b2DestroyBody( b2BodyId bodyId )
{
	World* world = GetWorldFromIndex( bodyId.world0 );
	if ( world != nullptr )
	{
		world->DestroyBody( bodyId );
	}
}

b2Vec2 b2Body_GetLocalPoint( b2BodyId bodyId, b2Pos worldPoint )
{
	return b2InvTransformWorldPoint( b2Body_GetTransform( bodyId ), worldPoint );
}

b2Pos b2Body_GetWorldPoint( b2BodyId bodyId, b2Vec2 localPoint )
{
	return b2TransformWorldPoint( b2Body_GetTransform( bodyId ), localPoint );
}

b2Vec2 b2Body_GetLocalVector( b2BodyId bodyId, b2Vec2 worldVector )
{
	return b2InvRotateVector( b2Body_GetTransform( bodyId ).q, worldVector );
}

b2Vec2 b2Body_GetWorldVector( b2BodyId bodyId, b2Vec2 localVector )
{
	return b2RotateVector( b2Body_GetTransform( bodyId ).q, localVector );
}

void b2Body_SetTransform( b2BodyId bodyId, b2Pos position, b2Rot rotation )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr || world->locked )
	{
		return;
	}
	world->SetBodyTransform( body, position, rotation );
}

// Velocity that carries the body to the target in one step (reference src/body.c:842-893).
void b2Body_SetTargetTransform( b2BodyId bodyId, b2WorldTransform target, float timeStep, bool wake )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr || body->setIndex == kDisabledSet || body->type == b2_staticBody || timeStep <= 0.0f )
	{
		return;
	}
	if ( body->setIndex != kAwakeSet && wake == false )
	{
		return;
	}
	world->PullBodies();
	const BodySim& sim = world->bodySims[body->id];
	b2Vec2 delta = b2SubPos( b2TransformWorldPoint( target, sim.localCenter ), sim.center );
	float invTimeStep = 1.0f / timeStep;
	b2Vec2 linearVelocity = b2MulSV( invTimeStep, delta );
	float deltaAngle = b2RelativeAngle( sim.transform.q, target.q );
	float angularVelocity = invTimeStep * deltaAngle;

	if ( body->setIndex != kAwakeSet )
	{
		float maxVelocity = b2Length( linearVelocity ) + b2AbsFloat( angularVelocity ) * sim.maxExtent;
		if ( maxVelocity < world->bodyAux[body->id].sleepThreshold )
		{
			return;
		}
		world->WakeBodyPublic( body );
	}
	BodyState& state = world->bodyStates[body->id];
	state.linearVelocity = linearVelocity;
	state.angularVelocity = angularVelocity;
	world->bodiesDirty = true;
}

// Shared prologue of the force and impulse calls (reference src/body.c:935-1095): dynamic,
// enabled bodies only; optionally wake a sleeping body; act only on awake bodies.
static World* AwakeDynamicBody( b2BodyId bodyId, bool wake, Body** bodyOut )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr || body->type != b2_dynamicBody || body->setIndex == kDisabledSet )
	{
		return nullptr;
	}
	world->PullBodies();
	if ( wake && body->setIndex >= kFirstSleepingSet )
	{
		world->WakeBodyPublic( body );
	}
	if ( body->setIndex != kAwakeSet )
	{
		return nullptr;
	}
	world->bodiesDirty = true;
	*bodyOut = body;
	return world;
}

static void LimitVelocity( BodyState* state, float maxLinearSpeed )
{
	float v2 = b2LengthSquared( state->linearVelocity );
	if ( v2 > maxLinearSpeed * maxLinearSpeed )
	{
		state->linearVelocity = b2MulSV( maxLinearSpeed / sqrtf( v2 ), state->linearVelocity );
	}
}

void b2Body_ApplyForce( b2BodyId bodyId, b2Vec2 force, b2Pos point, bool wake )
{
	Body* body;
	World* world = AwakeDynamicBody( bodyId, wake, &body );
	if ( world != nullptr )
	{
		BodySim& sim = world->bodySims[body->id];
		sim.force = b2Add( sim.force, force );
		sim.torque += b2Cross( b2SubPos( point, sim.center ), force );
	}
}

void b2Body_ApplyForceToCenter( b2BodyId bodyId, b2Vec2 force, bool wake )
{
	Body* body;
	World* world = AwakeDynamicBody( bodyId, wake, &body );
	if ( world != nullptr )
	{
		BodySim& sim = world->bodySims[body->id];
		sim.force = b2Add( sim.force, force );
	}
}

void b2Body_ApplyTorque( b2BodyId bodyId, float torque, bool wake )
{
	Body* body;
	World* world = AwakeDynamicBody( bodyId, wake, &body );
	if ( world != nullptr )
	{
		world->bodySims[body->id].torque += torque;
	}
}

void b2Body_ApplyLinearImpulse( b2BodyId bodyId, b2Vec2 impulse, b2Pos point, bool wake )
{
	Body* body;
	World* world = AwakeDynamicBody( bodyId, wake, &body );
	if ( world != nullptr )
	{
		const BodySim& sim = world->bodySims[body->id];
		BodyState& state = world->bodyStates[body->id];
		state.linearVelocity = b2MulAdd( state.linearVelocity, sim.invMass, impulse );
		state.angularVelocity += sim.invInertia * b2Cross( b2SubPos( point, sim.center ), impulse );
		LimitVelocity( &state, world->maxLinearSpeed );
	}
}

void b2Body_ApplyLinearImpulseToCenter( b2BodyId bodyId, b2Vec2 impulse, bool wake )
{
	Body* body;
	World* world = AwakeDynamicBody( bodyId, wake, &body );
	if ( world != nullptr )
	{
		const BodySim& sim = world->bodySims[body->id];
		BodyState& state = world->bodyStates[body->id];
		state.linearVelocity = b2MulAdd( state.linearVelocity, sim.invMass, impulse );
		LimitVelocity( &state, world->maxLinearSpeed );
	}
}

void b2Body_ApplyAngularImpulse( b2BodyId bodyId, float impulse, bool wake )
{
	Body* body;
	World* world = AwakeDynamicBody( bodyId, wake, &body );
	if ( world != nullptr )
	{
		world->bodyStates[body->id].angularVelocity += world->bodySims[body->id].invInertia * impulse;
	}
}

void b2Body_ApplyMassFromShapes( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr || world->locked )
	{
		return;
	}
	world->ApplyMassFromShapes( body );
}

// Damping, gravity scale and sleep settings (reference src/body.c:1466-1550, 1629-1696)
static World* UnlockedBody( b2BodyId bodyId, Body** body )
{
	World* world = WorldFromBody( bodyId, body );
	return world != nullptr && world->locked == false ? world : nullptr;
}

void b2Body_SetLinearDamping( b2BodyId bodyId, float linearDamping )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->PullBodies();
		world->bodySims[body->id].linearDamping = linearDamping;
		world->bodiesDirty = true;
	}
}

float b2Body_GetLinearDamping( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? world->bodySims[body->id].linearDamping : 0.0f;
}

void b2Body_SetAngularDamping( b2BodyId bodyId, float angularDamping )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->PullBodies();
		world->bodySims[body->id].angularDamping = angularDamping;
		world->bodiesDirty = true;
	}
}

float b2Body_GetAngularDamping( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? world->bodySims[body->id].angularDamping : 0.0f;
}

void b2Body_SetGravityScale( b2BodyId bodyId, float gravityScale )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->PullBodies();
		world->bodySims[body->id].gravityScale = gravityScale;
		world->bodiesDirty = true;
	}
}

float b2Body_GetGravityScale( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? world->bodySims[body->id].gravityScale : 0.0f;
}

void b2Body_SetSleepThreshold( b2BodyId bodyId, float sleepThreshold )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->PullBodies();
		world->bodyAux[body->id].sleepThreshold = sleepThreshold;
		world->bodiesDirty = true;
	}
}

float b2Body_GetSleepThreshold( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? world->bodyAux[body->id].sleepThreshold : 0.0f;
}

void b2Body_EnableSleep( b2BodyId bodyId, bool enableSleep )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->EnableBodySleep( body, enableSleep );
	}
}

bool b2Body_IsSleepEnabled( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr && ( body->flags & kEnableSleep ) != 0;
}

void b2Body_SetBullet( b2BodyId bodyId, bool flag )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->SetBodyBullet( body, flag );
	}
}

bool b2Body_IsBullet( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr && ( body->flags & kIsBullet ) != 0;
}

void b2Body_SetAwake( b2BodyId bodyId, bool awake )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->SetBodyAwake( body, awake );
	}
}

void b2Body_WakeTouching( b2BodyId bodyId )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->PullBodies();
		world->WakeTouching( body );
	}
}

// Contact data (reference src/body.c:436-486, src/shape.c:1749-1800, src/contact.c:45-92)
static World* WorldFromContact( b2ContactId id, Contact** out )
{
	World* world = GetWorldFromIndex( id.world0 );
	if ( world == nullptr )
	{
		return nullptr;
	}
	int index = id.index1 - 1;
	if ( index < 0 || index >= (int)world->contacts.size() )
	{
		return nullptr;
	}
	Contact* contact = &world->contacts[index];
	if ( contact->contactId != index || contact->generation != id.generation )
	{
		return nullptr;
	}
	*out = contact;
	return world;
}

static b2ContactData ContactData( World* world, const Contact& contact )
{
	b2ContactData data = {};
	data.contactId = world->ContactIdOf( contact );
	data.shapeIdA = world->ShapeIdOf( contact.shapeIdA );
	data.shapeIdB = world->ShapeIdOf( contact.shapeIdB );
	data.manifold = world->PulledContact( contact.contactId ).manifold;
	return data;
}

bool b2Contact_IsValid( b2ContactId id )
{
	Contact* contact;
	return WorldFromContact( id, &contact ) != nullptr;
}

b2ContactData b2Contact_GetData( b2ContactId contactId )
{
	Contact* contact;
	World* world = WorldFromContact( contactId, &contact );
	return world != nullptr ? ContactData( world, *contact ) : b2ContactData{};
}

int b2Body_GetContactCapacity( b2BodyId bodyId )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	return world != nullptr ? body->contactCount : 0;
}

int b2Body_GetContactData( b2BodyId bodyId, b2ContactData* contactData, int capacity )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world == nullptr )
	{
		return 0;
	}
	int contactKey = body->headContactKey;
	int index = 0;
	while ( contactKey != B2_NULL_INDEX && index < capacity )
	{
		const Contact& contact = world->contacts[contactKey >> 1];
		if ( contact.flags & kContactTouching )
		{
			contactData[index++] = ContactData( world, contact );
		}
		contactKey = contact.edges[contactKey & 1].nextKey;
	}
	return index;
}

// Type changes and enabling (reference src/body.c:1134-1316, 1699-1858)
void b2Body_SetType( b2BodyId bodyId, b2BodyType type )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->SetBodyType( body, type );
	}
}

void b2Body_Disable( b2BodyId bodyId )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->DisableBody( body );
	}
}

void b2Body_Enable( b2BodyId bodyId )
{
	Body* body;
	World* world = UnlockedBody( bodyId, &body );
	if ( world != nullptr )
	{
		world->EnableBody( body );
	}
}

bool b2Body_IsEnabled( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr && body->setIndex != kDisabledSet;
}

int b2Body_GetShapes( b2BodyId bodyId, b2ShapeId* shapeArray, int capacity )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr )
	{
		return 0;
	}
	int shapeId = body->headShapeId;
	int shapeCount = 0;
	while ( shapeId != B2_NULL_INDEX && shapeCount < capacity )
	{
		const Shape& shape = world->shapes[shapeId];
		shapeArray[shapeCount] = b2ShapeId{ shape.id + 1, bodyId.world0, shape.generation };
		shapeCount += 1;
		shapeId = shape.nextShapeId;
	}
	return shapeCount;
}

int b2Body_GetJointCount( b2BodyId bodyId )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	return world != nullptr ? body->jointCount : 0;
}

int b2Body_GetJoints( b2BodyId bodyId, b2JointId* jointArray, int capacity )
{
	Body* body;
	World* world = WorldFromBody( bodyId, &body );
	if ( world == nullptr )
	{
		return 0;
	}
	int jointKey = body->headJointKey;
	int jointCount = 0;
	while ( jointKey != B2_NULL_INDEX && jointCount < capacity )
	{
		const Joint& joint = world->joints[jointKey >> 1];
		jointArray[jointCount] = b2JointId{ joint.jointId + 1, bodyId.world0, joint.generation };
		jointCount += 1;
		jointKey = joint.edges[jointKey & 1].nextKey;
	}
	return jointCount;
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

void b2Shape_EnableSensorEvents( b2ShapeId shapeId, bool flag )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world != nullptr && world->locked == false && shape->enableSensorEvents != flag )
	{
		world->PullShapes();
		shape->enableSensorEvents = flag;
		uint32_t& flags = world->shapeData[shape->id].flags;
		flags = flag ? ( flags | kShapeEnableSensorEvents ) : ( flags & ~kShapeEnableSensorEvents );
		world->shapesDirty = true;
	}
}

int b2Shape_GetSensorCapacity( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world == nullptr || world->locked || shape->sensorIndex == B2_NULL_INDEX )
	{
		return 0;
	}
	return (int)world->sensors[shape->sensorIndex].overlaps2.size();
}

int b2Shape_GetSensorData( b2ShapeId shapeId, b2ShapeId* visitorIds, int capacity )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world == nullptr || world->locked || shape->sensorIndex == B2_NULL_INDEX )
	{
		return 0;
	}
	const std::vector<SensorVisitor>& overlaps = world->sensors[shape->sensorIndex].overlaps2;
	int count = std::min( (int)overlaps.size(), capacity );
	for ( int i = 0; i < count; ++i )
	{
		visitorIds[i] = b2ShapeId{ overlaps[i].shapeId + 1, shapeId.world0, overlaps[i].generation };
	}
	return count;
}

bool b2Shape_AreSensorEventsEnabled( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr && shape->enableSensorEvents;
}

void b2Shape_EnablePreSolveEvents( b2ShapeId shapeId, bool flag )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world != nullptr && world->locked == false )
	{
		shape->enablePreSolveEvents = flag;
	}
}

bool b2Shape_ArePreSolveEventsEnabled( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr && shape->enablePreSolveEvents;
}

// Materials, density, filter and user data (reference src/shape.c:1190-1316, 1367-1391)
static World* UnlockedShape( b2ShapeId shapeId, Shape** shape )
{
	World* world = WorldFromShape( shapeId, shape );
	return world != nullptr && world->locked == false ? world : nullptr;
}

void b2Shape_SetFriction( b2ShapeId shapeId, float friction )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	if ( world != nullptr )
	{
		b2SurfaceMaterial m = shape->material;
		m.friction = friction;
		world->SetShapeMaterial( shape, m );
	}
}

float b2Shape_GetFriction( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? shape->material.friction : 0.0f;
}

void b2Shape_SetRestitution( b2ShapeId shapeId, float restitution )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	if ( world != nullptr )
	{
		b2SurfaceMaterial m = shape->material;
		m.restitution = restitution;
		world->SetShapeMaterial( shape, m );
	}
}

float b2Shape_GetRestitution( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? shape->material.restitution : 0.0f;
}

void b2Shape_SetUserMaterial( b2ShapeId shapeId, uint64_t material )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	if ( world != nullptr )
	{
		b2SurfaceMaterial m = shape->material;
		m.userMaterialId = material;
		world->SetShapeMaterial( shape, m );
	}
}

uint64_t b2Shape_GetUserMaterial( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? shape->material.userMaterialId : 0;
}

void b2Shape_SetSurfaceMaterial( b2ShapeId shapeId, const b2SurfaceMaterial* surfaceMaterial )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	if ( world != nullptr )
	{
		world->SetShapeMaterial( shape, *surfaceMaterial );
	}
}

b2SurfaceMaterial b2Shape_GetSurfaceMaterial( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? shape->material : b2SurfaceMaterial{};
}

void b2Shape_SetDensity( b2ShapeId shapeId, float density, bool updateBodyMass )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	if ( world != nullptr )
	{
		world->SetShapeDensity( shape, density, updateBodyMass );
	}
}

float b2Shape_GetDensity( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? shape->density : 0.0f;
}

void b2Shape_SetFilter( b2ShapeId shapeId, b2Filter filter )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	if ( world != nullptr )
	{
		world->SetShapeFilter( shape, filter );
	}
}

b2Filter b2Shape_GetFilter( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? shape->filter : b2Filter{};
}

void b2Shape_SetUserData( b2ShapeId shapeId, void* userData )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world != nullptr )
	{
		shape->userData = userData;
	}
}

void* b2Shape_GetUserData( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? shape->userData : nullptr;
}

void b2DestroyShape( b2ShapeId shapeId, bool updateBodyMass )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	// a segment of a chain goes with its chain (reference src/shape.c:413-418)
	if ( world == nullptr || ( shape->type == b2_chainSegmentShape && world->shapeData[shape->id].chainSegment.chainId != B2_NULL_INDEX ) )
	{
		return;
	}
	world->DestroyShape( shape, updateBodyMass );
}

int b2Shape_GetContactCapacity( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	return world != nullptr && shape->sensorIndex == B2_NULL_INDEX ? world->bodies[shape->bodyId].contactCount : 0;
}

int b2Shape_GetContactData( b2ShapeId shapeId, b2ContactData* contactData, int capacity )
{
	Shape* shape;
	World* world = UnlockedShape( shapeId, &shape );
	if ( world == nullptr || shape->sensorIndex != B2_NULL_INDEX )
	{
		return 0;
	}
	const Body& body = world->bodies[shape->bodyId];
	int contactKey = body.headContactKey;
	int index = 0;
	while ( contactKey != B2_NULL_INDEX && index < capacity )
	{
		const Contact& contact = world->contacts[contactKey >> 1];
		if ( ( contact.shapeIdA == shape->id || contact.shapeIdB == shape->id ) && ( contact.flags & kContactTouching ) )
		{
			contactData[index++] = ContactData( world, contact );
		}
		contactKey = contact.edges[contactKey & 1].nextKey;
	}
	return index;
}

b2WorldId b2Shape_GetWorld( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr ? b2WorldId{ (uint16_t)( shapeId.world0 + 1 ), world->generation } : b2_nullWorldId;
}

bool b2Shape_IsSensor( b2ShapeId shapeId )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	return world != nullptr && shape->sensorIndex != B2_NULL_INDEX;
}

// Geometry accessors (reference src/shape.c:1489-1604). A shape of another type returns a
// zeroed value where the reference asserts.
extern "C++" template <typename T>
static T ShapeGeometry( b2ShapeId shapeId, b2ShapeType type, T ShapeDev::*member )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world == nullptr || shape->type != type )
	{
		return T{};
	}
	return world->shapeData[shape->id].*member;
}

b2Circle b2Shape_GetCircle( b2ShapeId shapeId )
{
	return ShapeGeometry( shapeId, b2_circleShape, &ShapeDev::circle );
}

b2Segment b2Shape_GetSegment( b2ShapeId shapeId )
{
	return ShapeGeometry( shapeId, b2_segmentShape, &ShapeDev::segment );
}

b2Capsule b2Shape_GetCapsule( b2ShapeId shapeId )
{
	return ShapeGeometry( shapeId, b2_capsuleShape, &ShapeDev::capsule );
}

b2Polygon b2Shape_GetPolygon( b2ShapeId shapeId )
{
	return ShapeGeometry( shapeId, b2_polygonShape, &ShapeDev::polygon );
}

static void SetGeometry( b2ShapeId shapeId, b2ShapeType type, const void* geometry )
{
	Shape* shape;
	World* world = WorldFromShape( shapeId, &shape );
	if ( world == nullptr || world->locked )
	{
		return;
	}
	world->SetShapeGeometry( shape, type, geometry );
}

void b2Shape_SetCircle( b2ShapeId shapeId, const b2Circle* circle )
{
	SetGeometry( shapeId, b2_circleShape, circle );
}

void b2Shape_SetCapsule( b2ShapeId shapeId, const b2Capsule* capsule )
{
	if ( b2DistanceSquared( capsule->center1, capsule->center2 ) <= B2_LINEAR_SLOP * B2_LINEAR_SLOP )
	{
		return;
	}
	SetGeometry( shapeId, b2_capsuleShape, capsule );
}

void b2Shape_SetSegment( b2ShapeId shapeId, const b2Segment* segment )
{
	SetGeometry( shapeId, b2_segmentShape, segment );
}

void b2Shape_SetPolygon( b2ShapeId shapeId, const b2Polygon* polygon )
{
	SetGeometry( shapeId, b2_polygonShape, polygon );
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

void b2DestroyJoint( b2JointId jointId )
{
	World* world = GetWorldFromIndex( jointId.world0 );
	if ( world != nullptr )
	{
		world->DestroyJoint( jointId );
	}
}

// The joint's solver record, current on the host. Edits mark the joint array for upload.
static JointSim* JointSimOf( b2JointId jointId, World** worldOut, bool edit, int type = -1 )
{
	Joint* joint;
	World* world = WorldFromJoint( jointId, &joint );
	if ( world == nullptr || ( type >= 0 && joint->type != type ) )
	{
		return nullptr;
	}
	world->PullJoints();
	if ( edit )
	{
		world->jointsDirty = true;
	}
	if ( worldOut != nullptr )
	{
		*worldOut = world;
	}
	return &world->jointSims[joint->jointId];
}

void b2Joint_SetUserData( b2JointId jointId, void* userData )
{
	Joint* joint;
	if ( WorldFromJoint( jointId, &joint ) != nullptr )
	{
		joint->userData = userData;
	}
}

void* b2Joint_GetUserData( b2JointId jointId )
{
	Joint* joint;
	return WorldFromJoint( jointId, &joint ) != nullptr ? joint->userData : nullptr;
}

// Event thresholds (reference src/joint.c:1369-1404); a finite threshold turns on the step's
// joint-event gather.
void b2Joint_SetForceThreshold( b2JointId jointId, float threshold )
{
	JointSim* js = JointSimOf( jointId, nullptr, true );
	if ( js != nullptr )
	{
		js->forceThreshold = threshold;
	}
}

float b2Joint_GetForceThreshold( b2JointId jointId )
{
	JointSim* js = JointSimOf( jointId, nullptr, false );
	return js != nullptr ? js->forceThreshold : 0.0f;
}

void b2Joint_SetTorqueThreshold( b2JointId jointId, float threshold )
{
	JointSim* js = JointSimOf( jointId, nullptr, true );
	if ( js != nullptr )
	{
		js->torqueThreshold = threshold;
	}
}

float b2Joint_GetTorqueThreshold( b2JointId jointId )
{
	JointSim* js = JointSimOf( jointId, nullptr, false );
	return js != nullptr ? js->torqueThreshold : 0.0f;
}

b2Transform b2Joint_GetLocalFrameA( b2JointId jointId )
{
	JointSim* js = JointSimOf( jointId, nullptr, false );
	return js != nullptr ? js->localFrameA : b2Transform_identity;
}

b2Transform b2Joint_GetLocalFrameB( b2JointId jointId )
{
	JointSim* js = JointSimOf( jointId, nullptr, false );
	return js != nullptr ? js->localFrameB : b2Transform_identity;
}

void b2Joint_SetLocalFrameA( b2JointId jointId, b2Transform localFrame )
{
	JointSim* js = JointSimOf( jointId, nullptr, true );
	if ( js != nullptr )
	{
		js->localFrameA = localFrame;
	}
}

void b2Joint_SetLocalFrameB( b2JointId jointId, b2Transform localFrame )
{
	JointSim* js = JointSimOf( jointId, nullptr, true );
	if ( js != nullptr )
	{
		js->localFrameB = localFrame;
	}
}

// Revolute joint accessors (reference src/revolute_joint.c:17-203). Toggling the spring, motor
// or limit clears the matching accumulated impulse, as the reference does.
#define B2G_REVOLUTE( EDIT )                                                                           \
	World* world;                                                                                      \
	JointSim* js = JointSimOf( jointId, &world, EDIT, b2_revoluteJoint );                              \
	(void)world

void b2RevoluteJoint_EnableSpring( b2JointId jointId, bool enableSpring )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr && enableSpring != ( js->enableSpring != 0 ) )
	{
		js->enableSpring = enableSpring;
		js->springImpulse = 0.0f;
	}
}

bool b2RevoluteJoint_IsSpringEnabled( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr && js->enableSpring != 0;
}

void b2RevoluteJoint_SetSpringHertz( b2JointId jointId, float hertz )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr )
	{
		js->hertz = hertz;
	}
}

float b2RevoluteJoint_GetSpringHertz( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr ? js->hertz : 0.0f;
}

void b2RevoluteJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr )
	{
		js->dampingRatio = dampingRatio;
	}
}

float b2RevoluteJoint_GetSpringDampingRatio( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr ? js->dampingRatio : 0.0f;
}

void b2RevoluteJoint_SetTargetAngle( b2JointId jointId, float angle )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr )
	{
		js->targetAngle = angle;
	}
}

float b2RevoluteJoint_GetTargetAngle( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr ? js->targetAngle : 0.0f;
}

float b2RevoluteJoint_GetAngle( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	if ( js == nullptr )
	{
		return 0.0f;
	}
	world->PullBodies();
	b2Rot qA = b2MulRot( world->bodySims[js->bodyIdA].transform.q, js->localFrameA.q );
	b2Rot qB = b2MulRot( world->bodySims[js->bodyIdB].transform.q, js->localFrameB.q );
	return b2RelativeAngle( qA, qB );
}

void b2RevoluteJoint_EnableLimit( b2JointId jointId, bool enableLimit )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr && enableLimit != ( js->enableLimit != 0 ) )
	{
		js->enableLimit = enableLimit;
		js->lowerImpulse = 0.0f;
		js->upperImpulse = 0.0f;
	}
}

bool b2RevoluteJoint_IsLimitEnabled( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr && js->enableLimit != 0;
}

float b2RevoluteJoint_GetLowerLimit( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr ? js->lowerAngle : 0.0f;
}

float b2RevoluteJoint_GetUpperLimit( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr ? js->upperAngle : 0.0f;
}

void b2RevoluteJoint_SetLimits( b2JointId jointId, float lower, float upper )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr )
	{
		float lowerAngle = b2MinFloat( lower, upper );
		float upperAngle = b2MaxFloat( lower, upper );
		js->lowerAngle = b2ClampFloat( lowerAngle, -0.99f * B2_PI, 0.99f * B2_PI );
		js->upperAngle = b2ClampFloat( upperAngle, -0.99f * B2_PI, 0.99f * B2_PI );
	}
}

void b2RevoluteJoint_EnableMotor( b2JointId jointId, bool enableMotor )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr && enableMotor != ( js->enableMotor != 0 ) )
	{
		js->enableMotor = enableMotor;
		js->motorImpulse = 0.0f;
	}
}

bool b2RevoluteJoint_IsMotorEnabled( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr && js->enableMotor != 0;
}

void b2RevoluteJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr )
	{
		js->motorSpeed = motorSpeed;
	}
}

float b2RevoluteJoint_GetMotorSpeed( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr ? js->motorSpeed : 0.0f;
}

float b2RevoluteJoint_GetMotorTorque( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr ? world->inv_h * js->motorImpulse : 0.0f;
}

void b2RevoluteJoint_SetMaxMotorTorque( b2JointId jointId, float torque )
{
	B2G_REVOLUTE( true );
	if ( js != nullptr )
	{
		js->maxMotorTorque = torque;
	}
}

float b2RevoluteJoint_GetMaxMotorTorque( b2JointId jointId )
{
	B2G_REVOLUTE( false );
	return js != nullptr ? js->maxMotorTorque : 0.0f;
}

#undef B2G_REVOLUTE

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
		for ( int k = 0; k < kBatchBuckets; ++k )
		{
			stats->bucketWorlds[k] = bs.bucketWorlds[k];
			stats->bucketThreads[k] = bs.bucketThreads[k];
			stats->bucketWarmByBody[k] = bs.bucketWarmByBody[k];
			stats->bucketWorkgroups[k] = bs.bucketGrid[k];
			stats->bucketMs[k] = bs.bucketMs[k];
		}
	}
	return n;
}

void b2amd_Synchronize( void )
{
	DeviceSynchronize();
}

int b2amd_SetBatchSolverThreads( int threads )
{
	return SetBatchSolverThreads( threads ) ? 0 : -1;
}

int b2amd_SetBatchTuning( int bucket, int threads, int warmByBody )
{
	return SetBatchTuning( bucket, threads, warmByBody ) ? 0 : -1;
}

int b2amd_GetBatchTuning( int bucket, int* maxBodies, int* threads, int* warmByBody, int* tableThreads, int* tableWarmByBody )
{
	if ( bucket < 0 || bucket >= kBatchBuckets )
	{
		return -1;
	}
	BatchTuning t = BatchTuningFor( bucket );
	BatchTuning table = BatchTuningTable( bucket );
	if ( maxBodies != nullptr )
	{
		*maxBodies = t.maxBodies;
	}
	if ( threads != nullptr )
	{
		*threads = t.threads;
	}
	if ( warmByBody != nullptr )
	{
		*warmByBody = t.warmByBody;
	}
	if ( tableThreads != nullptr )
	{
		*tableThreads = table.threads;
	}
	if ( tableWarmByBody != nullptr )
	{
		*tableWarmByBody = table.warmByBody;
	}
	return 0;
}

int b2amd_SpinWaitActive( void )
{
	return SpinWaitActive() ? 1 : 0;
}

int b2amd_CollideBatch( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
						b2LocalManifold* manifolds )
{
	return CollideBatch( shapesA, shapesB, transforms, count, manifolds );
}

int b2amd_CollideBatchNoAlloc( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
							   b2LocalManifold* manifolds )
{
	return CollideBatchNoAlloc( shapesA, shapesB, transforms, count, manifolds );
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
