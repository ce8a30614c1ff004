// Host-side world bookkeeping for b2World_Step (see world.h).
//
// Each block below restates a reference routine whose ordering decisions the GPU results depend
// on; the reference file:line is given at each function. Arithmetic that runs every step lives in
// the GPU kernels (device.hip); the host only runs creation-time arithmetic (shape margins, mass
// properties, initial AABBs) and integer bookkeeping.

#include "world.h"

#include "device.h"

#include <algorithm>
#include <cfloat>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

namespace b2g
{

static b2LogFcn* s_logFcn = nullptr;
static b2AssertFcn* s_assertFcn = nullptr;

void Log( const char* fmt, ... )
{
	char buffer[512];
	va_list args;
	va_start( args, fmt );
	vsnprintf( buffer, sizeof( buffer ), fmt, args );
	va_end( args );
	if ( s_logFcn != nullptr )
	{
		s_logFcn( buffer );
	}
	else
	{
		fprintf( stderr, "box2d_amd: %s\n", buffer );
	}
}

void Fatal( const char* fmt, ... )
{
	char buffer[512];
	va_list args;
	va_start( args, fmt );
	vsnprintf( buffer, sizeof( buffer ), fmt, args );
	va_end( args );
	fprintf( stderr, "box2d_amd fatal: %s\n", buffer );
	fflush( stderr );
	abort();
}

// ---------------------------------------------------------------------------------------------
// World registry

static World* s_worlds[B2_MAX_WORLDS] = {};
static uint16_t s_generations[B2_MAX_WORLDS] = {};

World* GetWorld( b2WorldId id )
{
	if ( id.index1 < 1 || id.index1 > B2_MAX_WORLDS )
	{
		return nullptr;
	}
	World* world = s_worlds[id.index1 - 1];
	if ( world == nullptr || world->generation != id.generation )
	{
		return nullptr;
	}
	return world;
}

World* GetWorldFromIndex( int index )
{
	if ( index < 0 || index >= B2_MAX_WORLDS )
	{
		return nullptr;
	}
	return s_worlds[index];
}

// ---------------------------------------------------------------------------------------------
// Shape helpers (reference src/shape.c:36-87, 704-721, 826-887)

static float ComputeShapeMargin( const ShapeDev& s )
{
	float margin = 0.0f;
	switch ( s.type )
	{
		case b2_capsuleShape:
			margin = 0.5f * b2Distance( s.capsule.center2, s.capsule.center1 ) + s.capsule.radius;
			break;
		case b2_circleShape:
			margin = s.circle.radius;
			break;
		case b2_polygonShape:
		{
			float maxExtentSqr = 0.0f;
			for ( int i = 0; i < s.polygon.count; ++i )
			{
				float distanceSqr = b2DistanceSquared( s.polygon.vertices[i], s.polygon.centroid );
				maxExtentSqr = b2MaxFloat( maxExtentSqr, distanceSqr );
			}
			margin = sqrtf( maxExtentSqr );
		}
		break;
		case b2_segmentShape:
			margin = 0.5f * b2Distance( s.segment.point1, s.segment.point2 );
			break;
		case b2_chainSegmentShape:
			margin = 0.5f * b2Distance( s.chainSegment.segment.point1, s.chainSegment.segment.point2 );
			break;
		default:
			return B2_MAX_AABB_MARGIN;
	}
	return b2MinFloat( B2_MAX_AABB_MARGIN, B2_AABB_MARGIN_FRACTION * margin );
}

static b2Vec2 ShapeCentroid( const ShapeDev& s )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return b2Lerp( s.capsule.center1, s.capsule.center2, 0.5f );
		case b2_circleShape:
			return s.circle.center;
		case b2_polygonShape:
			return s.polygon.centroid;
		case b2_segmentShape:
			return b2Lerp( s.segment.point1, s.segment.point2, 0.5f );
		case b2_chainSegmentShape:
			return b2Lerp( s.chainSegment.segment.point1, s.chainSegment.segment.point2, 0.5f );
		default:
			return b2Vec2{ 0.0f, 0.0f };
	}
}

struct ShapeExtent
{
	float minExtent;
	float maxExtent;
};

static ShapeExtent ComputeShapeExtent( const ShapeDev& s, b2Vec2 localCenter )
{
	ShapeExtent extent = { 0.0f, 0.0f };
	switch ( s.type )
	{
		case b2_capsuleShape:
		{
			float radius = s.capsule.radius;
			extent.minExtent = radius;
			b2Vec2 c1 = b2Sub( s.capsule.center1, localCenter );
			b2Vec2 c2 = b2Sub( s.capsule.center2, localCenter );
			extent.maxExtent = sqrtf( b2MaxFloat( b2LengthSquared( c1 ), b2LengthSquared( c2 ) ) ) + radius;
		}
		break;
		case b2_circleShape:
		{
			float radius = s.circle.radius;
			extent.minExtent = radius;
			extent.maxExtent = b2Length( b2Sub( s.circle.center, localCenter ) ) + radius;
		}
		break;
		case b2_polygonShape:
		{
			const b2Polygon& poly = s.polygon;
			float minExtent = B2_HUGE;
			float maxExtentSqr = 0.0f;
			for ( int i = 0; i < poly.count; ++i )
			{
				b2Vec2 v = poly.vertices[i];
				float planeOffset = b2Dot( poly.normals[i], b2Sub( v, poly.centroid ) );
				minExtent = b2MinFloat( minExtent, planeOffset );
				float distanceSqr = b2LengthSquared( b2Sub( v, localCenter ) );
				maxExtentSqr = b2MaxFloat( maxExtentSqr, distanceSqr );
			}
			extent.minExtent = minExtent + poly.radius;
			extent.maxExtent = sqrtf( maxExtentSqr ) + poly.radius;
		}
		break;
		case b2_segmentShape:
		{
			b2Vec2 c1 = b2Sub( s.segment.point1, localCenter );
			b2Vec2 c2 = b2Sub( s.segment.point2, localCenter );
			extent.maxExtent = sqrtf( b2MaxFloat( b2LengthSquared( c1 ), b2LengthSquared( c2 ) ) );
		}
		break;
		case b2_chainSegmentShape:
		{
			b2Vec2 c1 = b2Sub( s.chainSegment.segment.point1, localCenter );
			b2Vec2 c2 = b2Sub( s.chainSegment.segment.point2, localCenter );
			extent.maxExtent = sqrtf( b2MaxFloat( b2LengthSquared( c1 ), b2LengthSquared( c2 ) ) );
		}
		break;
		default:
			break;
	}
	return extent;
}

static b2MassData ComputeShapeMass( const ShapeDev& s, float density )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return b2ComputeCapsuleMass( &s.capsule, density );
		case b2_circleShape:
			return b2ComputeCircleMass( &s.circle, density );
		case b2_polygonShape:
			return b2ComputePolygonMass( &s.polygon, density );
		default:
			return b2MassData{ 0.0f, { 0.0f, 0.0f }, 0.0f };
	}
}

// Fat AABB on the host for shape creation (same construction as the GPU finalize stage,
// reference src/geometry.c:402-505).
static b2AABB HostFatShapeAABB( const ShapeDev& s, b2Transform xf, float extra )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
		{
			b2Vec2 v1 = b2TransformWorldPoint( xf, s.capsule.center1 );
			b2Vec2 v2 = b2TransformWorldPoint( xf, s.capsule.center2 );
			double r = (double)s.capsule.radius + (double)extra;
			double v1x = v1.x, v1y = v1.y, v2x = v2.x, v2y = v2.y;
			return b2AABB{ { (float)( ( v1x < v2x ? v1x : v2x ) - r ), (float)( ( v1y < v2y ? v1y : v2y ) - r ) },
						   { (float)( ( v1x > v2x ? v1x : v2x ) + r ), (float)( ( v1y > v2y ? v1y : v2y ) + r ) } };
		}
		case b2_circleShape:
		{
			b2Vec2 c = b2TransformWorldPoint( xf, s.circle.center );
			double r = (double)s.circle.radius + (double)extra;
			double cx = c.x, cy = c.y;
			return b2AABB{ { (float)( cx - r ), (float)( cy - r ) }, { (float)( cx + r ), (float)( cy + r ) } };
		}
		case b2_polygonShape:
		{
			b2Vec2 v = b2TransformWorldPoint( xf, s.polygon.vertices[0] );
			double lx = v.x, ly = v.y, ux = v.x, uy = v.y;
			for ( int i = 1; i < s.polygon.count; ++i )
			{
				v = b2TransformWorldPoint( xf, s.polygon.vertices[i] );
				double vx = v.x, vy = v.y;
				lx = vx < lx ? vx : lx;
				ly = vy < ly ? vy : ly;
				ux = vx > ux ? vx : ux;
				uy = vy > uy ? vy : uy;
			}
			double r = (double)s.polygon.radius + (double)extra;
			return b2AABB{ { (float)( lx - r ), (float)( ly - r ) }, { (float)( ux + r ), (float)( uy + r ) } };
		}
		case b2_segmentShape:
		case b2_chainSegmentShape:
		{
			const b2Segment& seg = s.type == b2_segmentShape ? s.segment : s.chainSegment.segment;
			b2Vec2 v1 = b2TransformWorldPoint( xf, seg.point1 );
			b2Vec2 v2 = b2TransformWorldPoint( xf, seg.point2 );
			double e = (double)extra;
			double v1x = v1.x, v1y = v1.y, v2x = v2.x, v2y = v2.y;
			return b2AABB{ { (float)( ( v1x < v2x ? v1x : v2x ) - e ), (float)( ( v1y < v2y ? v1y : v2y ) - e ) },
						   { (float)( ( v1x > v2x ? v1x : v2x ) + e ), (float)( ( v1y > v2y ? v1y : v2y ) + e ) } };
		}
		default:
			return b2AABB{ { 0.0f, 0.0f }, { 0.0f, 0.0f } };
	}
}

// Contact registry: which shape type pairs have a manifold and which order is primary
// (reference src/contact.c:166-206).
struct ContactRegister
{
	bool exists;
	bool primary;
};

static ContactRegister s_registers[b2_shapeTypeCount][b2_shapeTypeCount];

static void AddType( b2ShapeType type1, b2ShapeType type2 )
{
	s_registers[type1][type2] = { true, true };
	if ( type1 != type2 )
	{
		s_registers[type2][type1] = { true, false };
	}
}

static void InitializeContactRegisters()
{
	static bool initialized = false;
	if ( initialized )
	{
		return;
	}
	AddType( b2_circleShape, b2_circleShape );
	AddType( b2_capsuleShape, b2_circleShape );
	AddType( b2_capsuleShape, b2_capsuleShape );
	AddType( b2_polygonShape, b2_circleShape );
	AddType( b2_polygonShape, b2_capsuleShape );
	AddType( b2_polygonShape, b2_polygonShape );
	AddType( b2_segmentShape, b2_circleShape );
	AddType( b2_segmentShape, b2_capsuleShape );
	AddType( b2_segmentShape, b2_polygonShape );
	AddType( b2_chainSegmentShape, b2_circleShape );
	AddType( b2_chainSegmentShape, b2_capsuleShape );
	AddType( b2_chainSegmentShape, b2_polygonShape );
	initialized = true;
}

static bool CanCollide( b2ShapeType a, b2ShapeType b )
{
	return s_registers[a][b].exists;
}

static bool ShouldShapesCollide( b2Filter filterA, b2Filter filterB )
{
	if ( filterA.groupIndex == filterB.groupIndex && filterA.groupIndex != 0 )
	{
		return filterA.groupIndex > 0;
	}
	return ( filterA.maskBits & filterB.categoryBits ) != 0 && ( filterA.categoryBits & filterB.maskBits ) != 0;
}

// ---------------------------------------------------------------------------------------------
// Broadphase proxy bookkeeping (reference src/broad_phase.c:82-147)

void BroadPhase::UnBufferMove( int proxyKey )
{
	b2BodyType proxyType = B2G_PROXY_TYPE( proxyKey );
	int proxyId = B2G_PROXY_ID( proxyKey );
	if ( movedProxies[proxyType].GetBit( proxyId ) )
	{
		movedProxies[proxyType].ClearBit( proxyId );
		for ( size_t i = 0; i < moveArray.size(); ++i )
		{
			if ( moveArray[i] == proxyKey )
			{
				RemoveSwap( moveArray, (int)i );
				break;
			}
		}
	}
}

int BroadPhase::CreateProxy( b2AABB aabb, b2BodyType proxyType, uint64_t categoryBits, int shapeIndex, bool forcePairCreation )
{
	int proxyId = trees[proxyType].CreateProxy( aabb, categoryBits, (uint64_t)shapeIndex );
	int proxyKey = B2G_PROXY_KEY( proxyId, proxyType );
	if ( proxyType != b2_staticBody || forcePairCreation )
	{
		BufferMove( proxyKey );
	}
	return proxyKey;
}

void BroadPhase::DestroyProxy( int proxyKey )
{
	UnBufferMove( proxyKey );
	trees[B2G_PROXY_TYPE( proxyKey )].DestroyProxy( B2G_PROXY_ID( proxyKey ) );
}

void BroadPhase::MoveProxy( int proxyKey, b2AABB aabb )
{
	trees[B2G_PROXY_TYPE( proxyKey )].MoveProxy( B2G_PROXY_ID( proxyKey ), aabb );
	BufferMove( proxyKey );
}

void BroadPhase::EnlargeProxy( int proxyKey, b2AABB aabb )
{
	// the ancestors are updated before the trees' internal nodes are next read: in the next
	// rebuild, which runs while the GPU solves, or before a host query
	trees[B2G_PROXY_TYPE( proxyKey )].EnlargeProxyDeferred( B2G_PROXY_ID( proxyKey ), aabb );
	BufferMove( proxyKey );
}

// ---------------------------------------------------------------------------------------------
// World

World::World( const b2WorldDef* def, int worldIndex, uint16_t gen )
{
	InitializeContactRegisters();

	worldId = (uint16_t)worldIndex;
	generation = gen;

	// static, disabled and awake solver sets (reference src/physics_world.c:206-227)
	for ( int i = 0; i < 3; ++i )
	{
		int setId = solverSetIdPool.Alloc();
		SolverSet set;
		set.setIndex = setId;
		solverSets.push_back( set );
	}

	gravity = def->gravity;
	hitEventThreshold = def->hitEventThreshold;
	restitutionThreshold = def->restitutionThreshold;
	maxLinearSpeed = def->maximumLinearSpeed;
	contactSpeed = def->contactSpeed;
	contactHertz = def->contactHertz;
	contactDampingRatio = def->contactDampingRatio;
	contactRecycleDistance = B2_CONTACT_RECYCLE_DISTANCE;
	enableSleep = def->enableSleep;
	enableContinuous = def->enableContinuous;
	enableContactSoftening = def->enableContactSoftening;
	userData = def->userData;
	workerCount = def->workerCount;

	if ( b2GetLengthUnitsPerMeter() != 1.0f )
	{
		// the device code's slop and speculative distances are constants for one unit per meter
		Fatal( "b2SetLengthUnitsPerMeter(%g) is not supported by the GPU step (1 unit per meter only)",
			   (double)b2GetLengthUnitsPerMeter() );
	}

	frictionCallback = def->frictionCallback;
	restitutionCallback = def->restitutionCallback;

	device = new DeviceWorld();
}

World::~World()
{
	delete pairFinder;
	delete device;
}

Body* World::GetBody( b2BodyId id )
{
	int index = id.index1 - 1;
	if ( index < 0 || index >= (int)bodies.size() )
	{
		return nullptr;
	}
	Body* body = &bodies[index];
	if ( body->id != index || body->generation != id.generation )
	{
		return nullptr;
	}
	return body;
}

void World::PullBodies()
{
	if ( bodiesStale )
	{
		device->DownloadBodies( *this );
		bodiesStale = false;
	}
}

void World::PushBodiesIfDirty()
{
	if ( bodiesDirty )
	{
		device->UploadBodies( *this );
		bodiesDirty = false;
	}
}

void World::PullShapes()
{
	if ( shapesStale )
	{
		device->DownloadShapes( *this );
		shapesStale = false;
	}
}

void World::PullContacts()
{
	if ( contactsStale || contactSims.size() < contacts.size() )
	{
		contactSims.resize( contacts.size() );
		device->DownloadContacts( (int)contacts.size(), contactSims.data() );
		contactsStale = false;
	}
}

const ContactSim& World::PulledContact( int contactId )
{
	PullContacts();
	return contactSims[contactId];
}

void World::PullJoints()
{
	if ( jointsStale )
	{
		device->DownloadJoints( *this );
		jointsStale = false;
	}
}

// Deep state hash (reference src/world_snapshot.c:1025-1187, FNV-1a 64 from src/recording.h:18-38):
// body transforms, awake velocities and set indices; contact set/color/local indices and manifold
// impulses; joint indices and accumulated impulses by type; id pool states; the solver set count.
// The device copies are pulled first, so the hash describes the state after the last step.
namespace
{
constexpr uint64_t kFnvInit = 14695981039346656037ull;
constexpr uint64_t kFnvPrime = 1099511628211ull;

uint64_t MixWord( uint64_t hash, uint32_t word )
{
	return ( hash ^ (uint64_t)word ) * kFnvPrime;
}

uint64_t MixFloat( uint64_t hash, float f )
{
	uint32_t word;
	memcpy( &word, &f, 4 );
	return MixWord( hash, word );
}

uint64_t MixInt( uint64_t hash, int v )
{
	return MixWord( hash, (uint32_t)v );
}

uint64_t MixVec2Bytes( uint64_t hash, b2Vec2 v )
{
	// the reference mixes a b2Vec2 byte by byte
	unsigned char bytes[8];
	memcpy( bytes, &v, 8 );
	for ( unsigned char b : bytes )
	{
		hash = ( hash ^ (uint64_t)b ) * kFnvPrime;
	}
	return hash;
}

uint64_t MixJointImpulses( uint64_t hash, const JointSim& j )
{
	switch ( j.type )
	{
		case b2_distanceJoint:
			hash = MixFloat( hash, j.distanceImpulse );
			hash = MixFloat( hash, j.lowerImpulse );
			hash = MixFloat( hash, j.upperImpulse );
			return MixFloat( hash, j.motorImpulse );
		case b2_moverJoint:
			hash = MixFloat( hash, j.linearVelocityImpulse.x );
			return MixFloat( hash, j.linearVelocityImpulse.y );
		case b2_motorJoint:
			hash = MixFloat( hash, j.linearVelocityImpulse.x );
			hash = MixFloat( hash, j.linearVelocityImpulse.y );
			hash = MixFloat( hash, j.angularVelocityImpulse );
			hash = MixFloat( hash, j.linearSpringImpulse.x );
			hash = MixFloat( hash, j.linearSpringImpulse.y );
			return MixFloat( hash, j.angularSpringImpulse );
		case b2_pogoJoint:
			hash = MixFloat( hash, j.pogoImpulse );
			return MixFloat( hash, j.pogoVelocity );
		case b2_prismaticJoint:
			hash = MixVec2Bytes( hash, j.pointImpulse );
			hash = MixFloat( hash, j.springImpulse );
			hash = MixFloat( hash, j.motorImpulse );
			hash = MixFloat( hash, j.lowerImpulse );
			return MixFloat( hash, j.upperImpulse );
		case b2_revoluteJoint:
			hash = MixVec2Bytes( hash, j.linearImpulse );
			hash = MixFloat( hash, j.springImpulse );
			hash = MixFloat( hash, j.motorImpulse );
			hash = MixFloat( hash, j.lowerImpulse );
			return MixFloat( hash, j.upperImpulse );
		case b2_weldJoint:
			hash = MixVec2Bytes( hash, j.linearImpulse );
			return MixFloat( hash, j.angularImpulse );
		case b2_wheelJoint:
			hash = MixFloat( hash, j.perpImpulse );
			hash = MixFloat( hash, j.motorImpulse );
			hash = MixFloat( hash, j.springImpulse );
			hash = MixFloat( hash, j.lowerImpulse );
			return MixFloat( hash, j.upperImpulse );
		default:
			return hash;
	}
}

uint64_t MixPool( uint64_t hash, const IdPool& pool )
{
	hash = MixInt( hash, pool.nextIndex );
	return MixInt( hash, pool.Count() );
}
} // namespace

uint64_t World::StateHash()
{
	PullBodies();
	PullJoints();
	uint64_t hash = kFnvInit;

	for ( int i = 0; i < (int)bodies.size(); ++i )
	{
		const Body& body = bodies[i];
		if ( body.id != i )
		{
			continue;
		}
		const BodySim& sim = bodySims[i];
		hash = MixFloat( hash, sim.transform.p.x );
		hash = MixFloat( hash, sim.transform.p.y );
		hash = MixFloat( hash, sim.transform.q.c );
		hash = MixFloat( hash, sim.transform.q.s );
		if ( body.setIndex == kAwakeSet )
		{
			const BodyState& state = bodyStates[i];
			hash = MixFloat( hash, state.linearVelocity.x );
			hash = MixFloat( hash, state.linearVelocity.y );
			hash = MixFloat( hash, state.angularVelocity );
		}
		hash = MixInt( hash, body.setIndex );
		hash = MixInt( hash, body.localIndex );
	}

	for ( int i = 0; i < (int)contacts.size(); ++i )
	{
		const Contact& contact = contacts[i];
		if ( contact.contactId != i )
		{
			continue;
		}
		hash = MixInt( hash, contact.setIndex );
		hash = MixInt( hash, contact.colorIndex );
		hash = MixInt( hash, contact.localIndex );
		const b2Manifold& m = PulledContact( i ).manifold;
		hash = MixInt( hash, m.pointCount );
		for ( int p = 0; p < m.pointCount; ++p )
		{
			hash = MixFloat( hash, m.points[p].normalImpulse );
			hash = MixFloat( hash, m.points[p].tangentImpulse );
			hash = MixFloat( hash, m.points[p].totalNormalImpulse );
		}
	}

	for ( int i = 0; i < (int)joints.size(); ++i )
	{
		const Joint& joint = joints[i];
		if ( joint.jointId != i )
		{
			continue;
		}
		hash = MixInt( hash, joint.setIndex );
		hash = MixInt( hash, joint.colorIndex );
		hash = MixInt( hash, joint.localIndex );
		hash = MixJointImpulses( hash, jointSims[i] );
	}

	hash = MixPool( hash, bodyIdPool );
	hash = MixPool( hash, shapeIdPool );
	hash = MixPool( hash, chainIdPool );
	hash = MixPool( hash, contactIdPool );
	hash = MixPool( hash, jointIdPool );
	hash = MixPool( hash, islandIdPool );
	hash = MixPool( hash, solverSetIdPool );
	return MixInt( hash, (int)solverSets.size() );
}

// ---------------------------------------------------------------------------------------------
// Islands (reference src/island.c)

Island* World::CreateIsland( int setIndex )
{
	islandsDirty = true;
	int islandId = islandIdPool.Alloc();
	if ( islandId == (int)islands.size() )
	{
		islands.emplace_back();
	}

	SolverSet& set = solverSets[setIndex];
	Island* island = &islands[islandId];
	island->setIndex = setIndex;
	island->localIndex = (int)set.islandIds.size();
	island->islandId = islandId;
	island->bodies.clear();
	island->contacts.clear();
	island->joints.clear();
	island->constraintRemoveCount = 0;
	set.islandIds.push_back( islandId );
	return island;
}

void World::DestroyIsland( int islandId )
{
	islandsDirty = true;
	if ( splitIslandId == islandId )
	{
		splitIslandId = B2_NULL_INDEX;
	}

	Island* island = &islands[islandId];
	SolverSet& set = solverSets[island->setIndex];
	int localIndex = island->localIndex;
	int lastIndex = (int)set.islandIds.size() - 1;
	int moveIslandId = set.islandIds[lastIndex];
	set.islandIds[localIndex] = set.islandIds[lastIndex];
	islands[moveIslandId].localIndex = localIndex;
	set.islandIds.pop_back();

	island->bodies.clear();
	island->contacts.clear();
	island->joints.clear();
	island->constraintRemoveCount = 0;
	island->localIndex = B2_NULL_INDEX;
	island->islandId = B2_NULL_INDEX;
	island->setIndex = B2_NULL_INDEX;
	islandIdPool.Free( islandId );
}

void World::CreateIslandForBody( int setIndex, Body* body )
{
	islandsDirty = true;
	Island* island = CreateIsland( setIndex );
	island->bodies.push_back( body->id );
	body->islandId = island->islandId;
	body->islandIndex = 0;
}

int World::MergeIslands( int islandIdA, int islandIdB )
{
	islandsDirty = true;
	if ( islandIdA == islandIdB )
	{
		return islandIdA;
	}
	if ( islandIdA == B2_NULL_INDEX )
	{
		return islandIdB;
	}
	if ( islandIdB == B2_NULL_INDEX )
	{
		return islandIdA;
	}

	Island* islandA = &islands[islandIdA];
	Island* islandB = &islands[islandIdB];
	Island* bigIsland;
	Island* smallIsland;
	if ( islandA->bodies.size() >= islandB->bodies.size() )
	{
		bigIsland = islandA;
		smallIsland = islandB;
	}
	else
	{
		bigIsland = islandB;
		smallIsland = islandA;
	}

	int bigIslandId = bigIsland->islandId;

	for ( int bodyId : smallIsland->bodies )
	{
		Body& body = bodies[bodyId];
		body.islandId = bigIslandId;
		body.islandIndex = (int)bigIsland->bodies.size();
		bigIsland->bodies.push_back( bodyId );
	}

	for ( const ContactLink& link : smallIsland->contacts )
	{
		Contact& contact = contacts[link.contactId];
		contact.islandId = bigIslandId;
		contact.islandIndex = (int)bigIsland->contacts.size();
		bigIsland->contacts.push_back( link );
	}

	for ( const JointLink& link : smallIsland->joints )
	{
		Joint& joint = joints[link.jointId];
		joint.islandId = bigIslandId;
		joint.islandIndex = (int)bigIsland->joints.size();
		bigIsland->joints.push_back( link );
	}

	bigIsland->constraintRemoveCount += smallIsland->constraintRemoveCount;
	DestroyIsland( smallIsland->islandId );
	return bigIslandId;
}

void World::AddContactToIsland( int islandId, Contact* contact )
{
	islandsDirty = true;
	Island* island = &islands[islandId];
	contact->islandId = islandId;
	contact->islandIndex = (int)island->contacts.size();
	island->contacts.push_back( ContactLink{ contact->contactId, contact->edges[0].bodyId, contact->edges[1].bodyId } );
}

void World::LinkContact( Contact* contact )
{
	int bodyIdA = contact->edges[0].bodyId;
	int bodyIdB = contact->edges[1].bodyId;

	if ( bodies[bodyIdA].setIndex == kAwakeSet && bodies[bodyIdB].setIndex >= kFirstSleepingSet )
	{
		WakeSolverSet( bodies[bodyIdB].setIndex );
	}
	if ( bodies[bodyIdB].setIndex == kAwakeSet && bodies[bodyIdA].setIndex >= kFirstSleepingSet )
	{
		WakeSolverSet( bodies[bodyIdA].setIndex );
	}

	int islandIdA = bodies[bodyIdA].islandId;
	int islandIdB = bodies[bodyIdB].islandId;
	int finalIslandId = MergeIslands( islandIdA, islandIdB );
	AddContactToIsland( finalIslandId, contact );
}

void World::UnlinkContact( Contact* contact )
{
	islandsDirty = true;
	int islandId = contact->islandId;
	Island* island = &islands[islandId];
	int removeIndex = contact->islandIndex;
	int movedIndex = RemoveSwap( island->contacts, removeIndex );
	if ( movedIndex != B2_NULL_INDEX )
	{
		contacts[island->contacts[removeIndex].contactId].islandIndex = removeIndex;
	}
	contact->islandId = B2_NULL_INDEX;
	contact->islandIndex = B2_NULL_INDEX;
	island->constraintRemoveCount += 1;
}

void World::AddJointToIsland( int islandId, Joint* joint )
{
	islandsDirty = true;
	Island* island = &islands[islandId];
	joint->islandId = islandId;
	joint->islandIndex = (int)island->joints.size();
	island->joints.push_back( JointLink{ joint->jointId, joint->edges[0].bodyId, joint->edges[1].bodyId } );
}

void World::LinkJoint( Joint* joint )
{
	Body* bodyA = &bodies[joint->edges[0].bodyId];
	Body* bodyB = &bodies[joint->edges[1].bodyId];

	if ( bodyA->setIndex == kAwakeSet && bodyB->setIndex >= kFirstSleepingSet )
	{
		WakeSolverSet( bodyB->setIndex );
	}
	else if ( bodyB->setIndex == kAwakeSet && bodyA->setIndex >= kFirstSleepingSet )
	{
		WakeSolverSet( bodyA->setIndex );
	}

	int finalIslandId = MergeIslands( bodyA->islandId, bodyB->islandId );
	AddJointToIsland( finalIslandId, joint );
}

static inline int IslandFindParent( std::vector<int>& parents, int node )
{
	while ( parents[node] != node )
	{
		int grandParent = parents[parents[node]];
		parents[node] = grandParent;
		node = grandParent;
	}
	return node;
}

// Union-find split of an island whose constraints were removed (reference src/island.c:389-651).
void World::SplitIsland( int baseId )
{
	islandsDirty = true;
	Island* baseIsland = &islands[baseId];
	int baseBodyCount = (int)baseIsland->bodies.size();
	std::vector<int> baseBodyIds = baseIsland->bodies;
	std::vector<ContactLink> baseContacts = baseIsland->contacts;
	std::vector<JointLink> baseJoints = baseIsland->joints;

	std::vector<int> parents( baseBodyCount ), contactCounts( baseBodyCount, 0 ), jointCounts( baseBodyCount, 0 ),
		ranks( baseBodyCount, 0 );
	for ( int i = 0; i < baseBodyCount; ++i )
	{
		parents[i] = i;
	}

	auto unite = [&]( int node1, int node2 ) {
		int root1 = IslandFindParent( parents, node1 );
		int root2 = IslandFindParent( parents, node2 );
		if ( root1 != root2 )
		{
			if ( ranks[root1] < ranks[root2] )
			{
				parents[root1] = root2;
				contactCounts[root2] += contactCounts[root1];
				jointCounts[root2] += jointCounts[root1];
			}
			else if ( ranks[root1] > ranks[root2] )
			{
				parents[root2] = root1;
				contactCounts[root1] += contactCounts[root2];
				jointCounts[root1] += jointCounts[root2];
			}
			else
			{
				parents[root2] = root1;
				ranks[root1] += 1;
				contactCounts[root1] += contactCounts[root2];
				jointCounts[root1] += jointCounts[root2];
			}
		}
	};

	for ( const ContactLink& link : baseContacts )
	{
		int islandIndexA = bodies[link.bodyIdA].islandIndex;
		int islandIndexB = bodies[link.bodyIdB].islandIndex;
		if ( islandIndexA != B2_NULL_INDEX && islandIndexB != B2_NULL_INDEX )
		{
			unite( islandIndexA, islandIndexB );
			contactCounts[IslandFindParent( parents, islandIndexA )] += 1;
		}
		else
		{
			int islandIndex = islandIndexA != B2_NULL_INDEX ? islandIndexA : islandIndexB;
			contactCounts[IslandFindParent( parents, islandIndex )] += 1;
		}
	}

	for ( const JointLink& link : baseJoints )
	{
		int islandIndexA = bodies[link.bodyIdA].islandIndex;
		int islandIndexB = bodies[link.bodyIdB].islandIndex;
		if ( islandIndexA != B2_NULL_INDEX && islandIndexB != B2_NULL_INDEX )
		{
			unite( islandIndexA, islandIndexB );
			jointCounts[IslandFindParent( parents, islandIndexA )] += 1;
		}
		else
		{
			int islandIndex = islandIndexA != B2_NULL_INDEX ? islandIndexA : islandIndexB;
			jointCounts[IslandFindParent( parents, islandIndex )] += 1;
		}
	}

	int componentCount = 0;
	for ( int i = 0; i < baseBodyCount; ++i )
	{
		parents[i] = IslandFindParent( parents, i );
		if ( parents[i] == i )
		{
			componentCount += 1;
		}
	}

	if ( componentCount == 1 )
	{
		baseIsland->constraintRemoveCount = 0;
		return;
	}

	baseIsland->bodies.clear();
	baseIsland->contacts.clear();
	baseIsland->joints.clear();

	std::vector<int> rootMap( baseBodyCount, B2_NULL_INDEX );
	int islandCount = 0;
	for ( int i = 0; i < baseBodyCount; ++i )
	{
		int rootIndex = parents[i];
		if ( rootMap[rootIndex] == B2_NULL_INDEX )
		{
			rootMap[rootIndex] = islandCount;
			islandCount += 1;
		}
	}

	std::vector<int> islandIds( islandCount );
	for ( int i = 0; i < islandCount; ++i )
	{
		Island* newIsland = CreateIsland( kAwakeSet );
		islandIds[i] = newIsland->islandId;
	}

	for ( int i = 0; i < baseBodyCount; ++i )
	{
		int bodyId = baseBodyIds[i];
		int root = IslandFindParent( parents, i );
		int newIslandId = islandIds[rootMap[root]];
		Body& body = bodies[bodyId];
		Island& newIsland = islands[newIslandId];
		body.islandId = newIslandId;
		body.islandIndex = (int)newIsland.bodies.size();
		newIsland.bodies.push_back( bodyId );
	}

	for ( const ContactLink& link : baseContacts )
	{
		Contact& contact = contacts[link.contactId];
		int targetIslandId = bodies[link.bodyIdA].islandId != B2_NULL_INDEX ? bodies[link.bodyIdA].islandId : bodies[link.bodyIdB].islandId;
		Island& targetIsland = islands[targetIslandId];
		contact.islandId = targetIslandId;
		contact.islandIndex = (int)targetIsland.contacts.size();
		targetIsland.contacts.push_back( link );
	}

	for ( const JointLink& link : baseJoints )
	{
		Joint& joint = joints[link.jointId];
		int targetIslandId = bodies[link.bodyIdA].islandId != B2_NULL_INDEX ? bodies[link.bodyIdA].islandId : bodies[link.bodyIdB].islandId;
		Island& targetIsland = islands[targetIslandId];
		joint.islandId = targetIslandId;
		joint.islandIndex = (int)targetIsland.joints.size();
		targetIsland.joints.push_back( link );
	}

	DestroyIsland( baseId );
}

// ---------------------------------------------------------------------------------------------
// Solver sets and sleep (reference src/solver_set.c)

void World::DestroySolverSet( int setIndex )
{
	SolverSet& set = solverSets[setIndex];
	set.bodyIds.clear();
	set.contactIds.clear();
	set.jointIds.clear();
	set.islandIds.clear();
	set.setIndex = B2_NULL_INDEX;
	solverSetIdPool.Free( setIndex );
}

void World::WakeSolverSet( int setIndex )
{
	islandsDirty = true;
	// Waking rewrites body states, so the host copy must be current first.
	PullBodies();

	SolverSet& awakeSet = solverSets[kAwakeSet];
	SolverSet& disabledSet = solverSets[kDisabledSet];

	std::vector<int> setBodies = solverSets[setIndex].bodyIds;
	for ( int bodyId : setBodies )
	{
		Body& body = bodies[bodyId];
		body.setIndex = kAwakeSet;
		body.localIndex = (int)awakeSet.bodyIds.size();
		bodyAux[bodyId].sleepTime = 0.0f;
		awakeSet.bodyIds.push_back( bodyId );

		BodyState& state = bodyStates[bodyId];
		state = BodyState{ { 0.0f, 0.0f }, 0.0f, bodyAux[bodyId].flags, { 0.0f, 0.0f }, { 1.0f, 0.0f } };

		int contactKey = body.headContactKey;
		while ( contactKey != B2_NULL_INDEX )
		{
			int edgeIndex = contactKey & 1;
			int contactId = contactKey >> 1;
			Contact& contact = contacts[contactId];
			contactKey = contact.edges[edgeIndex].nextKey;

			if ( contact.setIndex != kDisabledSet )
			{
				continue;
			}

			int localIndex = contact.localIndex;
			contact.setIndex = kAwakeSet;
			contact.localIndex = (int)awakeSet.contactIds.size();
			awakeSet.contactIds.push_back( contactId );

			int movedLocalIndex = RemoveSwap( disabledSet.contactIds, localIndex );
			if ( movedLocalIndex != B2_NULL_INDEX )
			{
				contacts[disabledSet.contactIds[localIndex]].localIndex = localIndex;
			}
		}
	}

	std::vector<int> setContacts = solverSets[setIndex].contactIds;
	for ( int contactId : setContacts )
	{
		Contact& contact = contacts[contactId];
		AddContactToGraph( &contact );
		contact.setIndex = kAwakeSet;
	}

	std::vector<int> setJoints = solverSets[setIndex].jointIds;
	for ( int jointId : setJoints )
	{
		Joint& joint = joints[jointId];
		AddJointToGraph( &joint );
		joint.setIndex = kAwakeSet;
	}

	std::vector<int> setIslands = solverSets[setIndex].islandIds;
	for ( int islandId : setIslands )
	{
		Island& island = islands[islandId];
		island.setIndex = kAwakeSet;
		island.localIndex = (int)solverSets[kAwakeSet].islandIds.size();
		solverSets[kAwakeSet].islandIds.push_back( islandId );
	}

	DestroySolverSet( setIndex );
	bodiesDirty = true;
	listsDirty = true;
}

void World::RemoveBodyFromAwakeSet( Body* body )
{
	SolverSet& awakeSet = solverSets[kAwakeSet];
	int localIndex = body->localIndex;
	int movedIndex = RemoveSwap( awakeSet.bodyIds, localIndex );
	if ( movedIndex != B2_NULL_INDEX )
	{
		bodies[awakeSet.bodyIds[localIndex]].localIndex = localIndex;
	}
}

void World::TrySleepIsland( int islandId )
{
	islandsDirty = true;
	Island* island = &islands[islandId];
	if ( island->constraintRemoveCount > 0 && island->bodies.size() > 1 )
	{
		return;
	}

	int sleepSetId = solverSetIdPool.Alloc();
	if ( sleepSetId == (int)solverSets.size() )
	{
		solverSets.emplace_back();
	}
	SolverSet& sleepSet = solverSets[sleepSetId];
	sleepSet = SolverSet{};
	sleepSet.setIndex = sleepSetId;

	SolverSet& awakeSet = solverSets[kAwakeSet];
	SolverSet& disabledSet = solverSets[kDisabledSet];

	for ( int bodyId : island->bodies )
	{
		Body* body = &bodies[bodyId];

		if ( body->bodyMoveIndex != B2_NULL_INDEX )
		{
			moveEvents[body->bodyMoveIndex].fellAsleep = true;
			body->bodyMoveIndex = B2_NULL_INDEX;
		}

		int sleepBodyIndex = (int)sleepSet.bodyIds.size();
		sleepSet.bodyIds.push_back( bodyId );
		RemoveBodyFromAwakeSet( body );
		body->setIndex = sleepSetId;
		body->localIndex = sleepBodyIndex;

		int contactKey = body->headContactKey;
		while ( contactKey != B2_NULL_INDEX )
		{
			int contactId = contactKey >> 1;
			int edgeIndex = contactKey & 1;
			Contact* contact = &contacts[contactId];
			contactKey = contact->edges[edgeIndex].nextKey;

			if ( contact->setIndex == kDisabledSet )
			{
				continue;
			}
			if ( contact->colorIndex != B2_NULL_INDEX )
			{
				continue;
			}

			int otherEdgeIndex = edgeIndex ^ 1;
			int otherBodyId = contact->edges[otherEdgeIndex].bodyId;
			if ( bodies[otherBodyId].setIndex == kAwakeSet )
			{
				continue;
			}

			int localIndex = contact->localIndex;
			contact->setIndex = kDisabledSet;
			contact->localIndex = (int)disabledSet.contactIds.size();
			disabledSet.contactIds.push_back( contactId );

			int movedLocalIndex = RemoveSwap( awakeSet.contactIds, localIndex );
			if ( movedLocalIndex != B2_NULL_INDEX )
			{
				contacts[awakeSet.contactIds[localIndex]].localIndex = localIndex;
			}
		}
	}

	for ( const ContactLink& link : island->contacts )
	{
		Contact* contact = &contacts[link.contactId];
		int colorIndex = contact->colorIndex;
		GraphColor& color = colors[colorIndex];

		if ( colorIndex != kOverflowIndex )
		{
			color.bodySet.ClearBit( contact->edges[0].bodyId );
			color.bodySet.ClearBit( contact->edges[1].bodyId );
		}

		int localIndex = contact->localIndex;
		int sleepContactIndex = (int)sleepSet.contactIds.size();
		sleepSet.contactIds.push_back( link.contactId );

		int movedLocalIndex = RemoveSwap( color.contactIds, localIndex );
		if ( movedLocalIndex != B2_NULL_INDEX )
		{
			contacts[color.contactIds[localIndex]].localIndex = localIndex;
		}

		contact->setIndex = sleepSetId;
		contact->colorIndex = B2_NULL_INDEX;
		contact->localIndex = sleepContactIndex;
	}

	for ( const JointLink& link : island->joints )
	{
		Joint* joint = &joints[link.jointId];
		int colorIndex = joint->colorIndex;
		int localIndex = joint->localIndex;
		GraphColor& color = colors[colorIndex];

		if ( colorIndex != kOverflowIndex )
		{
			color.bodySet.ClearBit( joint->edges[0].bodyId );
			color.bodySet.ClearBit( joint->edges[1].bodyId );
		}

		int sleepJointIndex = (int)sleepSet.jointIds.size();
		sleepSet.jointIds.push_back( link.jointId );

		int movedIndex = RemoveSwap( color.jointIds, localIndex );
		if ( movedIndex != B2_NULL_INDEX )
		{
			joints[color.jointIds[localIndex]].localIndex = localIndex;
		}

		joint->setIndex = sleepSetId;
		joint->colorIndex = B2_NULL_INDEX;
		joint->localIndex = sleepJointIndex;
	}

	{
		int islandIndex = island->localIndex;
		sleepSet.islandIds.push_back( islandId );
		int movedIslandIndex = RemoveSwap( awakeSet.islandIds, islandIndex );
		if ( movedIslandIndex != B2_NULL_INDEX )
		{
			islands[awakeSet.islandIds[islandIndex]].localIndex = islandIndex;
		}
		island->setIndex = sleepSetId;
		island->localIndex = 0;
	}

	if ( splitIslandId == islandId )
	{
		splitIslandId = B2_NULL_INDEX;
	}

	listsDirty = true;
}

void World::MergeSolverSets( int setId1, int setId2 )
{
	if ( solverSets[setId1].bodyIds.size() < solverSets[setId2].bodyIds.size() )
	{
		std::swap( setId1, setId2 );
	}

	SolverSet& set1 = solverSets[setId1];
	SolverSet& set2 = solverSets[setId2];

	for ( int bodyId : set2.bodyIds )
	{
		Body& body = bodies[bodyId];
		body.setIndex = setId1;
		body.localIndex = (int)set1.bodyIds.size();
		set1.bodyIds.push_back( bodyId );
	}

	for ( int contactId : set2.contactIds )
	{
		Contact& contact = contacts[contactId];
		contact.setIndex = setId1;
		contact.localIndex = (int)set1.contactIds.size();
		set1.contactIds.push_back( contactId );
	}

	for ( int jointId : set2.jointIds )
	{
		Joint& joint = joints[jointId];
		joint.setIndex = setId1;
		joint.localIndex = (int)set1.jointIds.size();
		set1.jointIds.push_back( jointId );
	}

	for ( int islandId : set2.islandIds )
	{
		Island& island = islands[islandId];
		island.setIndex = setId1;
		island.localIndex = (int)set1.islandIds.size();
		set1.islandIds.push_back( islandId );
	}

	DestroySolverSet( setId2 );
}

bool World::WakeBody( Body* body )
{
	if ( body->setIndex >= kFirstSleepingSet )
	{
		WakeSolverSet( body->setIndex );
		return true;
	}
	return false;
}

// ---------------------------------------------------------------------------------------------
// Bodies (reference src/body.c:174-332, 524-656)

b2BodyId World::CreateBody( const b2BodyDef* def )
{
	if ( locked )
	{
		return b2_nullBodyId;
	}

	PullBodies();

	bool isAwake = ( def->isAwake || def->enableSleep == false ) && def->isEnabled;

	int setId;
	if ( def->isEnabled == false )
	{
		setId = kDisabledSet;
	}
	else if ( def->type == b2_staticBody )
	{
		setId = kStaticSet;
	}
	else if ( isAwake )
	{
		setId = kAwakeSet;
	}
	else
	{
		setId = solverSetIdPool.Alloc();
		if ( setId == (int)solverSets.size() )
		{
			solverSets.emplace_back();
		}
		solverSets[setId] = SolverSet{};
		solverSets[setId].setIndex = setId;
	}

	int bodyId = bodyIdPool.Alloc();

	uint32_t lockFlags = 0;
	lockFlags |= def->motionLocks.linearX ? kLockLinearX : 0;
	lockFlags |= def->motionLocks.linearY ? kLockLinearY : 0;
	lockFlags |= def->motionLocks.angularZ ? kLockAngularZ : 0;

	if ( bodyId == (int)bodies.size() )
	{
		bodies.emplace_back();
		bodySims.emplace_back();
		bodyStates.emplace_back();
		bodyAux.emplace_back();
	}

	BodySim& sim = bodySims[bodyId];
	sim = BodySim{};
	sim.transform.p = def->position;
	sim.transform.q = def->rotation;
	sim.center = def->position;
	sim.rotation0 = sim.transform.q;
	sim.center0 = sim.center;
	sim.minExtent = B2_HUGE;
	sim.maxExtent = 0.0f;
	sim.linearDamping = def->linearDamping;
	sim.angularDamping = def->angularDamping;
	sim.gravityScale = def->gravityScale;
	sim.flags = lockFlags;
	sim.flags |= def->isBullet ? kIsBullet : 0;
	sim.flags |= def->allowFastRotation ? kAllowFastRotation : 0;
	sim.flags |= def->type == b2_dynamicBody ? kDynamicFlag : 0;
	sim.flags |= def->enableSleep ? kEnableSleep : 0;
	sim.flags |= def->enableContactRecycling ? kEnableContactRecycling : 0;
	sim.headShapeId = B2_NULL_INDEX;

	BodyState& state = bodyStates[bodyId];
	state = BodyState{ { 0.0f, 0.0f }, 0.0f, 0u, { 0.0f, 0.0f }, { 1.0f, 0.0f } };
	if ( setId == kAwakeSet )
	{
		state.linearVelocity = def->linearVelocity;
		state.angularVelocity = def->angularVelocity;
		state.flags = sim.flags;
	}

	Body& body = bodies[bodyId];
	uint16_t gen = body.generation;
	body = Body{};
	body.generation = gen + 1;
	if ( def->name != nullptr )
	{
		strncpy( body.name, def->name, B2_NAME_LENGTH );
		body.name[B2_NAME_LENGTH] = 0;
	}
	body.userData = def->userData;
	body.setIndex = setId;
	body.localIndex = (int)solverSets[setId].bodyIds.size();
	solverSets[setId].bodyIds.push_back( bodyId );
	body.id = bodyId;
	body.type = def->type;
	body.flags = sim.flags;

	BodyAux& aux = bodyAux[bodyId];
	aux = BodyAux{};
	aux.sleepThreshold = def->sleepThreshold;
	aux.sleepTime = 0.0f;
	aux.safetyFactor = def->safetyFactor;
	aux.flags = sim.flags;
	aux.type = def->type;

	if ( setId >= kAwakeSet )
	{
		CreateIslandForBody( setId, &body );
	}

	bodiesDirty = true;
	listsDirty = true;

	return b2BodyId{ bodyId + 1, worldId, body.generation };
}

void World::UpdateBodyMassData( Body* body )
{
	BodySim& sim = bodySims[body->id];
	body->flags &= ~kDirtyMass;
	bodyAux[body->id].flags &= ~kDirtyMass;

	body->mass = 0.0f;
	body->inertia = 0.0f;
	sim.invMass = 0.0f;
	sim.invInertia = 0.0f;
	sim.localCenter = b2Vec2{ 0.0f, 0.0f };
	sim.minExtent = B2_HUGE;
	sim.maxExtent = 0.0f;

	if ( body->type != b2_dynamicBody )
	{
		sim.center = sim.transform.p;
		sim.center0 = sim.center;

		if ( body->type == b2_kinematicBody )
		{
			int shapeId = body->headShapeId;
			while ( shapeId != B2_NULL_INDEX )
			{
				ShapeExtent extent = ComputeShapeExtent( shapeData[shapeId], b2Vec2{ 0.0f, 0.0f } );
				sim.minExtent = b2MinFloat( sim.minExtent, extent.minExtent );
				sim.maxExtent = b2MaxFloat( sim.maxExtent, extent.maxExtent );
				shapeId = shapes[shapeId].nextShapeId;
			}
		}
		bodiesDirty = true;
		return;
	}

	std::vector<b2MassData> masses;
	b2Vec2 localCenter = b2Vec2{ 0.0f, 0.0f };
	int shapeId = body->headShapeId;
	while ( shapeId != B2_NULL_INDEX )
	{
		const Shape& s = shapes[shapeId];
		shapeId = s.nextShapeId;
		if ( s.density == 0.0f )
		{
			masses.push_back( b2MassData{ 0.0f, { 0.0f, 0.0f }, 0.0f } );
			continue;
		}
		b2MassData massData = ComputeShapeMass( shapeData[s.id], s.density );
		body->mass += massData.mass;
		localCenter = b2MulAdd( localCenter, massData.mass, massData.center );
		masses.push_back( massData );
	}

	if ( body->mass > 0.0f )
	{
		sim.invMass = 1.0f / body->mass;
		localCenter = b2MulSV( sim.invMass, localCenter );
	}

	for ( const b2MassData& massData : masses )
	{
		if ( massData.mass == 0.0f )
		{
			continue;
		}
		b2Vec2 offset = b2Sub( localCenter, massData.center );
		float inertia = massData.rotationalInertia + massData.mass * b2Dot( offset, offset );
		body->inertia += inertia;
	}

	if ( body->inertia > 0.0f && ( body->flags & kLockAngularZ ) == 0 )
	{
		sim.invInertia = 1.0f / body->inertia;
	}
	else
	{
		body->inertia = 0.0f;
		sim.invInertia = 0.0f;
	}

	b2Vec2 oldCenter = sim.center;
	sim.localCenter = localCenter;
	sim.center = b2TransformWorldPoint( sim.transform, sim.localCenter );
	sim.center0 = sim.center;

	if ( body->setIndex == kAwakeSet )
	{
		BodyState& state = bodyStates[body->id];
		b2Vec2 deltaLinear = b2CrossSV( state.angularVelocity, b2SubPos( sim.center, oldCenter ) );
		state.linearVelocity = b2Add( state.linearVelocity, deltaLinear );
	}

	shapeId = body->headShapeId;
	while ( shapeId != B2_NULL_INDEX )
	{
		ShapeExtent extent = ComputeShapeExtent( shapeData[shapeId], localCenter );
		sim.minExtent = b2MinFloat( sim.minExtent, extent.minExtent );
		sim.maxExtent = b2MaxFloat( sim.maxExtent, extent.maxExtent );
		shapeId = shapes[shapeId].nextShapeId;
	}

	bodiesDirty = true;
}

bool World::ShouldBodiesCollide( Body* bodyA, Body* bodyB )
{
	if ( bodyA->type != b2_dynamicBody && bodyB->type != b2_dynamicBody )
	{
		return false;
	}

	int jointKey;
	int otherBodyId;
	if ( bodyA->jointCount < bodyB->jointCount )
	{
		jointKey = bodyA->headJointKey;
		otherBodyId = bodyB->id;
	}
	else
	{
		jointKey = bodyB->headJointKey;
		otherBodyId = bodyA->id;
	}

	while ( jointKey != B2_NULL_INDEX )
	{
		int jointId = jointKey >> 1;
		int edgeIndex = jointKey & 1;
		int otherEdgeIndex = edgeIndex ^ 1;
		Joint& joint = joints[jointId];
		if ( joint.collideConnected == false && joint.edges[otherEdgeIndex].bodyId == otherBodyId )
		{
			return false;
		}
		jointKey = joint.edges[edgeIndex].nextKey;
	}
	return true;
}

// ---------------------------------------------------------------------------------------------
// Shapes (reference src/shape.c:89-305)

void World::CreateShapeProxy( Shape* shape, bool forcePairCreation )
{
	Body& body = bodies[shape->bodyId];
	ShapeDev& sd = shapeData[shape->id];
	b2Transform transform = bodySims[body.id].transform;

	b2AABB aabb = HostFatShapeAABB( sd, transform, B2_SPECULATIVE_DISTANCE );
	sd.aabb = aabb;
	float margin = body.type == b2_staticBody ? B2_SPECULATIVE_DISTANCE : sd.aabbMargin;
	b2AABB fatAABB;
	fatAABB.lowerBound.x = aabb.lowerBound.x - margin;
	fatAABB.lowerBound.y = aabb.lowerBound.y - margin;
	fatAABB.upperBound.x = aabb.upperBound.x + margin;
	fatAABB.upperBound.y = aabb.upperBound.y + margin;
	sd.fatAABB = fatAABB;

	shape->proxyKey = broadPhase.CreateProxy( fatAABB, body.type, shape->filter.categoryBits, shape->id, forcePairCreation );
	sd.flags |= kShapeHasProxy;
	shapesDirty = true;
}

b2ShapeId World::CreateShape( b2BodyId bodyId, const b2ShapeDef* def, const void* geometry, b2ShapeType shapeType )
{
	if ( locked )
	{
		return b2_nullShapeId;
	}

	Body* body = GetBody( bodyId );
	if ( body == nullptr )
	{
		return b2_nullShapeId;
	}

	PullBodies();
	PullShapes();

	int shapeId = CreateShapeInternal( body, def, geometry, shapeType );
	Shape* shape = &shapes[shapeId];

	if ( def->updateBodyMass == true )
	{
		UpdateBodyMassData( body );
	}
	else if ( ( body->flags & kDirtyMass ) == 0 )
	{
		body->flags |= kDirtyMass;
		bodyAux[body->id].flags |= kDirtyMass;
	}

	shapesDirty = true;
	bodiesDirty = true;
	return b2ShapeId{ shapeId + 1, worldId, shape->generation };
}

// Shape record, proxy and body shape-list link (reference src/shape.c:108-207). Mass is the
// caller's business: public shapes update it, chain segments do not.
int World::CreateShapeInternal( Body* body, const b2ShapeDef* def, const void* geometry, b2ShapeType shapeType )
{
	int shapeId = shapeIdPool.Alloc();
	if ( shapeId == (int)shapes.size() )
	{
		shapes.emplace_back();
		shapeData.emplace_back();
	}

	Shape* shape = &shapes[shapeId];
	uint16_t gen = shape->generation;
	*shape = Shape{};
	shape->generation = gen + 1;

	ShapeDev& sd = shapeData[shapeId];
	memset( &sd, 0, sizeof( sd ) );
	switch ( shapeType )
	{
		case b2_capsuleShape:
			sd.capsule = *(const b2Capsule*)geometry;
			break;
		case b2_circleShape:
			sd.circle = *(const b2Circle*)geometry;
			break;
		case b2_polygonShape:
			sd.polygon = *(const b2Polygon*)geometry;
			break;
		case b2_segmentShape:
			sd.segment = *(const b2Segment*)geometry;
			break;
		case b2_chainSegmentShape:
			sd.chainSegment = *(const b2ChainSegment*)geometry;
			break;
		default:
			break;
	}

	shape->id = shapeId;
	shape->bodyId = body->id;
	shape->type = shapeType;
	shapeFilterVersion += 1;
	shape->density = def->density;
	shape->material = def->material;
	shape->filter = def->filter;
	shape->userData = def->userData;
	shape->enableSensorEvents = def->enableSensorEvents;
	shape->enableContactEvents = def->enableContactEvents;
	shape->enableCustomFiltering = def->enableCustomFiltering;
	shape->enableHitEvents = def->enableHitEvents;
	shape->enablePreSolveEvents = def->enablePreSolveEvents;
	shape->proxyKey = B2_NULL_INDEX;

	sd.type = shapeType;
	sd.bodyId = body->id;
	sd.flags = ( def->enableHitEvents ? kShapeEnableHitEvents : 0 ) | ( def->enableContactEvents ? kShapeEnableContactEvents : 0 ) |
			   ( def->isSensor ? kShapeIsSensor : 0 ) | ( def->enableSensorEvents ? kShapeEnableSensorEvents : 0 );
	sd.friction = def->material.friction;
	sd.restitution = def->material.restitution;
	sd.rollingResistance = def->material.rollingResistance;
	sd.tangentSpeed = def->material.tangentSpeed;
	sd.filter = def->filter;
	sd.localCentroid = ShapeCentroid( sd );
	sd.aabbMargin = ComputeShapeMargin( sd );
	sd.enlargedAABB = 0;

	if ( def->material.restitution > 0.0f )
	{
		restitutionShapeCount += 1;
	}

	if ( body->setIndex != kDisabledSet )
	{
		CreateShapeProxy( shape, def->invokeContactCreation || def->isSensor );
	}

	// head insert into the body's shape list
	if ( body->headShapeId != B2_NULL_INDEX )
	{
		shapes[body->headShapeId].prevShapeId = shapeId;
	}
	shape->prevShapeId = B2_NULL_INDEX;
	shape->nextShapeId = body->headShapeId;
	sd.nextShapeId = body->headShapeId;
	body->headShapeId = shapeId;
	bodySims[body->id].headShapeId = shapeId;
	body->shapeCount += 1;

	// sensors (reference src/shape.c:189-197)
	if ( def->isSensor )
	{
		shape->sensorIndex = (int)sensors.size();
		sensors.emplace_back();
		sensors.back().shapeId = shapeId;
	}
	return shapeId;
}

// Chain of one-sided segments with ghost vertices (reference src/shape.c:439-571).
b2ChainId World::CreateChain( b2BodyId bodyId, const b2ChainDef* def )
{
	if ( locked )
	{
		return b2ChainId{ 0, 0, 0 };
	}

	Body* body = GetBody( bodyId );
	if ( body == nullptr || def->count < 4 || ( def->materialCount != 1 && def->materialCount != def->count ) )
	{
		return b2ChainId{ 0, 0, 0 };
	}

	PullBodies();
	PullShapes();

	int chainId = chainIdPool.Alloc();
	if ( chainId == (int)chains.size() )
	{
		chains.emplace_back();
	}

	ChainShape* chain = &chains[chainId];
	chain->id = chainId;
	chain->bodyId = body->id;
	chain->nextChainId = body->headChainId;
	chain->generation += 1;
	chain->materials.assign( def->materials, def->materials + def->materialCount );
	body->headChainId = chainId;

	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.userData = def->userData;
	shapeDef.filter = def->filter;
	shapeDef.enableSensorEvents = def->enableSensorEvents;
	shapeDef.enableContactEvents = false;
	shapeDef.enableHitEvents = false;

	const b2Vec2* points = def->points;
	int n = def->count;
	int materialCount = def->materialCount;

	auto addSegment = [&]( b2Vec2 ghost1, b2Vec2 p1, b2Vec2 p2, b2Vec2 ghost2, int materialIndex ) {
		b2ChainSegment segment;
		segment.ghost1 = ghost1;
		segment.segment.point1 = p1;
		segment.segment.point2 = p2;
		segment.ghost2 = ghost2;
		segment.chainId = chainId;
		shapeDef.material = def->materials[materialCount == 1 ? 0 : materialIndex];
		int shapeId = CreateShapeInternal( body, &shapeDef, &segment, b2_chainSegmentShape );
		chains[chainId].shapeIndices.push_back( shapeId );
	};

	if ( def->isLoop )
	{
		int prevIndex = n - 1;
		for ( int i = 0; i < n - 2; ++i )
		{
			addSegment( points[prevIndex], points[i], points[i + 1], points[i + 2], i );
			prevIndex = i;
		}
		addSegment( points[n - 3], points[n - 2], points[n - 1], points[0], n - 2 );
		addSegment( points[n - 2], points[n - 1], points[0], points[1], n - 1 );
	}
	else
	{
		for ( int i = 0; i < n - 3; ++i )
		{
			// the material belongs to the leading point of the solid segment
			addSegment( points[i], points[i + 1], points[i + 2], points[i + 3], i + 1 );
		}
	}

	shapesDirty = true;
	bodiesDirty = true;
	return b2ChainId{ chainId + 1, worldId, chains[chainId].generation };
}

// ---------------------------------------------------------------------------------------------
// Contacts (reference src/contact.c:214-501)

void World::CreateContact( Shape* shapeA, Shape* shapeB )
{
	b2ShapeType type1 = shapeA->type;
	b2ShapeType type2 = shapeB->type;

	if ( s_registers[type1][type2].exists == false )
	{
		return;
	}

	if ( s_registers[type1][type2].primary == false )
	{
		CreateContact( shapeB, shapeA );
		return;
	}

	Body* bodyA = &bodies[shapeA->bodyId];
	Body* bodyB = &bodies[shapeB->bodyId];

	int setIndex;
	if ( bodyA->setIndex == kAwakeSet || bodyB->setIndex == kAwakeSet )
	{
		setIndex = kAwakeSet;
	}
	else
	{
		setIndex = kDisabledSet;
	}

	SolverSet& set = solverSets[setIndex];

	int contactId = contactIdPool.Alloc();
	if ( contactId == (int)contacts.size() )
	{
		contacts.emplace_back();
	}

	int shapeIdA = shapeA->id;
	int shapeIdB = shapeB->id;

	Contact* contact = &contacts[contactId];
	uint32_t gen = contact->generation;
	*contact = Contact{};
	contact->contactId = contactId;
	contact->generation = gen + 1;
	contact->setIndex = setIndex;
	contact->colorIndex = B2_NULL_INDEX;
	contact->localIndex = (int)set.contactIds.size();
	contact->islandId = B2_NULL_INDEX;
	contact->islandIndex = B2_NULL_INDEX;
	contact->shapeIdA = shapeIdA;
	contact->shapeIdB = shapeIdB;
	contact->flags = 0;

	if ( ( bodyA->flags & kEnableContactRecycling ) != 0 && ( bodyB->flags & kEnableContactRecycling ) != 0 )
	{
		contact->flags |= kContactRecycle;
	}

	if ( shapeA->enableContactEvents || shapeB->enableContactEvents )
	{
		contact->flags |= kContactEnableContactEvents;
	}

	{
		contact->edges[0].bodyId = shapeA->bodyId;
		contact->edges[0].prevKey = B2_NULL_INDEX;
		contact->edges[0].nextKey = bodyA->headContactKey;
		int keyA = ( contactId << 1 ) | 0;
		int headContactKey = bodyA->headContactKey;
		if ( headContactKey != B2_NULL_INDEX )
		{
			contacts[headContactKey >> 1].edges[headContactKey & 1].prevKey = keyA;
		}
		bodyA->headContactKey = keyA;
		bodyA->contactCount += 1;
	}

	{
		contact->edges[1].bodyId = shapeB->bodyId;
		contact->edges[1].prevKey = B2_NULL_INDEX;
		contact->edges[1].nextKey = bodyB->headContactKey;
		int keyB = ( contactId << 1 ) | 1;
		int headContactKey = bodyB->headContactKey;
		if ( headContactKey != B2_NULL_INDEX )
		{
			contacts[headContactKey >> 1].edges[headContactKey & 1].prevKey = keyB;
		}
		bodyB->headContactKey = keyB;
		bodyB->contactCount += 1;
	}

	broadPhase.pairSet.Insert( B2G_SHAPE_PAIR_KEY( shapeIdA, shapeIdB ) );

	set.contactIds.push_back( contactId );

	ContactInit init;
	init.contactId = contactId;
	init.shapeIdA = shapeIdA;
	init.shapeIdB = shapeIdB;
	init.bodyIdA = shapeA->bodyId;
	init.bodyIdB = shapeB->bodyId;
	init.friction = MixFriction( *shapeA, *shapeB );
	init.restitution = MixRestitution( *shapeA, *shapeB );
	init.simFlags = contact->flags;
	if ( shapeA->enablePreSolveEvents || shapeB->enablePreSolveEvents )
	{
		init.simFlags |= kSimEnablePreSolveEvents;
	}
	pendingContactInits.push_back( init );

	listsDirty = true;
}

void World::DestroyContact( Contact* contact )
{
	broadPhase.pairSet.Erase( B2G_SHAPE_PAIR_KEY( contact->shapeIdA, contact->shapeIdB ) );

	ContactEdge* edgeA = contact->edges + 0;
	ContactEdge* edgeB = contact->edges + 1;
	int bodyIdA = edgeA->bodyId;
	int bodyIdB = edgeB->bodyId;
	Body* bodyA = &bodies[bodyIdA];
	Body* bodyB = &bodies[bodyIdB];

	bool touching = ( contact->flags & kContactTouching ) != 0;

	// end touch event (reference src/contact.c:389-412)
	if ( touching && ( contact->flags & kContactEnableContactEvents ) != 0 )
	{
		contactEndEvents[endEventArrayIndex].push_back(
			b2ContactEndTouchEvent{ ShapeIdOf( contact->shapeIdA ), ShapeIdOf( contact->shapeIdB ), ContactIdOf( *contact ) } );
	}

	if ( edgeA->prevKey != B2_NULL_INDEX )
	{
		contacts[edgeA->prevKey >> 1].edges[edgeA->prevKey & 1].nextKey = edgeA->nextKey;
	}
	if ( edgeA->nextKey != B2_NULL_INDEX )
	{
		contacts[edgeA->nextKey >> 1].edges[edgeA->nextKey & 1].prevKey = edgeA->prevKey;
	}

	int contactId = contact->contactId;
	int edgeKeyA = ( contactId << 1 ) | 0;
	if ( bodyA->headContactKey == edgeKeyA )
	{
		bodyA->headContactKey = edgeA->nextKey;
	}
	bodyA->contactCount -= 1;

	if ( edgeB->prevKey != B2_NULL_INDEX )
	{
		contacts[edgeB->prevKey >> 1].edges[edgeB->prevKey & 1].nextKey = edgeB->nextKey;
	}
	if ( edgeB->nextKey != B2_NULL_INDEX )
	{
		contacts[edgeB->nextKey >> 1].edges[edgeB->nextKey & 1].prevKey = edgeB->prevKey;
	}

	int edgeKeyB = ( contactId << 1 ) | 1;
	if ( bodyB->headContactKey == edgeKeyB )
	{
		bodyB->headContactKey = edgeB->nextKey;
	}
	bodyB->contactCount -= 1;

	if ( contact->islandId != B2_NULL_INDEX )
	{
		UnlinkContact( contact );
	}

	if ( contact->colorIndex != B2_NULL_INDEX )
	{
		RemoveContactFromGraph( bodyIdA, bodyIdB, contact->colorIndex, contact->localIndex );
	}
	else
	{
		SolverSet& set = solverSets[contact->setIndex];
		int movedIndex = RemoveSwap( set.contactIds, contact->localIndex );
		if ( movedIndex != B2_NULL_INDEX )
		{
			contacts[set.contactIds[contact->localIndex]].localIndex = contact->localIndex;
		}
	}

	contact->contactId = B2_NULL_INDEX;
	contact->setIndex = B2_NULL_INDEX;
	contact->colorIndex = B2_NULL_INDEX;
	contact->localIndex = B2_NULL_INDEX;
	contactIdPool.Free( contactId );

	if ( touching )
	{
		WakeBody( bodyA );
		WakeBody( bodyB );
	}

	listsDirty = true;
}

// ---------------------------------------------------------------------------------------------
// Destruction (reference src/body.c:120-156, 346-436; src/joint.c:721-840; src/island.c:311-337)

void World::UnlinkJoint( Joint* joint )
{
	islandsDirty = true;
	if ( joint->islandId == B2_NULL_INDEX )
	{
		return;
	}
	Island* island = &islands[joint->islandId];
	int removeIndex = joint->islandIndex;
	int movedIndex = RemoveSwap( island->joints, removeIndex );
	if ( movedIndex != B2_NULL_INDEX )
	{
		joints[island->joints[removeIndex].jointId].islandIndex = removeIndex;
	}
	joint->islandId = B2_NULL_INDEX;
	joint->islandIndex = B2_NULL_INDEX;
	island->constraintRemoveCount += 1;
}

void World::RemoveBodyFromIsland( Body* body )
{
	islandsDirty = true;
	if ( body->islandId == B2_NULL_INDEX )
	{
		return;
	}
	int islandId = body->islandId;
	Island* island = &islands[islandId];
	int localIndex = body->islandIndex;
	int movedBodyId = island->bodies.back();
	island->bodies[localIndex] = movedBodyId;
	bodies[movedBodyId].islandIndex = localIndex;
	island->bodies.pop_back();
	if ( island->bodies.empty() )
	{
		DestroyIsland( islandId );
	}
	body->islandId = B2_NULL_INDEX;
	body->islandIndex = B2_NULL_INDEX;
}

void World::DestroyJointInternal( Joint* joint )
{
	int jointId = joint->jointId;
	JointEdge* edgeA = joint->edges + 0;
	JointEdge* edgeB = joint->edges + 1;
	Body* bodyA = &bodies[edgeA->bodyId];
	Body* bodyB = &bodies[edgeB->bodyId];

	// remove from the two bodies' joint lists
	if ( edgeA->prevKey != B2_NULL_INDEX )
	{
		joints[edgeA->prevKey >> 1].edges[edgeA->prevKey & 1].nextKey = edgeA->nextKey;
	}
	if ( edgeA->nextKey != B2_NULL_INDEX )
	{
		joints[edgeA->nextKey >> 1].edges[edgeA->nextKey & 1].prevKey = edgeA->prevKey;
	}
	if ( bodyA->headJointKey == ( ( jointId << 1 ) | 0 ) )
	{
		bodyA->headJointKey = edgeA->nextKey;
	}
	bodyA->jointCount -= 1;

	if ( edgeB->prevKey != B2_NULL_INDEX )
	{
		joints[edgeB->prevKey >> 1].edges[edgeB->prevKey & 1].nextKey = edgeB->nextKey;
	}
	if ( edgeB->nextKey != B2_NULL_INDEX )
	{
		joints[edgeB->nextKey >> 1].edges[edgeB->nextKey & 1].prevKey = edgeB->prevKey;
	}
	if ( bodyB->headJointKey == ( ( jointId << 1 ) | 1 ) )
	{
		bodyB->headJointKey = edgeB->nextKey;
	}
	bodyB->jointCount -= 1;

	if ( joint->islandId != B2_NULL_INDEX )
	{
		UnlinkJoint( joint );
	}

	// remove from the solver set that owns the joint
	int setIndex = joint->setIndex;
	int localIndex = joint->localIndex;
	if ( setIndex == kAwakeSet )
	{
		RemoveJointFromGraph( edgeA->bodyId, edgeB->bodyId, joint->colorIndex, localIndex );
	}
	else
	{
		SolverSet& set = solverSets[setIndex];
		int movedIndex = RemoveSwap( set.jointIds, localIndex );
		if ( movedIndex != B2_NULL_INDEX )
		{
			joints[set.jointIds[localIndex]].localIndex = localIndex;
		}
	}

	joint->setIndex = B2_NULL_INDEX;
	joint->localIndex = B2_NULL_INDEX;
	joint->colorIndex = B2_NULL_INDEX;
	joint->jointId = B2_NULL_INDEX;
	jointIdPool.Free( jointId );

	WakeBody( bodyA );
	WakeBody( bodyB );
	jointsDirty = true;
	listsDirty = true;
}

void World::DestroyJoint( b2JointId id )
{
	if ( locked )
	{
		return;
	}
	int index = id.index1 - 1;
	if ( index < 0 || index >= (int)joints.size() || joints[index].jointId != index || joints[index].generation != id.generation )
	{
		return;
	}
	// the device holds the other joints' impulses; the host copy must be current before the
	// joint array is uploaded again
	PullJoints();
	DestroyJointInternal( &joints[index] );
}

void World::DestroyBody( b2BodyId id )
{
	if ( locked )
	{
		return;
	}
	Body* body = GetBody( id );
	if ( body == nullptr )
	{
		return;
	}
	PullBodies();
	PullShapes();
	PullJoints();

	// attached joints
	int edgeKey = body->headJointKey;
	while ( edgeKey != B2_NULL_INDEX )
	{
		Joint* joint = &joints[edgeKey >> 1];
		edgeKey = joint->edges[edgeKey & 1].nextKey;
		DestroyJointInternal( joint );
	}

	// attached contacts
	int contactKey = body->headContactKey;
	while ( contactKey != B2_NULL_INDEX )
	{
		Contact* contact = &contacts[contactKey >> 1];
		contactKey = contact->edges[contactKey & 1].nextKey;
		DestroyContact( contact );
	}

	// shapes and their broadphase proxies
	int shapeId = body->headShapeId;
	while ( shapeId != B2_NULL_INDEX )
	{
		Shape* shape = &shapes[shapeId];
		if ( shape->sensorIndex != B2_NULL_INDEX )
		{
			DestroySensor( shape );
		}
		if ( shape->proxyKey != B2_NULL_INDEX )
		{
			broadPhase.DestroyProxy( shape->proxyKey );
			shape->proxyKey = B2_NULL_INDEX;
		}
		shapeData[shapeId].flags &= ~kShapeHasProxy;
		if ( shape->material.restitution > 0.0f )
		{
			restitutionShapeCount -= 1;
		}
		shapeIdPool.Free( shapeId );
		shape->id = B2_NULL_INDEX;
		shapeId = shape->nextShapeId;
	}

	// chains (their shapes went above)
	int chainId = body->headChainId;
	while ( chainId != B2_NULL_INDEX )
	{
		ChainShape* chain = &chains[chainId];
		chain->shapeIndices.clear();
		chain->materials.clear();
		chainIdPool.Free( chainId );
		chain->id = B2_NULL_INDEX;
		chainId = chain->nextChainId;
	}

	RemoveBodyFromIsland( body );

	int setIndex = body->setIndex;
	SolverSet& set = solverSets[setIndex];
	int movedIndex = RemoveSwap( set.bodyIds, body->localIndex );
	if ( movedIndex != B2_NULL_INDEX )
	{
		bodies[set.bodyIds[body->localIndex]].localIndex = body->localIndex;
	}
	if ( setIndex >= kFirstSleepingSet && set.bodyIds.empty() )
	{
		DestroySolverSet( setIndex );
	}

	bodyIdPool.Free( body->id );
	body->setIndex = B2_NULL_INDEX;
	body->localIndex = B2_NULL_INDEX;
	body->id = B2_NULL_INDEX;

	// the static tree copy used by continuous collision and the shape array go up again
	shapesDirty = true;
	listsDirty = true;
}

// b2TransferBody (reference src/solver_set.c:511-545): move a body between solver sets, keeping
// the reference's array order (remove-swap in the source, append to the target).
void World::TransferBody( int targetSetIndex, int sourceSetIndex, Body* body )
{
	if ( targetSetIndex == sourceSetIndex )
	{
		return;
	}
	int sourceIndex = body->localIndex;
	int targetIndex = (int)solverSets[targetSetIndex].bodyIds.size();
	solverSets[targetSetIndex].bodyIds.push_back( body->id );

	body->flags &= ~kTransientFlags;
	bodySims[body->id].flags &= ~kTransientFlags;
	bodyAux[body->id].flags &= ~kTransientFlags;

	SolverSet& source = solverSets[sourceSetIndex];
	int movedIndex = RemoveSwap( source.bodyIds, sourceIndex );
	if ( movedIndex != B2_NULL_INDEX )
	{
		bodies[source.bodyIds[sourceIndex]].localIndex = sourceIndex;
	}
	if ( sourceSetIndex != kAwakeSet && targetSetIndex == kAwakeSet )
	{
		// a body entering the awake set starts from the identity state with its flags
		bodyStates[body->id] = BodyState{ { 0.0f, 0.0f }, 0.0f, body->flags, { 0.0f, 0.0f }, { 1.0f, 0.0f } };
	}
	else if ( targetSetIndex != kAwakeSet )
	{
		// outside the awake set the state is the identity (constraint lanes read a static body's)
		bodyStates[body->id] = BodyState{ { 0.0f, 0.0f }, 0.0f, 0u, { 0.0f, 0.0f }, { 1.0f, 0.0f } };
	}
	body->setIndex = targetSetIndex;
	body->localIndex = targetIndex;
	bodiesDirty = true;
	listsDirty = true;
}

// b2TransferJoint (reference src/solver_set.c:547-605)
void World::TransferJoint( int targetSetIndex, int sourceSetIndex, Joint* joint )
{
	if ( targetSetIndex == sourceSetIndex )
	{
		return;
	}
	int localIndex = joint->localIndex;
	int colorIndex = joint->colorIndex;
	if ( targetSetIndex == kAwakeSet )
	{
		AddJointToGraph( joint );
		joint->setIndex = kAwakeSet;
	}
	else
	{
		SolverSet& target = solverSets[targetSetIndex];
		joint->setIndex = targetSetIndex;
		joint->localIndex = (int)target.jointIds.size();
		joint->colorIndex = B2_NULL_INDEX;
		target.jointIds.push_back( joint->jointId );
	}
	if ( sourceSetIndex == kAwakeSet )
	{
		RemoveJointFromGraph( joint->edges[0].bodyId, joint->edges[1].bodyId, colorIndex, localIndex );
	}
	else
	{
		SolverSet& source = solverSets[sourceSetIndex];
		int movedIndex = RemoveSwap( source.jointIds, localIndex );
		if ( movedIndex != B2_NULL_INDEX )
		{
			joints[source.jointIds[localIndex]].localIndex = localIndex;
		}
	}
	listsDirty = true;
}

void World::DestroyBodyContacts( Body* body )
{
	int contactKey = body->headContactKey;
	while ( contactKey != B2_NULL_INDEX )
	{
		Contact* contact = &contacts[contactKey >> 1];
		contactKey = contact->edges[contactKey & 1].nextKey;
		DestroyContact( contact );
	}
}

void World::DestroyShapeProxy( Shape* shape )
{
	if ( shape->proxyKey != B2_NULL_INDEX )
	{
		broadPhase.DestroyProxy( shape->proxyKey );
		shape->proxyKey = B2_NULL_INDEX;
	}
	shapeData[shape->id].flags &= ~kShapeHasProxy;
	shapesDirty = true;
}

// A destroyed sensor reports the end of its overlaps; the last sensor takes its slot
// (reference src/shape.c:365-402, src/sensor.c:392-429).
void World::DestroySensor( Shape* shape )
{
	int index = shape->sensorIndex;
	Sensor& sensor = sensors[index];
	b2ShapeId sensorId = ShapeIdOf( shape->id );
	for ( const SensorVisitor& v : sensor.overlaps2 )
	{
		sensorEndEvents[endEventArrayIndex].push_back( b2SensorEndTouchEvent{ sensorId, b2ShapeId{ v.shapeId + 1, worldId, v.generation } } );
	}
	int last = (int)sensors.size() - 1;
	if ( index != last )
	{
		sensors[index] = std::move( sensors[last] );
		shapes[sensors[index].shapeId].sensorIndex = index;
	}
	sensors.pop_back();
	shape->sensorIndex = B2_NULL_INDEX;
}

void World::SetBodyTypeFlag( Body* body, b2BodyType type )
{
	body->type = type;
	body->flags = type == b2_dynamicBody ? ( body->flags | kDynamicFlag ) : ( body->flags & ~kDynamicFlag );
	bodyAux[body->id].type = type;
}

// b2Body_SetType (reference src/body.c:1134-1316)
void World::SetBodyType( Body* body, b2BodyType type )
{
	b2BodyType originalType = body->type;
	if ( originalType == type )
	{
		return;
	}
	PullBodies();
	PullShapes();
	PullJoints();

	if ( body->setIndex == kDisabledSet )
	{
		SetBodyTypeFlag( body, type );
		SyncBodyFlags( body );
		UpdateBodyMassData( body );
		return;
	}

	DestroyBodyContacts( body );
	WakeBody( body );

	// joints go to the static set first so they take consistent colors below
	int jointKey = body->headJointKey;
	while ( jointKey != B2_NULL_INDEX )
	{
		Joint* joint = &joints[jointKey >> 1];
		jointKey = joint->edges[jointKey & 1].nextKey;
		if ( joint->setIndex == kDisabledSet )
		{
			continue;
		}
		WakeBody( &bodies[joint->edges[0].bodyId] );
		WakeBody( &bodies[joint->edges[1].bodyId] );
		UnlinkJoint( joint );
		TransferJoint( kStaticSet, joint->setIndex, joint );
	}

	SetBodyTypeFlag( body, type );
	TransferBody( type == b2_staticBody ? kStaticSet : kAwakeSet, body->setIndex, body );

	if ( originalType == b2_staticBody )
	{
		CreateIslandForBody( kAwakeSet, body );
	}
	else if ( type == b2_staticBody )
	{
		RemoveBodyFromIsland( body );
	}

	jointKey = body->headJointKey;
	while ( jointKey != B2_NULL_INDEX )
	{
		Joint* joint = &joints[jointKey >> 1];
		jointKey = joint->edges[jointKey & 1].nextKey;
		if ( joint->setIndex == kDisabledSet )
		{
			continue;
		}
		const Body& bodyA = bodies[joint->edges[0].bodyId];
		const Body& bodyB = bodies[joint->edges[1].bodyId];
		if ( bodyA.type == b2_dynamicBody || bodyB.type == b2_dynamicBody )
		{
			TransferJoint( kAwakeSet, kStaticSet, joint );
		}
	}

	// proxies go to the tree of the new type
	int shapeId = body->headShapeId;
	while ( shapeId != B2_NULL_INDEX )
	{
		Shape* shape = &shapes[shapeId];
		shapeId = shape->nextShapeId;
		DestroyShapeProxy( shape );
		CreateShapeProxy( shape, true );
	}

	jointKey = body->headJointKey;
	while ( jointKey != B2_NULL_INDEX )
	{
		int edgeIndex = jointKey & 1;
		Joint* joint = &joints[jointKey >> 1];
		jointKey = joint->edges[edgeIndex].nextKey;
		const Body& other = bodies[joint->edges[edgeIndex ^ 1].bodyId];
		if ( other.setIndex == kDisabledSet )
		{
			continue;
		}
		if ( body->type != b2_dynamicBody && other.type != b2_dynamicBody )
		{
			continue;
		}
		LinkJoint( joint );
	}

	SyncBodyFlags( body );
	UpdateBodyMassData( body );
	shapesDirty = true;
	bodiesDirty = true;
	jointsDirty = true;
	listsDirty = true;
}

// b2Body_Disable (reference src/body.c:1699-1768)
void World::DisableBody( Body* body )
{
	if ( body->setIndex == kDisabledSet )
	{
		return;
	}
	PullBodies();
	PullShapes();
	PullJoints();

	DestroyBodyContacts( body );
	int setIndex = body->setIndex;

	int jointKey = body->headJointKey;
	while ( jointKey != B2_NULL_INDEX )
	{
		Joint* joint = &joints[jointKey >> 1];
		jointKey = joint->edges[jointKey & 1].nextKey;
		if ( joint->setIndex == kDisabledSet )
		{
			continue;
		}
		UnlinkJoint( joint );
		TransferJoint( kDisabledSet, joint->setIndex, joint );
	}

	int shapeId = body->headShapeId;
	while ( shapeId != B2_NULL_INDEX )
	{
		Shape* shape = &shapes[shapeId];
		shapeId = shape->nextShapeId;
		DestroyShapeProxy( shape );
	}

	RemoveBodyFromIsland( body );
	TransferBody( kDisabledSet, setIndex, body );
	shapesDirty = true;
	bodiesDirty = true;
	jointsDirty = true;
	listsDirty = true;
}

// b2Body_Enable (reference src/body.c:1770-1858)
void World::EnableBody( Body* body )
{
	if ( body->setIndex != kDisabledSet )
	{
		return;
	}
	PullBodies();
	PullShapes();
	PullJoints();

	int setId = body->type == b2_staticBody ? kStaticSet : kAwakeSet;
	TransferBody( setId, kDisabledSet, body );

	int shapeId = body->headShapeId;
	while ( shapeId != B2_NULL_INDEX )
	{
		Shape* shape = &shapes[shapeId];
		shapeId = shape->nextShapeId;
		CreateShapeProxy( shape, true );
	}

	if ( setId != kStaticSet )
	{
		CreateIslandForBody( setId, body );
	}

	int jointKey = body->headJointKey;
	while ( jointKey != B2_NULL_INDEX )
	{
		Joint* joint = &joints[jointKey >> 1];
		jointKey = joint->edges[jointKey & 1].nextKey;
		const Body& bodyA = bodies[joint->edges[0].bodyId];
		const Body& bodyB = bodies[joint->edges[1].bodyId];
		if ( bodyA.setIndex == kDisabledSet || bodyB.setIndex == kDisabledSet )
		{
			continue;
		}
		int jointSetId;
		if ( bodyA.setIndex == kStaticSet && bodyB.setIndex == kStaticSet )
		{
			jointSetId = kStaticSet;
		}
		else if ( bodyA.setIndex == kStaticSet )
		{
			jointSetId = bodyB.setIndex;
		}
		else
		{
			jointSetId = bodyA.setIndex;
		}
		TransferJoint( jointSetId, kDisabledSet, joint );
		if ( jointSetId != kStaticSet )
		{
			LinkJoint( joint );
		}
	}
	shapesDirty = true;
	bodiesDirty = true;
	jointsDirty = true;
	listsDirty = true;
}

// b2DestroyShape (reference src/shape.c:320-431): unlink the shape from its body's list (and the
// device's copy of the list), drop its proxy and contacts, free its id.
void World::DestroyShape( Shape* shape, bool updateBodyMass )
{
	PullBodies();
	PullShapes();
	Body* body = &bodies[shape->bodyId];
	int shapeId = shape->id;

	if ( shape->prevShapeId != B2_NULL_INDEX )
	{
		shapes[shape->prevShapeId].nextShapeId = shape->nextShapeId;
		shapeData[shape->prevShapeId].nextShapeId = shape->nextShapeId;
	}
	if ( shape->nextShapeId != B2_NULL_INDEX )
	{
		shapes[shape->nextShapeId].prevShapeId = shape->prevShapeId;
	}
	if ( shapeId == body->headShapeId )
	{
		body->headShapeId = shape->nextShapeId;
		bodySims[body->id].headShapeId = shape->nextShapeId;
	}
	body->shapeCount -= 1;

	if ( shape->proxyKey != B2_NULL_INDEX )
	{
		broadPhase.DestroyProxy( shape->proxyKey );
		shape->proxyKey = B2_NULL_INDEX;
	}
	shapeData[shapeId].flags &= ~kShapeHasProxy;

	int contactKey = body->headContactKey;
	while ( contactKey != B2_NULL_INDEX )
	{
		Contact* contact = &contacts[contactKey >> 1];
		contactKey = contact->edges[contactKey & 1].nextKey;
		if ( contact->shapeIdA == shapeId || contact->shapeIdB == shapeId )
		{
			DestroyContact( contact );
		}
	}

	if ( shape->sensorIndex != B2_NULL_INDEX )
	{
		DestroySensor( shape );
	}

	if ( shape->material.restitution > 0.0f )
	{
		restitutionShapeCount -= 1;
	}
	shapeIdPool.Free( shapeId );
	shape->id = B2_NULL_INDEX;
	shapeData[shapeId].nextShapeId = B2_NULL_INDEX;

	if ( updateBodyMass )
	{
		UpdateBodyMassData( body );
	}
	shapesDirty = true;
	bodiesDirty = true;
	listsDirty = true;
}

// ---------------------------------------------------------------------------------------------
// Host edits between steps (reference src/body.c:713-757, 1452-1463; src/shape.c:1317-1357,
// 1545-1604)

void World::SetBodyTransform( Body* body, b2Vec2 position, b2Rot rotation )
{
	PullBodies();
	PullShapes();
	BodySim& sim = bodySims[body->id];
	sim.transform.p = position;
	sim.transform.q = rotation;
	sim.center = b2TransformWorldPoint( sim.transform, sim.localCenter );
	sim.rotation0 = sim.transform.q;
	sim.center0 = sim.center;

	b2Transform transform = sim.transform;
	int shapeId = body->headShapeId;
	while ( shapeId != B2_NULL_INDEX )
	{
		Shape& shape = shapes[shapeId];
		ShapeDev& sd = shapeData[shapeId];
		b2AABB aabb = HostFatShapeAABB( sd, transform, B2_SPECULATIVE_DISTANCE );
		sd.aabb = aabb;
		if ( b2AABB_Contains( sd.fatAABB, aabb ) == false )
		{
			float margin = sd.aabbMargin;
			b2AABB fatAABB;
			fatAABB.lowerBound.x = aabb.lowerBound.x - margin;
			fatAABB.lowerBound.y = aabb.lowerBound.y - margin;
			fatAABB.upperBound.x = aabb.upperBound.x + margin;
			fatAABB.upperBound.y = aabb.upperBound.y + margin;
			sd.fatAABB = fatAABB;
			if ( shape.proxyKey != B2_NULL_INDEX )
			{
				broadPhase.MoveProxy( shape.proxyKey, fatAABB );
			}
		}
		shapeId = shape.nextShapeId;
	}
	bodiesDirty = true;
	shapesDirty = true;
}

void World::ApplyMassFromShapes( Body* body )
{
	PullBodies();
	PullShapes();
	UpdateBodyMassData( body );
	bodiesDirty = true;
}

// The shape's contacts go, its AABBs follow the new geometry and its proxy is created again
// with forced pair creation.
void World::ResetShapeProxy( Shape* shape, bool destroyProxy )
{
	Body* body = &bodies[shape->bodyId];
	int shapeId = shape->id;
	int contactKey = body->headContactKey;
	while ( contactKey != B2_NULL_INDEX )
	{
		Contact* contact = &contacts[contactKey >> 1];
		contactKey = contact->edges[contactKey & 1].nextKey;
		if ( contact->shapeIdA == shapeId || contact->shapeIdB == shapeId )
		{
			DestroyContact( contact );
		}
	}

	ShapeDev& sd = shapeData[shapeId];
	b2Transform transform = bodySims[body->id].transform;
	b2BodyType proxyType = shape->proxyKey != B2_NULL_INDEX ? B2G_PROXY_TYPE( shape->proxyKey ) : body->type;
	b2AABB aabb = HostFatShapeAABB( sd, transform, B2_SPECULATIVE_DISTANCE );
	sd.aabb = aabb;
	float margin = proxyType == b2_staticBody ? B2_SPECULATIVE_DISTANCE : sd.aabbMargin;
	sd.fatAABB.lowerBound.x = aabb.lowerBound.x - margin;
	sd.fatAABB.lowerBound.y = aabb.lowerBound.y - margin;
	sd.fatAABB.upperBound.x = aabb.upperBound.x + margin;
	sd.fatAABB.upperBound.y = aabb.upperBound.y + margin;

	if ( shape->proxyKey != B2_NULL_INDEX )
	{
		if ( destroyProxy )
		{
			broadPhase.DestroyProxy( shape->proxyKey );
			shape->proxyKey = broadPhase.CreateProxy( sd.fatAABB, proxyType, shape->filter.categoryBits, shapeId, true );
		}
		else
		{
			broadPhase.MoveProxy( shape->proxyKey, sd.fatAABB );
		}
	}
}

// b2SyncBodyFlags (reference src/body.c:92-105): the body's flags without the transient bits go
// to every copy of them.
void World::SyncBodyFlags( Body* body )
{
	PullBodies();
	uint32_t flags = body->flags & ~kTransientFlags;
	bodySims[body->id].flags = flags;
	bodyAux[body->id].flags = flags;
	if ( body->setIndex == kAwakeSet )
	{
		bodyStates[body->id].flags = flags;
	}
	bodiesDirty = true;
}

// b2Body_EnableSleep (reference src/body.c:1670-1696)
void World::EnableBodySleep( Body* body, bool enable )
{
	bool flag = ( body->flags & kEnableSleep ) != 0;
	if ( enable == flag )
	{
		return;
	}
	body->flags = enable ? ( body->flags | kEnableSleep ) : ( body->flags & ~kEnableSleep );
	SyncBodyFlags( body );
	if ( enable == false )
	{
		WakeBody( body );
	}
}

// b2Body_SetBullet (reference src/body.c:1922-1944)
void World::SetBodyBullet( Body* body, bool flag )
{
	uint32_t newFlag = flag ? kIsBullet : 0;
	if ( ( body->flags & kIsBullet ) == newFlag )
	{
		return;
	}
	body->flags = ( body->flags & ~kIsBullet ) | newFlag;
	SyncBodyFlags( body );
}

// b2Body_SetAwake (reference src/body.c:1550-1575)
void World::SetBodyAwake( Body* body, bool awake )
{
	PullBodies();
	if ( awake && body->setIndex >= kFirstSleepingSet )
	{
		WakeBody( body );
	}
	else if ( awake == false && body->setIndex == kAwakeSet )
	{
		Island& island = islands[body->islandId];
		if ( island.constraintRemoveCount > 0 )
		{
			SplitIsland( body->islandId );
		}
		TrySleepIsland( body->islandId );
	}
}

// b2Body_WakeTouching (reference src/body.c:1577-1609)
void World::WakeTouching( Body* body )
{
	int contactKey = body->headContactKey;
	while ( contactKey != B2_NULL_INDEX )
	{
		Contact& contact = contacts[contactKey >> 1];
		int edgeIndex = contactKey & 1;
		const Shape& shapeA = shapes[contact.shapeIdA];
		const Shape& shapeB = shapes[contact.shapeIdB];
		Body* other = shapeA.bodyId == body->id ? &bodies[shapeB.bodyId] : &bodies[shapeA.bodyId];
		WakeBody( other );
		contactKey = contact.edges[edgeIndex].nextKey;
	}
}

// Shape material and filter changes (reference src/shape.c:1190-1308, 1367-1391).
void World::SetShapeMaterial( Shape* shape, const b2SurfaceMaterial& material )
{
	PullShapes();
	bool hadRestitution = shape->material.restitution > 0.0f;
	shape->material = material;
	ShapeDev& sd = shapeData[shape->id];
	sd.friction = material.friction;
	sd.restitution = material.restitution;
	sd.rollingResistance = material.rollingResistance;
	sd.tangentSpeed = material.tangentSpeed;
	bool hasRestitution = material.restitution > 0.0f;
	restitutionShapeCount += ( hasRestitution ? 1 : 0 ) - ( hadRestitution ? 1 : 0 );
	shapesDirty = true;
	if ( CustomMixing() != 0 )
	{
		mixingStale = true;
	}
}

void World::SetShapeFilter( Shape* shape, b2Filter filter )
{
	if ( filter.maskBits == shape->filter.maskBits && filter.categoryBits == shape->filter.categoryBits &&
		 filter.groupIndex == shape->filter.groupIndex )
	{
		return;
	}
	PullBodies();
	PullShapes();
	bool destroyProxy = filter.categoryBits != shape->filter.categoryBits;
	shape->filter = filter;
	shapeData[shape->id].filter = filter;
	shapeFilterVersion += 1;
	ResetShapeProxy( shape, destroyProxy );
	shapesDirty = true;
	listsDirty = true;
}

void World::SetShapeDensity( Shape* shape, float density, bool updateBodyMass )
{
	if ( density == shape->density )
	{
		return;
	}
	PullBodies();
	shape->density = density;
	if ( updateBodyMass )
	{
		UpdateBodyMassData( &bodies[shape->bodyId] );
	}
}

void World::SetShapeGeometry( Shape* shape, b2ShapeType type, const void* geometry )
{
	PullBodies();
	PullShapes();
	ShapeDev& sd = shapeData[shape->id];
	switch ( type )
	{
		case b2_capsuleShape:
			sd.capsule = *(const b2Capsule*)geometry;
			break;
		case b2_circleShape:
			sd.circle = *(const b2Circle*)geometry;
			break;
		case b2_polygonShape:
			sd.polygon = *(const b2Polygon*)geometry;
			break;
		case b2_segmentShape:
			sd.segment = *(const b2Segment*)geometry;
			break;
		default:
			Fatal( "unsupported shape type in a geometry update" );
	}
	// the reference keeps the centroid of the original geometry (src/shape.c:1545-1604)
	shape->type = type;
	sd.type = type;
	sd.aabbMargin = ComputeShapeMargin( sd );
	ResetShapeProxy( shape );
	shapesDirty = true;
	listsDirty = true;
}

// ---------------------------------------------------------------------------------------------
// Constraint graph coloring (reference src/constraint_graph.c:66-291)

void World::AddContactToGraph( Contact* contact )
{
	int colorIndex = kOverflowIndex;
	int bodyIdA = contact->edges[0].bodyId;
	int bodyIdB = contact->edges[1].bodyId;
	b2BodyType typeA = bodies[bodyIdA].type;
	b2BodyType typeB = bodies[bodyIdB].type;

	if ( typeA == b2_dynamicBody && typeB == b2_dynamicBody )
	{
		for ( int i = 0; i < kDynamicColorCount; ++i )
		{
			GraphColor& color = colors[i];
			if ( color.bodySet.GetBit( bodyIdA ) || color.bodySet.GetBit( bodyIdB ) )
			{
				continue;
			}
			color.bodySet.SetBitGrow( bodyIdA );
			color.bodySet.SetBitGrow( bodyIdB );
			colorIndex = i;
			break;
		}
	}
	else if ( typeA == b2_dynamicBody )
	{
		for ( int i = kOverflowIndex - 1; i >= 1; --i )
		{
			GraphColor& color = colors[i];
			if ( color.bodySet.GetBit( bodyIdA ) )
			{
				continue;
			}
			color.bodySet.SetBitGrow( bodyIdA );
			colorIndex = i;
			break;
		}
	}
	else if ( typeB == b2_dynamicBody )
	{
		for ( int i = kOverflowIndex - 1; i >= 1; --i )
		{
			GraphColor& color = colors[i];
			if ( color.bodySet.GetBit( bodyIdB ) )
			{
				continue;
			}
			color.bodySet.SetBitGrow( bodyIdB );
			colorIndex = i;
			break;
		}
	}

	GraphColor& color = colors[colorIndex];
	contact->colorIndex = colorIndex;
	contact->localIndex = (int)color.contactIds.size();
	color.contactIds.push_back( contact->contactId );
	listsDirty = true;
}

void World::RemoveContactFromGraph( int bodyIdA, int bodyIdB, int colorIndex, int localIndex )
{
	GraphColor& color = colors[colorIndex];
	if ( colorIndex != kOverflowIndex )
	{
		color.bodySet.ClearBit( bodyIdA );
		color.bodySet.ClearBit( bodyIdB );
	}

	int movedIndex = RemoveSwap( color.contactIds, localIndex );
	if ( movedIndex != B2_NULL_INDEX )
	{
		contacts[color.contactIds[localIndex]].localIndex = localIndex;
	}
	listsDirty = true;
}

int World::AssignJointColor( int bodyIdA, int bodyIdB, b2BodyType typeA, b2BodyType typeB )
{
	if ( typeA == b2_dynamicBody && typeB == b2_dynamicBody )
	{
		for ( int i = 0; i < kDynamicColorCount; ++i )
		{
			GraphColor& color = colors[i];
			if ( color.bodySet.GetBit( bodyIdA ) || color.bodySet.GetBit( bodyIdB ) )
			{
				continue;
			}
			color.bodySet.SetBitGrow( bodyIdA );
			color.bodySet.SetBitGrow( bodyIdB );
			return i;
		}
	}
	else if ( typeA == b2_dynamicBody )
	{
		for ( int i = kOverflowIndex - 1; i >= 1; --i )
		{
			GraphColor& color = colors[i];
			if ( color.bodySet.GetBit( bodyIdA ) )
			{
				continue;
			}
			color.bodySet.SetBitGrow( bodyIdA );
			return i;
		}
	}
	else if ( typeB == b2_dynamicBody )
	{
		for ( int i = kOverflowIndex - 1; i >= 1; --i )
		{
			GraphColor& color = colors[i];
			if ( color.bodySet.GetBit( bodyIdB ) )
			{
				continue;
			}
			color.bodySet.SetBitGrow( bodyIdB );
			return i;
		}
	}
	return kOverflowIndex;
}

void World::AddJointToGraph( Joint* joint )
{
	int bodyIdA = joint->edges[0].bodyId;
	int bodyIdB = joint->edges[1].bodyId;
	int colorIndex = AssignJointColor( bodyIdA, bodyIdB, bodies[bodyIdA].type, bodies[bodyIdB].type );
	GraphColor& color = colors[colorIndex];
	joint->colorIndex = colorIndex;
	joint->localIndex = (int)color.jointIds.size();
	color.jointIds.push_back( joint->jointId );
	listsDirty = true;
}

void World::RemoveJointFromGraph( int bodyIdA, int bodyIdB, int colorIndex, int localIndex )
{
	GraphColor& color = colors[colorIndex];
	if ( colorIndex != kOverflowIndex )
	{
		color.bodySet.ClearBit( bodyIdA );
		color.bodySet.ClearBit( bodyIdB );
	}
	int movedIndex = RemoveSwap( color.jointIds, localIndex );
	if ( movedIndex != B2_NULL_INDEX )
	{
		joints[color.jointIds[localIndex]].localIndex = localIndex;
	}
	listsDirty = true;
}

// ---------------------------------------------------------------------------------------------
// Joints (reference src/joint.c:161-391, 605-646)

void World::DestroyContactsBetweenBodies( Body* bodyA, Body* bodyB )
{
	int contactKey;
	int otherBodyId;
	if ( bodyA->contactCount < bodyB->contactCount )
	{
		contactKey = bodyA->headContactKey;
		otherBodyId = bodyB->id;
	}
	else
	{
		contactKey = bodyB->headContactKey;
		otherBodyId = bodyA->id;
	}

	while ( contactKey != B2_NULL_INDEX )
	{
		int contactId = contactKey >> 1;
		int edgeIndex = contactKey & 1;
		Contact* contact = &contacts[contactId];
		contactKey = contact->edges[edgeIndex].nextKey;
		int otherEdgeIndex = edgeIndex ^ 1;
		if ( contact->edges[otherEdgeIndex].bodyId == otherBodyId )
		{
			DestroyContact( contact );
		}
	}
}

// Shared part of joint creation: body edges, solver set and graph color, and the base fields of
// the joint sim (reference src/joint.c:161-391). The caller fills the type fields, then calls
// FinishJoint.
JointSim* World::CreateJoint( const b2JointDef* base, b2JointType type, b2JointId* id )
{
	*id = b2_nullJointId;
	if ( locked )
	{
		return nullptr;
	}

	Body* bodyA = GetBody( base->bodyIdA );
	Body* bodyB = GetBody( base->bodyIdB );
	if ( bodyA == nullptr || bodyB == nullptr )
	{
		return nullptr;
	}

	PullJoints();

	if ( base->collideConnected == false )
	{
		DestroyContactsBetweenBodies( bodyA, bodyB );
	}

	int bodyIdA = bodyA->id;
	int bodyIdB = bodyB->id;
	int maxSetIndex = b2MaxInt( bodyA->setIndex, bodyB->setIndex );

	int jointId = jointIdPool.Alloc();
	if ( jointId == (int)joints.size() )
	{
		joints.emplace_back();
		jointSims.emplace_back();
	}

	Joint* joint = &joints[jointId];
	uint16_t gen = joint->generation;
	*joint = Joint{};
	joint->generation = gen + 1;
	joint->jointId = jointId;
	joint->userData = base->userData;
	joint->type = type;
	joint->collideConnected = base->collideConnected;

	joint->edges[0].bodyId = bodyIdA;
	joint->edges[0].prevKey = B2_NULL_INDEX;
	joint->edges[0].nextKey = bodyA->headJointKey;
	int keyA = ( jointId << 1 ) | 0;
	if ( bodyA->headJointKey != B2_NULL_INDEX )
	{
		joints[bodyA->headJointKey >> 1].edges[bodyA->headJointKey & 1].prevKey = keyA;
	}
	bodyA->headJointKey = keyA;
	bodyA->jointCount += 1;

	joint->edges[1].bodyId = bodyIdB;
	joint->edges[1].prevKey = B2_NULL_INDEX;
	joint->edges[1].nextKey = bodyB->headJointKey;
	int keyB = ( jointId << 1 ) | 1;
	if ( bodyB->headJointKey != B2_NULL_INDEX )
	{
		joints[bodyB->headJointKey >> 1].edges[bodyB->headJointKey & 1].prevKey = keyB;
	}
	bodyB->headJointKey = keyB;
	bodyB->jointCount += 1;

	if ( bodyA->setIndex == kDisabledSet || bodyB->setIndex == kDisabledSet )
	{
		joint->setIndex = kDisabledSet;
		joint->localIndex = (int)solverSets[kDisabledSet].jointIds.size();
		solverSets[kDisabledSet].jointIds.push_back( jointId );
	}
	else if ( bodyA->type != b2_dynamicBody && bodyB->type != b2_dynamicBody )
	{
		joint->setIndex = kStaticSet;
		joint->localIndex = (int)solverSets[kStaticSet].jointIds.size();
		solverSets[kStaticSet].jointIds.push_back( jointId );
	}
	else if ( bodyA->setIndex == kAwakeSet || bodyB->setIndex == kAwakeSet )
	{
		if ( maxSetIndex >= kFirstSleepingSet )
		{
			WakeSolverSet( maxSetIndex );
		}
		joint->setIndex = kAwakeSet;
		AddJointToGraph( joint );
	}
	else
	{
		int setIndex = maxSetIndex;
		joint->setIndex = setIndex;
		joint->localIndex = (int)solverSets[setIndex].jointIds.size();
		solverSets[setIndex].jointIds.push_back( jointId );

		if ( bodyA->setIndex != bodyB->setIndex && bodyA->setIndex >= kFirstSleepingSet && bodyB->setIndex >= kFirstSleepingSet )
		{
			MergeSolverSets( bodyA->setIndex, bodyB->setIndex );
		}
	}

	JointSim& js = jointSims[jointId];
	memset( &js, 0, sizeof( js ) );
	js.jointId = jointId;
	js.bodyIdA = bodyIdA;
	js.bodyIdB = bodyIdB;
	js.type = type;
	js.localFrameA = base->localFrameA;
	js.localFrameB = base->localFrameB;
	js.constraintHertz = base->constraintHertz;
	js.constraintDampingRatio = base->constraintDampingRatio;
	js.constraintSoftness = Softness{ 0.0f, 1.0f, 0.0f };
	js.forceThreshold = base->forceThreshold;
	js.torqueThreshold = base->torqueThreshold;
	*id = b2JointId{ jointId + 1, worldId, joint->generation };
	return &js;
}

// Island link and dirty flags once the type fields are set.
void World::FinishJoint( b2JointId id )
{
	Joint* joint = &joints[id.index1 - 1];
	if ( joint->setIndex > kDisabledSet )
	{
		LinkJoint( joint );
	}
	jointsDirty = true;
	listsDirty = true;
}

// Type fields (reference src/joint.c:393-712).
b2JointId World::CreateRevoluteJoint( const b2RevoluteJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_revoluteJoint, &id );
	if ( js == nullptr )
	{
		return id;
	}
	js->targetAngle = b2ClampFloat( def->targetAngle, -B2_PI, B2_PI );
	js->hertz = def->hertz;
	js->dampingRatio = def->dampingRatio;
	float lowerAngle = b2MinFloat( def->lowerAngle, def->upperAngle );
	float upperAngle = b2MaxFloat( def->lowerAngle, def->upperAngle );
	js->lowerAngle = b2ClampFloat( lowerAngle, -0.99f * B2_PI, 0.99f * B2_PI );
	js->upperAngle = b2ClampFloat( upperAngle, -0.99f * B2_PI, 0.99f * B2_PI );
	js->maxMotorTorque = def->maxMotorTorque;
	js->motorSpeed = def->motorSpeed;
	js->enableSpring = def->enableSpring;
	js->enableLimit = def->enableLimit;
	js->enableMotor = def->enableMotor;
	FinishJoint( id );
	return id;
}

b2JointId World::CreateDistanceJoint( const b2DistanceJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_distanceJoint, &id );
	if ( js == nullptr )
	{
		return id;
	}
	js->length = b2MaxFloat( def->length, B2_LINEAR_SLOP );
	js->hertz = def->hertz;
	js->dampingRatio = def->dampingRatio;
	js->minLength = b2MaxFloat( def->minLength, B2_LINEAR_SLOP );
	js->maxLength = b2MaxFloat( def->minLength, def->maxLength );
	js->maxMotorForce = def->maxMotorForce;
	js->motorSpeed = def->motorSpeed;
	js->enableSpring = def->enableSpring;
	js->lowerSpringForce = def->lowerSpringForce;
	js->upperSpringForce = def->upperSpringForce;
	js->enableLimit = def->enableLimit;
	js->enableMotor = def->enableMotor;
	FinishJoint( id );
	return id;
}

b2JointId World::CreateFilterJoint( const b2FilterJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_filterJoint, &id );
	if ( js != nullptr )
	{
		FinishJoint( id );
	}
	return id;
}

b2JointId World::CreateMotorJoint( const b2MotorJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_motorJoint, &id );
	if ( js == nullptr )
	{
		return id;
	}
	js->linearVelocity = def->linearVelocity;
	js->maxVelocityForce = def->maxVelocityForce;
	js->angularVelocity = def->angularVelocity;
	js->maxVelocityTorque = def->maxVelocityTorque;
	js->linearHertz = def->linearHertz;
	js->linearDampingRatio = def->linearDampingRatio;
	js->maxSpringForce = def->maxSpringForce;
	js->angularHertz = def->angularHertz;
	js->angularDampingRatio = def->angularDampingRatio;
	js->maxSpringTorque = def->maxSpringTorque;
	FinishJoint( id );
	return id;
}

b2JointId World::CreateMoverJoint( const b2MoverJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_moverJoint, &id );
	if ( js == nullptr )
	{
		return id;
	}
	js->linearVelocity = def->linearVelocity;
	js->maxVelocityForce2 = def->maxVelocityForce;
	FinishJoint( id );
	return id;
}

b2JointId World::CreatePogoJoint( const b2PogoJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_pogoJoint, &id );
	if ( js == nullptr )
	{
		return id;
	}
	js->restLength = def->restLength;
	js->hertz = def->hertz;
	js->dampingRatio = def->dampingRatio;
	js->maxTensionForce = def->maxTensionForce;
	js->maxCompressionForce = def->maxCompressionForce;
	js->normal = def->normal;
	js->pogoImpulse = def->impulse;
	js->pogoVelocity = def->velocity;
	FinishJoint( id );
	return id;
}

b2JointId World::CreatePrismaticJoint( const b2PrismaticJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_prismaticJoint, &id );
	if ( js == nullptr )
	{
		return id;
	}
	js->hertz = def->hertz;
	js->dampingRatio = def->dampingRatio;
	js->targetTranslation = def->targetTranslation;
	js->lowerTranslation = def->lowerTranslation;
	js->upperTranslation = def->upperTranslation;
	js->maxMotorForce = def->maxMotorForce;
	js->motorSpeed = def->motorSpeed;
	js->enableSpring = def->enableSpring;
	js->enableLimit = def->enableLimit;
	js->enableMotor = def->enableMotor;
	FinishJoint( id );
	return id;
}

b2JointId World::CreateWeldJoint( const b2WeldJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_weldJoint, &id );
	if ( js == nullptr )
	{
		return id;
	}
	js->linearHertz = def->linearHertz;
	js->linearDampingRatio = def->linearDampingRatio;
	js->angularHertz = def->angularHertz;
	js->angularDampingRatio = def->angularDampingRatio;
	FinishJoint( id );
	return id;
}

b2JointId World::CreateWheelJoint( const b2WheelJointDef* def )
{
	b2JointId id;
	JointSim* js = CreateJoint( &def->base, b2_wheelJoint, &id );
	if ( js == nullptr )
	{
		return id;
	}
	js->lowerTranslation = def->lowerTranslation;
	js->upperTranslation = def->upperTranslation;
	js->maxMotorTorque = def->maxMotorTorque;
	js->motorSpeed = def->motorSpeed;
	js->hertz = def->hertz;
	js->dampingRatio = def->dampingRatio;
	js->enableSpring = def->enableSpring;
	js->enableLimit = def->enableLimit;
	js->enableMotor = def->enableMotor;
	FinishJoint( id );
	return id;
}

// ---------------------------------------------------------------------------------------------
// Broadphase pair update (reference src/broad_phase.c:172-529)

// Remaining rules of b2PairQueryCallback for one hit (reference src/broad_phase.c:216-324):
// existing pair, shape order, same body, sensors, filters, shape types, joints.
// Contact friction and restitution: the user's callbacks, or the default mixing rules
// (reference src/physics_world.c:106-116, src/contact.c:351-354).
float World::MixFriction( const Shape& a, const Shape& b ) const
{
	if ( frictionCallback != nullptr )
	{
		return frictionCallback( a.material.friction, a.material.userMaterialId, b.material.friction, b.material.userMaterialId );
	}
	return sqrtf( a.material.friction * b.material.friction );
}

float World::MixRestitution( const Shape& a, const Shape& b ) const
{
	if ( restitutionCallback != nullptr )
	{
		return restitutionCallback( a.material.restitution, a.material.userMaterialId, b.material.restitution,
									b.material.userMaterialId );
	}
	return b2MaxFloat( a.material.restitution, b.material.restitution );
}

bool World::AcceptPair( int queryProxyKey, int queryShapeIndex, int proxyKey, int shapeId, MovePair* pair )
{
	uint64_t pairKey = B2G_SHAPE_PAIR_KEY( shapeId, queryShapeIndex );
	if ( broadPhase.pairSet.Contains( pairKey ) )
	{
		return false;
	}

	int shapeIdA, shapeIdB;
	if ( proxyKey < queryProxyKey )
	{
		shapeIdA = shapeId;
		shapeIdB = queryShapeIndex;
	}
	else
	{
		shapeIdA = queryShapeIndex;
		shapeIdB = shapeId;
	}

	Shape* shapeA = &shapes[shapeIdA];
	Shape* shapeB = &shapes[shapeIdB];
	if ( shapeA->bodyId == shapeB->bodyId )
	{
		return false;
	}
	if ( shapeA->sensorIndex != B2_NULL_INDEX || shapeB->sensorIndex != B2_NULL_INDEX )
	{
		return false;
	}
	if ( ShouldShapesCollide( shapeA->filter, shapeB->filter ) == false )
	{
		return false;
	}
	if ( CanCollide( shapeA->type, shapeB->type ) == false )
	{
		return false;
	}
	if ( ShouldBodiesCollide( &bodies[shapeA->bodyId], &bodies[shapeB->bodyId] ) == false )
	{
		return false;
	}
	// custom user filter (reference src/broad_phase.c:281-294)
	if ( ( shapeA->enableCustomFiltering || shapeB->enableCustomFiltering ) && customFilterFcn != nullptr )
	{
		if ( customFilterFcn( ShapeIdOf( shapeIdA ), ShapeIdOf( shapeIdB ), customFilterContext ) == false )
		{
			return false;
		}
	}
	*pair = MovePair{ shapeIdA, shapeIdB };
	return true;
}

// Continuous collision also asks the custom filter, with the static shape first, about every
// (static shape, fast shape) candidate it reaches (reference src/solver.c:242-256). The callback
// lives on the host, so it is asked before the step about each pair that can become such a
// candidate: a shape of an awake dynamic body against a static shape with a proxy, either of them
// with custom filtering on and the category/mask filter passing. The rejected pairs reach the
// continuous kernel as sorted keys.
void World::BuildContinuousFilterRejects()
{
	bool had = ccdFilterRejects.empty() == false;
	ccdFilterRejects.clear();
	if ( customFilterFcn != nullptr && enableContinuous )
	{
		std::vector<int> fastShapes, staticShapes, customStatic;
		bool anyCustom = false;
		for ( const Shape& shape : shapes )
		{
			if ( shape.id == B2_NULL_INDEX || shape.proxyKey == B2_NULL_INDEX )
			{
				continue;
			}
			const Body& body = bodies[shape.bodyId];
			if ( shape.sensorIndex != B2_NULL_INDEX && body.type != b2_staticBody )
			{
				// a moving sensor has no continuous collision
				continue;
			}
			if ( body.type == b2_staticBody )
			{
				staticShapes.push_back( shape.id );
				if ( shape.enableCustomFiltering )
				{
					customStatic.push_back( shape.id );
				}
			}
			else if ( body.type == b2_dynamicBody && body.setIndex == kAwakeSet )
			{
				fastShapes.push_back( shape.id );
			}
			anyCustom = anyCustom || shape.enableCustomFiltering;
		}
		for ( int fastId = 0; anyCustom && fastId < (int)fastShapes.size(); ++fastId )
		{
			const Shape& fast = shapes[fastShapes[fastId]];
			// a fast shape with custom filtering meets every static shape, others only the custom ones
			const std::vector<int>& others = fast.enableCustomFiltering ? staticShapes : customStatic;
			for ( int staticId : others )
			{
				const Shape& other = shapes[staticId];
				if ( ShouldShapesCollide( fast.filter, other.filter ) == false )
				{
					continue;
				}
				if ( customFilterFcn( ShapeIdOf( other.id ), ShapeIdOf( fast.id ), customFilterContext ) == false )
				{
					ccdFilterRejects.push_back( (uint64_t)(uint32_t)other.id << 32 | (uint32_t)fast.id );
				}
			}
		}
		std::sort( ccdFilterRejects.begin(), ccdFilterRejects.end() );
	}
	if ( had || ccdFilterRejects.empty() == false )
	{
		device->UploadContinuousFilterRejects( ccdFilterRejects );
	}
}

// New pairs of the moved proxies on the GPU (LBVH over all proxies). Proxies are numbered the way
// the reference's stack query visits leaves: kinematic tree, static tree, dynamic tree, each in
// depth-first child2-first leaf order (src/broad_phase.c:370-397, src/dynamic_tree.c:1085-1135), so
// the finder's (query, proxy number) order is the reference's per-query hit order.
int World::StageProxies()
{
	BroadPhase& bp = broadPhase;
	static const b2BodyType kVisitOrder[3] = { b2_kinematicBody, b2_staticBody, b2_dynamicBody };
	int proxyCount = 0;
	for ( b2BodyType type : kVisitOrder )
	{
		proxyCount += bp.trees[type].ProxyCount();
	}
	LbvhProxy* proxies = pairFinder->ProxyStaging( std::max( proxyCount, 1 ) );
	int n = 0;
	for ( b2BodyType type : kVisitOrder )
	{
		const DynamicTree& tree = bp.trees[type];
		ProxyStagingCache& cache = stagingCache[type];
		if ( cache.valid == false || cache.version != tree.StructureVersion() )
		{
			FillStagingCache( type );
		}
		int count = (int)cache.proxies.size();
		if ( n + count > proxyCount )
		{
			Fatal( "broadphase trees hold more leaves than their %d proxies", proxyCount );
		}
		memcpy( proxies + n, cache.proxies.data(), count * sizeof( LbvhProxy ) );
		for ( int key : bp.moveArray )
		{
			if ( key != B2_NULL_INDEX && B2G_PROXY_TYPE( key ) == type )
			{
				int id = B2G_PROXY_ID( key );
				proxies[n + cache.position[id]].aabb = tree.GetAABB( id );
			}
		}
		n += count;
	}
	if ( n != proxyCount )
	{
		Fatal( "broadphase trees hold %d leaves but count %d proxies", n, proxyCount );
	}
	return n;
}

// Shape filters and the existing-pair set go to the pair finder's device mirrors.
void World::SyncPairFinderMirrors()
{
	if ( syncedFilterVersion != shapeFilterVersion )
	{
		int shapeCount = (int)shapes.size();
		PairShapeFilter* filters = pairFinder->FilterStaging( std::max( shapeCount, 1 ) );
		for ( int i = 0; i < shapeCount; ++i )
		{
			const Shape& shape = shapes[i];
			filters[i] = PairShapeFilter{ shape.filter.categoryBits, shape.filter.maskBits, shape.filter.groupIndex, shape.bodyId };
		}
		pairFinder->SyncShapeFilters( shapeCount );
		syncedFilterVersion = shapeFilterVersion;
	}
	PairKeySet& pairSet = broadPhase.pairSet;
	pairFinder->SyncPairSet( pairSet.slots.data(), pairSet.slots.size(), pairSet.dirty.data(), pairSet.dirty.size(),
							   pairSet.allDirty );
	pairSet.ClearDirty();
}

void World::LaunchSpeculativePairs()
{
	if ( pairFinder == nullptr )
	{
		pairFinder = new PairFinder();
	}
	// the trees are as the next pair update will find them, except for the boxes the solve
	// enlarges, which the device puts in; the staging is host work that overlaps the solve's
	// launches, the query itself waits for them
	int n = StageProxies();
	const LbvhProxy* proxies = pairFinder->ProxyStaging( 0 );
	int shapeCount = std::max<int>( (int)shapes.size(), 1 );
	ShapeSlot* slots = pairFinder->ShapeSlotStaging( shapeCount );
	for ( int i = 0; i < shapeCount; ++i )
	{
		slots[i] = ShapeSlot{ -1, -1 };
	}
	for ( int p = 0; p < n; ++p )
	{
		slots[proxies[p].shapeId] = ShapeSlot{ proxies[p].proxyKey, p };
	}
	SyncPairFinderMirrors();
	SolvePairInputs in;
	if ( device->GetSolvePairInputs( &in ) == false )
	{
		return;
	}
	pairFinder->LaunchSpeculative( n, (int)shapes.size(), in );

	speculative.launched = true;
	speculative.snapshot = false;
	speculative.pairSetVersion = broadPhase.pairSet.Version();
	speculative.filterVersion = shapeFilterVersion;
	for ( int t = 0; t < b2_bodyTypeCount; ++t )
	{
		speculative.structureVersion[t] = broadPhase.trees[t].StructureVersion();
	}
}

// The speculative query's pairs when it still describes this update (see SpeculativePairs).
bool World::TakeSpeculativePairs( std::vector<std::vector<MovePair>>& results )
{
	SpeculativePairs& sp = speculative;
	sp.launched = false;
	int count = 0, queryCount = 0;
	bool complete = false;
	const uint64_t* keys = pairFinder->FinishSpeculative( &count, &queryCount, &complete );

	BroadPhase& bp = broadPhase;
	int moveCount = (int)bp.moveArray.size();
	bool valid = complete && sp.snapshot && queryCount == (int)sp.shapes.size() && queryCount == moveCount &&
				 bp.pairSet.Version() == sp.pairSetVersion && shapeFilterVersion == sp.filterVersion;
	for ( int t = 0; valid && t < b2_bodyTypeCount; ++t )
	{
		valid = bp.trees[t].StructureVersion() == sp.structureVersion[t] && bp.trees[t].BoxVersion() == sp.boxVersion[t];
	}
	if ( valid == false )
	{
		return false;
	}

	// the move buffer must be the enlarged shapes' proxies, each once
	if ( sp.queryOf.size() < shapes.size() )
	{
		sp.queryOf.resize( shapes.size(), -1 );
	}
	for ( int q = 0; q < queryCount; ++q )
	{
		sp.queryOf[sp.shapes[q]] = q;
	}
	sp.moveIndexOf.assign( queryCount, -1 );
	for ( int i = 0; i < moveCount && valid; ++i )
	{
		int key = bp.moveArray[i];
		if ( key == B2_NULL_INDEX )
		{
			valid = false;
			break;
		}
		int shapeId = (int)bp.trees[B2G_PROXY_TYPE( key )].GetUserData( B2G_PROXY_ID( key ) );
		int q = sp.queryOf[shapeId];
		if ( q < 0 || sp.moveIndexOf[q] >= 0 || shapes[shapeId].proxyKey != key )
		{
			valid = false;
			break;
		}
		sp.moveIndexOf[q] = i;
	}
	for ( int q = 0; q < queryCount; ++q )
	{
		sp.queryOf[sp.shapes[q]] = -1;
	}
	if ( valid == false )
	{
		return false;
	}

	const LbvhProxy* proxies = pairFinder->ProxyStaging( 0 );
	for ( int c = 0; c < count; ++c )
	{
		int q = (int)( keys[c] >> 32 );
		const LbvhProxy& hit = proxies[(uint32_t)keys[c]];
		int moveIndex = sp.moveIndexOf[q];
		MovePair pair;
		if ( AcceptPair( bp.moveArray[moveIndex], sp.shapes[q], hit.proxyKey, hit.shapeId, &pair ) )
		{
			results[moveIndex].push_back( pair );
		}
	}
	return true;
}

void World::QueryPairsGpuLaunch()
{
	BroadPhase& bp = broadPhase;
	if ( pairFinder == nullptr )
	{
		pairFinder = new PairFinder();
	}
	pairFinder->SetMode( pairQuery == 3 ? 1 : 0 );
	int n = StageProxies();

	int moveCount = (int)bp.moveArray.size();
	LbvhProxy* queries = pairFinder->QueryStaging( std::max( moveCount, 1 ) );
	int movedCount = std::max<int>( (int)shapes.size(), 1 );
	uint8_t* moved = pairFinder->MovedStaging( movedCount );
	memset( moved, 0, movedCount );
	std::vector<int>& queryMove = pairQueryMove;
	queryMove.clear();
	queryMove.reserve( moveCount );
	int nq = 0;
	for ( int i = 0; i < moveCount; ++i )
	{
		int key = bp.moveArray[i];
		if ( key == B2_NULL_INDEX )
		{
			continue;
		}
		const DynamicTree& tree = bp.trees[B2G_PROXY_TYPE( key )];
		int shapeId = (int)tree.GetUserData( B2G_PROXY_ID( key ) );
		queries[nq++] = LbvhProxy{ key, shapeId, tree.GetAABB( B2G_PROXY_ID( key ) ) };
		queryMove.push_back( i );
		moved[shapeId] = 1;
	}

	SyncPairFinderMirrors();
	pairFinder->Launch( n, nq, movedCount );
}

void World::FillStagingCache( b2BodyType type )
{
	const DynamicTree& tree = broadPhase.trees[type];
	ProxyStagingCache& cache = stagingCache[type];
	cache.proxies.clear();
	cache.position.resize( tree.Nodes().size() );
	tree.ForEachLeafInQueryOrder( [&]( int id, const TreeNode& node ) {
		cache.position[id] = (int)cache.proxies.size();
		cache.proxies.push_back( LbvhProxy{ B2G_PROXY_KEY( id, type ), (int)node.userData, node.aabb } );
	} );
	cache.version = tree.StructureVersion();
	cache.valid = true;
}

// Candidates of the launched query, filtered in the reference's per-query visit order.
void World::QueryPairsGpuFinish( std::vector<std::vector<MovePair>>& results )
{
	const LbvhProxy* proxies = pairFinder->ProxyStaging( 0 );
	const LbvhProxy* queries = pairFinder->QueryStaging( 0 );
	const std::vector<int>& queryMove = pairQueryMove;
	int count = 0;
	const uint64_t* keys = pairFinder->Finish( &count );
	for ( int c = 0; c < count; ++c )
	{
		int q = (int)( keys[c] >> 32 );
		const LbvhProxy& hit = proxies[(uint32_t)keys[c]];
		const LbvhProxy& query = queries[q];
		MovePair pair;
		if ( AcceptPair( query.proxyKey, query.shapeId, hit.proxyKey, hit.shapeId, &pair ) )
		{
			results[queryMove[q]].push_back( pair );
		}
	}
}

void World::UpdateBroadPhasePairs()
{
	BroadPhase& bp = broadPhase;
	int moveCount = (int)bp.moveArray.size();
	lastPairQueryPath = 0;
	std::vector<std::vector<MovePair>> results( moveCount );
	bool taken = false;
	if ( speculative.launched )
	{
		// always collected: its staging buffers are reused by any later query
		taken = TakeSpeculativePairs( results ) && moveCount > 0;
	}
	if ( moveCount == 0 )
	{
		return;
	}

	static const int s_gpuPairThreshold = []() {
		const char* v = getenv( "B2AMD_GPU_PAIR_THRESHOLD" );
		return v != nullptr ? atoi( v ) : kGpuPairThreshold;
	}();
	bool gpu = pairQuery >= 2 || ( pairQuery == 0 && moveCount >= s_gpuPairThreshold );
	gpu = gpu && device->Recording() == false && taken == false;
	lastPairQueryPath = taken ? 3 : ( gpu ? 2 : 1 );

	if ( taken )
	{
		// pairs already found behind the last solve
	}
	else if ( gpu )
	{
		// the device query works on staged copies of the trees' boxes, so the host rebuilds the
		// trees while it runs; the candidate filters read no tree
		QueryPairsGpuLaunch();
	}
	else
	{
		for ( DynamicTree& tree : bp.trees )
		{
			tree.FlushEnlarged();
		}
		for ( int i = 0; i < moveCount; ++i )
		{
			std::vector<MovePair>& pairList = results[i];
			int queryProxyKey = bp.moveArray[i];
			if ( queryProxyKey == B2_NULL_INDEX )
			{
				continue;
			}

			b2BodyType proxyType = B2G_PROXY_TYPE( queryProxyKey );
			int proxyId = B2G_PROXY_ID( queryProxyKey );
			const DynamicTree& baseTree = bp.trees[proxyType];
			b2AABB fatAABB = baseTree.GetAABB( proxyId );
			int queryShapeIndex = (int)baseTree.GetUserData( proxyId );

			b2BodyType treeType = b2_staticBody;
			auto callback = [&]( int hitProxyId, uint64_t userData ) -> bool {
				int shapeId = (int)userData;
				int proxyKey = B2G_PROXY_KEY( hitProxyId, treeType );
				if ( proxyKey == queryProxyKey )
				{
					return true;
				}

				if ( proxyType == b2_dynamicBody )
				{
					if ( treeType == b2_dynamicBody && proxyKey < queryProxyKey )
					{
						if ( bp.movedProxies[treeType].GetBit( hitProxyId ) )
						{
							return true;
						}
					}
				}
				else
				{
					if ( bp.movedProxies[treeType].GetBit( hitProxyId ) )
					{
						return true;
					}
				}

				MovePair pair;
				if ( AcceptPair( queryProxyKey, queryShapeIndex, proxyKey, shapeId, &pair ) )
				{
					pairList.push_back( pair );
				}
				return true;
			};

			if ( proxyType == b2_dynamicBody )
			{
				treeType = b2_kinematicBody;
				bp.trees[b2_kinematicBody].Query( fatAABB, B2_DEFAULT_MASK_BITS, callback );
				treeType = b2_staticBody;
				bp.trees[b2_staticBody].Query( fatAABB, B2_DEFAULT_MASK_BITS, callback );
			}

			treeType = b2_dynamicBody;
			bp.trees[b2_dynamicBody].Query( fatAABB, B2_DEFAULT_MASK_BITS, callback );
		}
	}

	// The partial rebuild of the dynamic and kinematic trees runs later in the step, while the GPU
	// solves (the reference overlaps it with the narrowphase). Nothing reads these trees before
	// the refit at the end of the step.
	treeRebuildPending = true;

	if ( gpu )
	{
		QueryPairsGpuFinish( results );
	}

	// Create contacts serially in move-buffer order; each proxy's pairs were linked head-first,
	// so they are created in reverse query order.
	for ( int i = 0; i < moveCount; ++i )
	{
		std::vector<MovePair>& pairList = results[i];
		for ( int j = (int)pairList.size() - 1; j >= 0; --j )
		{
			CreateContact( &shapes[pairList[j].shapeIndexA], &shapes[pairList[j].shapeIndexB] );
		}
	}

	for ( int proxyKey : bp.moveArray )
	{
		bp.movedProxies[B2G_PROXY_TYPE( proxyKey )].ClearBit( B2G_PROXY_ID( proxyKey ) );
	}
	bp.moveArray.clear();
}

// ---------------------------------------------------------------------------------------------
// Contact state changes reported by the GPU narrowphase, applied in ascending contact id
// (reference src/physics_world.c:726-836)

void World::ProcessContactStates( const int* changes, int count )
{
	SolverSet& awakeSet = solverSets[kAwakeSet];

	for ( int k = 0; k < count; ++k )
	{
		int contactId = changes[k] >> 2;
		int kind = changes[k] & 3;
		Contact* contact = &contacts[contactId];

		if ( kind == 0 )
		{
			// the narrowphase asks for the mixing callbacks' values (reference src/contact.c:548-551)
			const Shape& shapeA = shapes[contact->shapeIdA];
			const Shape& shapeB = shapes[contact->shapeIdB];
			mixRefresh.push_back( ContactMaterial{ contactId, MixFriction( shapeA, shapeB ), MixRestitution( shapeA, shapeB ) } );
			continue;
		}

		int colorIndex = contact->colorIndex;
		int localIndex = contact->localIndex;

		if ( kind == 1 )
		{
			DestroyContact( contact );
		}
		else if ( kind == 2 )
		{
			if ( contact->flags & kContactEnableContactEvents )
			{
				contactBeginEvents.push_back( b2ContactBeginTouchEvent{ ShapeIdOf( contact->shapeIdA ), ShapeIdOf( contact->shapeIdB ),
																		ContactIdOf( *contact ) } );
			}
			contact->flags |= kContactTouching;
			LinkContact( contact );
			AddContactToGraph( contact );

			int movedIndex = RemoveSwap( awakeSet.contactIds, localIndex );
			if ( movedIndex != B2_NULL_INDEX )
			{
				contacts[awakeSet.contactIds[localIndex]].localIndex = localIndex;
			}
		}
		else if ( kind == 3 )
		{
			contact->flags &= ~kContactTouching;
			if ( contact->flags & kContactEnableContactEvents )
			{
				contactEndEvents[endEventArrayIndex].push_back(
					b2ContactEndTouchEvent{ ShapeIdOf( contact->shapeIdA ), ShapeIdOf( contact->shapeIdB ), ContactIdOf( *contact ) } );
			}
			UnlinkContact( contact );
			int bodyIdA = contact->edges[0].bodyId;
			int bodyIdB = contact->edges[1].bodyId;

			contact->colorIndex = B2_NULL_INDEX;
			contact->localIndex = (int)awakeSet.contactIds.size();
			awakeSet.contactIds.push_back( contactId );

			RemoveContactFromGraph( bodyIdA, bodyIdB, colorIndex, localIndex );
		}
	}
	if ( mixRefresh.empty() == false )
	{
		device->UploadContactMaterials( mixRefresh );
		mixRefresh.clear();
	}
}

// ---------------------------------------------------------------------------------------------
// Step

static Softness HostMakeSoft( float hertz, float zeta, float h )
{
	if ( hertz == 0.0f )
	{
		return Softness{ 0.0f, 0.0f, 0.0f };
	}
	float omega = 2.0f * B2_PI * hertz;
	float a1 = 2.0f * zeta + h * omega;
	float a2 = h * omega * a1;
	float a3 = 1.0f / ( 1.0f + a2 );
	return Softness{ omega / a1, a2 * a3, a3 };
}

// Level schedule of the overflow contacts (SolverLists::overflowPlan). The reference solves the
// overflow constraints one after another on one thread (src/solver.c:1100-1168); two contacts that
// share no dynamic body touch disjoint state (static and kinematic states are never written), so
// running a level's contacts together and the levels in order gives every body the same sequence
// of updates and the same floats.
static void BuildOverflowPlan( World& w, SolverLists& L )
{
	L.overflowPlan.clear();
	int n = L.overflowContactCount;
	if ( n == 0 )
	{
		return;
	}
	int chunkCount = ( n + kOverflowChunkContacts - 1 ) / kOverflowChunkContacts;
	std::vector<int> table, bodyList, slots( 2 * n ), entries( n ), levelBounds;
	std::vector<int> stamp( w.bodies.size(), -1 ), slotOf( w.bodies.size() ), last( w.bodies.size() );
	std::vector<int> contactLevel( kOverflowChunkContacts ), levelCount;
	for ( int ch = 0; ch < chunkCount; ++ch )
	{
		int c0 = ch * kOverflowChunkContacts;
		int c1 = std::min( n, c0 + kOverflowChunkContacts );
		int b0 = (int)bodyList.size();
		int maxLevel = 0;
		for ( int k = c0; k < c1; ++k )
		{
			const Contact& contact = w.contacts[L.constraintContacts[L.overflowContactStart + k]];
			int level = 0;
			for ( int side = 0; side < 2; ++side )
			{
				int id = contact.edges[side].bodyId;
				if ( stamp[id] != ch )
				{
					stamp[id] = ch;
					slotOf[id] = (int)bodyList.size() - b0;
					bodyList.push_back( id );
					last[id] = -1;
				}
				slots[2 * k + side] = slotOf[id];
				if ( w.bodies[id].type == b2_dynamicBody )
				{
					level = std::max( level, last[id] + 1 );
				}
			}
			for ( int side = 0; side < 2; ++side )
			{
				int id = contact.edges[side].bodyId;
				if ( w.bodies[id].type == b2_dynamicBody )
				{
					last[id] = level;
				}
			}
			contactLevel[k - c0] = level;
			maxLevel = std::max( maxLevel, level );
		}

		// stable counting sort of the chunk's contacts by level
		levelCount.assign( maxLevel + 2, 0 );
		for ( int k = c0; k < c1; ++k )
		{
			levelCount[contactLevel[k - c0] + 1] += 1;
		}
		for ( int l = 0; l <= maxLevel; ++l )
		{
			levelCount[l + 1] += levelCount[l];
		}
		int levelBegin = (int)levelBounds.size();
		for ( int l = 0; l <= maxLevel + 1; ++l )
		{
			levelBounds.push_back( c0 + levelCount[l] );
		}
		for ( int k = c0; k < c1; ++k )
		{
			entries[c0 + levelCount[contactLevel[k - c0]]++] = k - c0;
		}
		int b1 = (int)bodyList.size();
		int levelEnd = (int)levelBounds.size() - 1;
		table.insert( table.end(), { c0, c1, b0, b1, levelBegin, levelEnd } );
	}

	std::vector<int>& plan = L.overflowPlan;
	plan.assign( 5, 0 );
	plan[0] = chunkCount;
	plan.insert( plan.end(), table.begin(), table.end() );
	plan[1] = (int)plan.size();
	plan.insert( plan.end(), bodyList.begin(), bodyList.end() );
	plan[2] = (int)plan.size();
	plan.insert( plan.end(), slots.begin(), slots.end() );
	plan[3] = (int)plan.size();
	plan.insert( plan.end(), entries.begin(), entries.end() );
	plan[4] = (int)plan.size();
	plan.insert( plan.end(), levelBounds.begin(), levelBounds.end() );
}

static void BuildSolverLists( World& w, SolverLists& L )
{
	L = SolverLists{};
	// the awake order may have changed: the island slots go up again
	w.islandsDirty = true;
	L.awakeBodies = w.solverSets[kAwakeSet].bodyIds;

	for ( int i = 0; i < kGraphColorCount; ++i )
	{
		L.collideContacts.insert( L.collideContacts.end(), w.colors[i].contactIds.begin(), w.colors[i].contactIds.end() );
	}
	const std::vector<int>& nonTouching = w.solverSets[kAwakeSet].contactIds;
	L.collideContacts.insert( L.collideContacts.end(), nonTouching.begin(), nonTouching.end() );

	for ( int i = 0; i < kOverflowIndex; ++i )
	{
		const GraphColor& color = w.colors[i];
		if ( color.contactIds.size() + color.jointIds.size() == 0 )
		{
			continue;
		}
		L.colorContactStart.push_back( (int)L.constraintContacts.size() );
		L.colorJointStart.push_back( (int)L.jointList.size() );
		L.constraintContacts.insert( L.constraintContacts.end(), color.contactIds.begin(), color.contactIds.end() );
		L.jointList.insert( L.jointList.end(), color.jointIds.begin(), color.jointIds.end() );
		L.activeColorCount += 1;
	}
	L.colorContactStart.push_back( (int)L.constraintContacts.size() );
	L.colorJointStart.push_back( (int)L.jointList.size() );

	const GraphColor& overflow = w.colors[kOverflowIndex];
	L.overflowContactStart = (int)L.constraintContacts.size();
	L.overflowContactCount = (int)overflow.contactIds.size();
	L.constraintContacts.insert( L.constraintContacts.end(), overflow.contactIds.begin(), overflow.contactIds.end() );
	L.overflowJointStart = (int)L.jointList.size();
	L.overflowJointCount = (int)overflow.jointIds.size();
	L.jointList.insert( L.jointList.end(), overflow.jointIds.begin(), overflow.jointIds.end() );

	BuildOverflowPlan( w, L );
}

// Per-body warm-start lists for the color-kernel path (SolverLists::warmStart/warmList). The warm
// start adds velocity deltas that depend only on impulses, anchors and delta rotations, never on
// velocities, so applying a body's deltas in stage order in one pass gives the same floats as the
// color-by-color stage.
static void BuildWarmLists( World& w, SolverLists& L )
{
	int nb = (int)L.awakeBodies.size();
	auto awakeDynamic = [&]( int bodyId ) {
		const Body& b = w.bodies[bodyId];
		return b.setIndex == kAwakeSet && b.type == b2_dynamicBody ? b.localIndex : -1;
	};
	// visit( awake index, entry ) for every entry in stage order
	auto forEachEntry = [&]( auto&& visit ) {
		auto contact = [&]( int k ) {
			const Contact& c = w.contacts[L.constraintContacts[k]];
			int a = awakeDynamic( c.edges[0].bodyId );
			int b = awakeDynamic( c.edges[1].bodyId );
			if ( a >= 0 )
			{
				visit( a, ( k << 3 ) | ( 1 << 2 ) );
			}
			if ( b >= 0 )
			{
				visit( b, ( k << 3 ) | ( a < 0 ? 1 << 2 : 0 ) | 1 );
			}
		};
		auto joint = [&]( int k ) {
			const Joint& j = w.joints[L.jointList[k]];
			int a = awakeDynamic( j.edges[0].bodyId );
			int b = awakeDynamic( j.edges[1].bodyId );
			if ( a >= 0 )
			{
				visit( a, ( k << 3 ) | 2 );
			}
			if ( b >= 0 )
			{
				visit( b, ( k << 3 ) | 2 | 1 );
			}
		};
		for ( int k = L.overflowJointStart; k < L.overflowJointStart + L.overflowJointCount; ++k )
		{
			joint( k );
		}
		for ( int k = L.overflowContactStart; k < L.overflowContactStart + L.overflowContactCount; ++k )
		{
			contact( k );
		}
		for ( int c = 0; c < L.activeColorCount; ++c )
		{
			for ( int k = L.colorJointStart[c]; k < L.colorJointStart[c + 1]; ++k )
			{
				joint( k );
			}
			for ( int k = L.colorContactStart[c]; k < L.colorContactStart[c + 1]; ++k )
			{
				contact( k );
			}
		}
	};
	L.warmStart.assign( nb + 1, 0 );
	forEachEntry( [&]( int body, int ) { L.warmStart[body + 1] += 1; } );
	for ( int i = 0; i < nb; ++i )
	{
		L.warmStart[i + 1] += L.warmStart[i];
	}
	L.warmList.resize( L.warmStart[nb] );
	std::vector<int> fill( L.warmStart.begin(), L.warmStart.end() - 1 );
	forEachEntry( [&]( int body, int entry ) { L.warmList[fill[body]++] = entry; } );
}

void World::Step( float timeStep, int subStepCount )
{
	if ( StepBegin( timeStep, subStepCount ) == false )
	{
		return;
	}
	StepMiddle();
	StepEnd();
}

// Step phase 1: pairs, uploads, collide launch. Returns false when the world is locked.
bool World::StepBegin( float timeStep, int subStepCount )
{
	if ( locked )
	{
		return false;
	}

	if ( solverSets[kAwakeSet].bodyIds.empty() == false )
	{
		// the solve's launches go to a helper thread, which starts spinning now
		device->ArmLaunchHelper();
	}

	moveEvents.clear();
	moveEventBodyIds.clear();
	moveEventsNeedTransforms = false;
	contactBeginEvents.clear();
	contactHitEvents.clear();
	jointEvents.clear();
	sensorBeginEvents.clear();
	profile = b2Profile{};
	recycledContactCount = 0;

	locked = true;
	stepTicks = b2GetTicks();

	{
		uint64_t pairTicks = b2GetTicks();
		UpdateBroadPhasePairs();
		profile.pairs = b2GetMilliseconds( pairTicks );
	}

	StepParams params = {};
	params.dt = timeStep;
	params.subStepCount = b2MaxInt( 1, subStepCount );
	if ( timeStep > 0.0f )
	{
		params.inv_dt = 1.0f / timeStep;
		params.h = timeStep / params.subStepCount;
		params.inv_h = params.subStepCount * params.inv_dt;
	}
	inv_h = params.inv_h;
	inv_dt = params.inv_dt;

	float hertz = b2MinFloat( contactHertz, 0.125f * params.inv_h );
	params.contactSoftness = HostMakeSoft( hertz, contactDampingRatio, params.h );
	params.staticSoftness = HostMakeSoft( 2.0f * hertz, contactDampingRatio, params.h );
	params.restitutionThreshold = restitutionThreshold;
	params.maxLinearVelocity = maxLinearSpeed;
	params.contactSpeed = contactSpeed;
	params.contactHertz = contactHertz;
	params.contactDampingRatio = contactDampingRatio;
	params.hitEventThreshold = hitEventThreshold;
	params.gravity = gravity;
	params.contactRecycleDistance = contactRecycleDistance;
	params.enableWarmStarting = enableWarmStarting ? 1 : 0;
	params.enableSleep = enableSleep ? 1 : 0;
	params.enableContinuous = enableContinuous ? 1 : 0;
	params.enableContactSoftening = enableContactSoftening ? 1 : 0;
	params.customMixing = CustomMixing();
	if ( mixingStale )
	{
		if ( params.customMixing != 0 )
		{
			device->MarkContactMixingStale();
		}
		mixingStale = false;
	}

	// Host edits made since the last step go to the GPU before any stage reads them.
	uint64_t setupTicks = b2GetTicks();
	PushBodiesIfDirty();
	if ( shapesDirty )
	{
		device->UploadShapes( *this );
		shapesDirty = false;
	}
	if ( jointsDirty )
	{
		device->UploadJoints( *this );
		jointsDirty = false;
		jointEventJointCount = 0;
		for ( const Joint& joint : joints )
		{
			if ( joint.jointId != B2_NULL_INDEX )
			{
				const JointSim& js = jointSims[joint.jointId];
				jointEventJointCount += js.forceThreshold < FLT_MAX || js.torqueThreshold < FLT_MAX ? 1 : 0;
			}
		}
	}
	params.jointEvents = jointEventJointCount > 0 ? 1 : 0;
	params.preSolve = preSolveFcn != nullptr ? 1 : 0;
	params.sensorHits = 0;
	for ( const Sensor& sensor : sensors )
	{
		const Shape& shape = shapes[sensor.shapeId];
		if ( shape.enableSensorEvents && bodies[shape.bodyId].type == b2_staticBody )
		{
			params.sensorHits = 1;
			break;
		}
	}
	// the pre-continuous callback decides hits inside the sweep: with it, every fast body sweeps
	// on the host
	hostContinuousStep = false;
	if ( preContinuousFcn != nullptr && enableContinuous )
	{
		for ( const Shape& shape : shapes )
		{
			if ( shape.id != B2_NULL_INDEX && shape.enablePreSolveEvents )
			{
				hostContinuousStep = true;
				break;
			}
		}
	}
	params.hostContinuous = hostContinuousStep ? 1 : 0;
	device->UploadContactInits( pendingContactInits, contactIdPool.Capacity() );
	pendingContactInits.clear();
	BuildContinuousFilterRejects();

	if ( listsDirty )
	{
		BuildSolverLists( *this, solverLists );
		device->UploadLists( solverLists );
		listsDirty = false;
		listsStableSteps = 0;
	}
	else
	{
		// A contact graph that stayed the same for two steps gets the per-body warm-start lists
		// (worlds whose graph changes every step keep the color-by-color warm start).
		listsStableSteps += 1;
		int nb = (int)solverLists.awakeBodies.size();
		if ( listsStableSteps >= 2 && solverLists.warmStart.empty() &&
			 device->WantsWarmLists( nb, (int)solverLists.constraintContacts.size(), solverLists.activeColorCount,
									 solverLists.overflowContactCount + solverLists.overflowJointCount ) )
		{
			BuildWarmLists( *this, solverLists );
			device->UploadWarmLists( solverLists );
		}
	}

	profile.solverSetup = b2GetMilliseconds( setupTicks );

	stepParams = params;
	collideTicks = b2GetTicks();
	// the narrowphase and the contact scatter rewrite the device contacts
	contactsStale = true;
	device->CollideLaunch( params );
	return true;
}

// Step phase 2: collide results, touch processing, solve launch.
void World::StepMiddle()
{
	CollideResult cr;
	device->CollideFinish( &cr );
	recycledContactCount = cr.recycledCount;
	if ( cr.deferred.empty() == false )
	{
		// contacts with pre-solve events: their narrowphase runs here, around the user callback
		PullBodies();
		PullShapes();
		PullContacts();
		device->CollideOnHost( *this, cr.deferred, stepParams, &cr.changes );
		std::sort( cr.changes.begin(), cr.changes.end() );
	}
	uint64_t stateTicks = b2GetTicks();
	if ( cr.changes.empty() == false )
	{
		ProcessContactStates( cr.changes.data(), (int)cr.changes.size() );
	}
	profile.prepareConstraints = b2GetMilliseconds( stateTicks );
	profile.collide = b2GetMilliseconds( collideTicks );

	solvePending = false;
	if ( stepParams.dt > 0.0f )
	{
		solveTicks = b2GetTicks();
		solvePending = SolveBegin( stepParams );
	}

	if ( treeRebuildPending )
	{
		broadPhase.trees[b2_dynamicBody].Rebuild( false );
		broadPhase.trees[b2_kinematicBody].Rebuild( false );
		treeRebuildPending = false;
	}
	// the trees' leaf order, staged ahead for the next GPU pair query: when the last update was on
	// the GPU, or when a world this large may need one after this solve (host work that overlaps
	// the solve)
	// (a batched world never runs the GPU pair query, so it stages nothing)
	const bool largeAwake = solvePending && (int)solverSets[kAwakeSet].bodyIds.size() >= kGpuPairThreshold &&
							device->Recording() == false;
	if ( lastPairQueryPath >= 2 || largeAwake )
	{
		for ( b2BodyType type : { b2_dynamicBody, b2_kinematicBody } )
		{
			const ProxyStagingCache& cache = stagingCache[type];
			if ( cache.valid == false || cache.version != broadPhase.trees[type].StructureVersion() )
			{
				FillStagingCache( type );
			}
		}
	}
	// the GPU pair finder's copies of the shape filters and the pair set follow the host's while
	// the GPU solves, so that a query launched after the solve uploads no backlog
	if ( largeAwake && pairFinder != nullptr )
	{
		SyncPairFinderMirrors();
	}

	// The next step's pair query behind this solve, when the last solve enlarged enough proxies
	// for a GPU query to pay.
	if ( solvePending && ( pairQuery == 2 || ( pairQuery == 0 && 2 * lastEnlargedShapeCount >= kGpuPairThreshold ) ) &&
		 device->Recording() == false )
	{
		LaunchSpeculativePairs();
	}
}

// Step phase 3: solve results, refit, sleep.
void World::StepEnd()
{
	if ( solvePending )
	{
		SolveEnd();
		profile.solve = b2GetMilliseconds( solveTicks );
		solvePending = false;
	}
	if ( speculative.launched && speculative.snapshot == false )
	{
		// the refit has put the enlarged boxes into the trees: later box changes invalidate
		for ( int t = 0; t < b2_bodyTypeCount; ++t )
		{
			speculative.boxVersion[t] = broadPhase.trees[t].BoxVersion();
		}
		speculative.snapshot = true;
	}
	if ( sensors.empty() == false )
	{
		uint64_t sensorTicks = b2GetTicks();
		OverlapSensors();
		profile.sensors = b2GetMilliseconds( sensorTicks );
	}
	profile.step = b2GetMilliseconds( stepTicks );

	// swap the end event buffers
	endEventArrayIndex = 1 - endEventArrayIndex;
	contactEndEvents[endEventArrayIndex].clear();
	sensorEndEvents[endEventArrayIndex].clear();
	locked = false;
}

b2ContactEvents World::GetContactEvents()
{
	b2ContactEvents events = {};
	if ( locked )
	{
		return events;
	}
	// the previous buffer holds the last step's end events
	std::vector<b2ContactEndTouchEvent>& ends = contactEndEvents[1 - endEventArrayIndex];
	events.beginEvents = contactBeginEvents.data();
	events.endEvents = ends.data();
	events.hitEvents = contactHitEvents.data();
	events.beginCount = (int)contactBeginEvents.size();
	events.endCount = (int)ends.size();
	events.hitCount = (int)contactHitEvents.size();
	return events;
}

b2SensorEvents World::GetSensorEvents()
{
	b2SensorEvents events = {};
	if ( locked )
	{
		return events;
	}
	std::vector<b2SensorEndTouchEvent>& ends = sensorEndEvents[1 - endEventArrayIndex];
	events.beginEvents = sensorBeginEvents.data();
	events.endEvents = ends.data();
	events.beginCount = (int)sensorBeginEvents.size();
	events.endCount = (int)ends.size();
	return events;
}

// Sensor overlaps of the step (reference src/sensor.c:140-390). The device tests every sensor
// against every shape (the tree query's box and mask tests, the callback's filters and the GJK
// distance); the host applies the custom filter, merges the continuous-collision hits, keeps each
// sensor's overlaps sorted by shape id and reports what began and ended, sensor by sensor.
void World::OverlapSensors()
{
	// continuous collision hits on static sensors, in this step's sensor hit lists
	std::vector<IdPair> hits;
	device->ContinuousSensorHits( &hits );
	for ( IdPair h : hits )
	{
		const Shape& sensorShape = shapes[h.first];
		if ( sensorShape.sensorIndex != B2_NULL_INDEX )
		{
			sensors[sensorShape.sensorIndex].hits.push_back( SensorVisitor{ h.second, shapes[h.second].generation } );
		}
	}

	int sensorCount = (int)sensors.size();
	std::vector<char> changed( sensorCount, 0 );
	std::vector<int> active;
	std::vector<int> activeShapes;
	for ( int i = 0; i < sensorCount; ++i )
	{
		Sensor& sensor = sensors[i];
		std::swap( sensor.overlaps1, sensor.overlaps2 );
		sensor.overlaps2 = sensor.hits;
		sensor.hits.clear();
		const Shape& shape = shapes[sensor.shapeId];
		if ( bodies[shape.bodyId].setIndex == kDisabledSet || shape.enableSensorEvents == false )
		{
			// the sensor drops every overlap
			changed[i] = sensor.overlaps1.empty() ? 0 : 1;
			continue;
		}
		active.push_back( i );
		activeShapes.push_back( sensor.shapeId );
	}

	std::vector<IdPair> overlaps;
	device->SensorOverlaps( activeShapes, &overlaps );
	for ( IdPair o : overlaps )
	{
		Sensor& sensor = sensors[active[o.first]];
		const Shape& sensorShape = shapes[sensor.shapeId];
		const Shape& other = shapes[o.second];
		if ( ( sensorShape.enableCustomFiltering || other.enableCustomFiltering ) && customFilterFcn != nullptr &&
			 customFilterFcn( ShapeIdOf( sensorShape.id ), ShapeIdOf( other.id ), customFilterContext ) == false )
		{
			continue;
		}
		sensor.overlaps2.push_back( SensorVisitor{ o.second, other.generation } );
	}

	for ( int i : active )
	{
		std::vector<SensorVisitor>& v = sensors[i].overlaps2;
		std::sort( v.begin(), v.end(), []( const SensorVisitor& a, const SensorVisitor& b ) { return a.shapeId < b.shapeId; } );
		v.erase( std::unique( v.begin(), v.end(), []( const SensorVisitor& a, const SensorVisitor& b ) { return a.shapeId == b.shapeId; } ),
				 v.end() );
		const std::vector<SensorVisitor>& w = sensors[i].overlaps1;
		bool same = v.size() == w.size();
		for ( size_t k = 0; same && k < v.size(); ++k )
		{
			same = v[k].shapeId == w[k].shapeId && v[k].generation == w[k].generation;
		}
		changed[i] = same ? 0 : 1;
	}

	for ( int i = 0; i < sensorCount; ++i )
	{
		if ( changed[i] == 0 )
		{
			continue;
		}
		const Sensor& sensor = sensors[i];
		b2ShapeId sensorId = ShapeIdOf( sensor.shapeId );
		const std::vector<SensorVisitor>& r1 = sensor.overlaps1;
		const std::vector<SensorVisitor>& r2 = sensor.overlaps2;
		auto endEvent = [&]( const SensorVisitor& v ) {
			sensorEndEvents[endEventArrayIndex].push_back( b2SensorEndTouchEvent{ sensorId, b2ShapeId{ v.shapeId + 1, worldId, v.generation } } );
		};
		auto beginEvent = [&]( const SensorVisitor& v ) {
			sensorBeginEvents.push_back( b2SensorBeginTouchEvent{ sensorId, b2ShapeId{ v.shapeId + 1, worldId, v.generation } } );
		};
		size_t i1 = 0, i2 = 0;
		while ( i1 < r1.size() && i2 < r2.size() )
		{
			const SensorVisitor& a = r1[i1];
			const SensorVisitor& b = r2[i2];
			if ( a.shapeId == b.shapeId )
			{
				if ( a.generation < b.generation )
				{
					endEvent( a );
					i1 += 1;
				}
				else if ( a.generation > b.generation )
				{
					beginEvent( b );
					i2 += 1;
				}
				else
				{
					i1 += 1;
					i2 += 1;
				}
			}
			else if ( a.shapeId < b.shapeId )
			{
				endEvent( a );
				i1 += 1;
			}
			else
			{
				beginEvent( b );
				i2 += 1;
			}
		}
		for ( ; i1 < r1.size(); ++i1 )
		{
			endEvent( r1[i1] );
		}
		for ( ; i2 < r2.size(); ++i2 )
		{
			beginEvent( r2[i2] );
		}
	}
}

bool World::SolveBegin( StepParams& params )
{
	stepIndex += 1;

	SolverSet& awakeSet = solverSets[kAwakeSet];
	int awakeBodyCount = (int)awakeSet.bodyIds.size();
	if ( awakeBodyCount == 0 )
	{
		return false;
	}

	// Woken bodies may have rewritten states on the host.
	PushBodiesIfDirty();

	if ( listsDirty )
	{
		BuildSolverLists( *this, solverLists );
		device->UploadLists( solverLists );
		listsDirty = false;
		listsStableSteps = 0;
	}

	// The island split changes only island bookkeeping, which the solve does not read, but the
	// finalize kernel's island sleep reduction must see the split islands, as the reference's
	// finalize does (src/solver.c:1475-1494): the split runs before the launch.
	uint64_t splitTicks = b2GetTicks();
	if ( splitIslandId != B2_NULL_INDEX )
	{
		SplitIsland( splitIslandId );
		splitIslandId = B2_NULL_INDEX;
	}
	profile.splitIslands = b2GetMilliseconds( splitTicks );

	// each awake body's island slot and id for the finalize kernel, resent when islands or the
	// awake order changed
	if ( enableSleep && islandsDirty )
	{
		const std::vector<int>& awake = awakeSet.bodyIds;
		islandSlotStaging.resize( awake.size() );
		islandIdStaging.resize( awake.size() );
		for ( size_t i = 0; i < awake.size(); ++i )
		{
			const Island& island = islands[bodies[awake[i]].islandId];
			islandSlotStaging[i] = island.localIndex | ( island.constraintRemoveCount > 0 ? (int)0x80000000u : 0 );
			islandIdStaging[i] = island.islandId;
		}
		device->UploadIslands( islandSlotStaging, islandIdStaging, (int)awakeSet.islandIds.size() );
		islandsDirty = false;
	}

	constraintTicks = b2GetTicks();
	device->SolveLaunchAsync( params, restitutionShapeCount > 0, enableSleep );

	// Move events: one per awake body in awake order, filled while the GPU solves (the awake set
	// cannot change before the solve ends; transforms are fetched lazily).
	uint64_t eventTicks = b2GetTicks();
	moveEventBodyIds = awakeSet.bodyIds;
	moveEvents.resize( awakeBodyCount );
	for ( int i = 0; i < awakeBodyCount; ++i )
	{
		int bodyId = moveEventBodyIds[i];
		Body& body = bodies[bodyId];
		body.bodyMoveIndex = i;
		moveEvents[i].userData = body.userData;
		moveEvents[i].bodyId = b2BodyId{ bodyId + 1, worldId, body.generation };
		moveEvents[i].fellAsleep = false;
	}
	moveEventsNeedTransforms = true;
	profile.transforms = b2GetMilliseconds( eventTicks );
	return true;
}

void World::SolveEnd()
{
	SolverSet& awakeSet = solverSets[kAwakeSet];
	int awakeBodyCount = (int)awakeSet.bodyIds.size();

	FinalizeResult fr;
	device->SolveFinish( &fr );
	profile.constraints = b2GetMilliseconds( constraintTicks );
	// A solve that enlarged many proxies after one that enlarged few: the next pair update would
	// run the GPU query as a round trip of its own, so the query is launched now instead, behind the
	// solve's results, and runs while the host refits (the speculative query's inputs are ready).
	if ( speculative.launched == false && pairQuery == 0 && 2 * (int)fr.enlargedShapes.size() >= kGpuPairThreshold &&
		 device->Recording() == false )
	{
		LaunchSpeculativePairs();
	}
	lastEnlargedShapeCount = (int)fr.enlargedShapes.size();
	if ( speculative.launched )
	{
		// the speculative query's queries, in the device list's slot order
		speculative.shapes = fr.enlargedShapes;
	}

	// joint events in ascending joint id (reference src/solver.c:1648-1700)
	for ( int jointId : fr.jointEvents )
	{
		const Joint& joint = joints[jointId];
		jointEvents.push_back( b2JointEvent{ b2JointId{ jointId + 1, worldId, joint.generation }, joint.userData } );
	}

	// hit events in ascending contact id (reference src/solver.c:1703-1820)
	for ( const HitRecord& hr : fr.hits )
	{
		b2ContactHitEvent event = {};
		event.shapeIdA = ShapeIdOf( hr.shapeIdA );
		event.shapeIdB = ShapeIdOf( hr.shapeIdB );
		event.contactId = ContactIdOf( contacts[hr.contactId] );
		event.point = hr.point;
		event.normal = hr.normal;
		event.approachSpeed = hr.approachSpeed;
		contactHitEvents.push_back( event );
	}

	bodiesStale = true;
	shapesStale = true;
	jointsStale = true;
	contactsStale = true;

	// Refit: enlarge proxies of enlarged bodies in awake order (reference src/solver.c:1822-1915).
	{
		uint64_t refitTicks = b2GetTicks();
		std::vector<int> shapeSlot;
		if ( fr.enlargedShapes.empty() == false )
		{
			for ( size_t k = 0; k < fr.enlargedShapes.size(); ++k )
			{
				int shapeId = fr.enlargedShapes[k];
				if ( (int)shapeSlot.size() <= shapeId )
				{
					shapeSlot.resize( shapeId + 1, -1 );
				}
				shapeSlot[shapeId] = (int)k;
			}
		}

		// Fast bodies the device left to the host: bullets (their moves are buffered here, in body
		// order, and they sweep after the refit; reference src/solver.c:1871-1885) and, in a
		// host-continuous step, every other fast body, which sweeps against the static tree here
		// as the reference's finalize does before its refit (src/solver.c:648-660).
		std::vector<int> bullets, hostSwept;
		std::vector<IdPair> hostSensorHits;
		bool pulled = false;
		for ( int awakeIndex : fr.enlargedBodies )
		{
			int bodyId = moveEventBodyIds[awakeIndex];
			if ( ( bodies[bodyId].flags & kIsBullet ) != 0 || hostContinuousStep )
			{
				if ( pulled == false )
				{
					PullBodies();
					PullShapes();
					broadPhase.trees[b2_staticBody].FlushEnlarged();
					pulled = true;
				}
				if ( bodyAux[bodyId].flags & kIsFast )
				{
					if ( bodySims[bodyId].flags & kIsBullet )
					{
						for ( int shapeId = bodies[bodyId].headShapeId; shapeId != B2_NULL_INDEX; shapeId = shapes[shapeId].nextShapeId )
						{
							broadPhase.BufferMove( shapes[shapeId].proxyKey );
						}
						bullets.push_back( bodyId );
						continue;
					}
					if ( hostContinuousStep )
					{
						DeviceWorld::ContinuousOnHost( *this, bodyId, &hostSensorHits );
						bodySims[bodyId].flags &= ~kEnlargeBounds;
						for ( int shapeId = bodies[bodyId].headShapeId; shapeId != B2_NULL_INDEX; shapeId = shapes[shapeId].nextShapeId )
						{
							ShapeDev& sd = shapeData[shapeId];
							if ( sd.enlargedAABB )
							{
								broadPhase.EnlargeProxy( shapes[shapeId].proxyKey, sd.fatAABB );
								sd.enlargedAABB = 0;
							}
						}
						hostSwept.push_back( bodyId );
						continue;
					}
				}
			}

			int shapeId = bodies[bodyId].headShapeId;
			while ( shapeId != B2_NULL_INDEX )
			{
				Shape& shape = shapes[shapeId];
				if ( shapeId < (int)shapeSlot.size() && shapeSlot[shapeId] >= 0 )
				{
					b2AABB fatAABB = fr.enlargedShapeAABBs[shapeSlot[shapeId]];
					broadPhase.EnlargeProxy( shape.proxyKey, fatAABB );
				}
				shapeId = shape.nextShapeId;
			}
		}
		profile.refit = b2GetMilliseconds( refitTicks );

		// Bullets (reference src/solver.c:1917-1980): each sweeps against the refit trees, then the
		// proxies of its enlarged shapes grow in the dynamic tree (their moves are buffered above).
		if ( bullets.empty() == false )
		{
			uint64_t bulletTicks = b2GetTicks();
			for ( DynamicTree& tree : broadPhase.trees )
			{
				tree.FlushEnlarged();
			}
			for ( int bodyId : bullets )
			{
				DeviceWorld::ContinuousOnHost( *this, bodyId, &hostSensorHits );
			}
			DynamicTree& dynamicTree = broadPhase.trees[b2_dynamicBody];
			for ( int bodyId : bullets )
			{
				BodySim& sim = bodySims[bodyId];
				if ( ( sim.flags & kEnlargeBounds ) == 0 )
				{
					continue;
				}
				sim.flags &= ~kEnlargeBounds;
				for ( int shapeId = bodies[bodyId].headShapeId; shapeId != B2_NULL_INDEX; shapeId = shapes[shapeId].nextShapeId )
				{
					ShapeDev& sd = shapeData[shapeId];
					if ( sd.enlargedAABB )
					{
						sd.enlargedAABB = 0;
						dynamicTree.EnlargeProxy( B2G_PROXY_ID( shapes[shapeId].proxyKey ), sd.fatAABB );
					}
				}
			}
			hostSwept.insert( hostSwept.end(), bullets.begin(), bullets.end() );
			profile.bullets = b2GetMilliseconds( bulletTicks );
		}
		if ( hostSwept.empty() == false )
		{
			device->UploadBodyShapes( *this, hostSwept );
		}
		for ( IdPair h : hostSensorHits )
		{
			const Shape& sensorShape = shapes[h.first];
			if ( sensorShape.sensorIndex != B2_NULL_INDEX )
			{
				sensors[sensorShape.sensorIndex].hits.push_back( SensorVisitor{ h.second, shapes[h.second].generation } );
			}
		}
	}

	// Island sleep (reference src/solver.c:2024-2080): the finalize kernel marked the islands a
	// body keeps awake and picked the sleepiest island that needs a split
	if ( enableSleep )
	{
		uint64_t sleepTicks = b2GetTicks();
		int islandCount = (int)awakeSet.islandIds.size();
		if ( (int)fr.islandBits.size() * 32 < islandCount )
		{
			Fatal( "island sleep: %d island bits read back for %d awake islands", (int)fr.islandBits.size() * 32, islandCount );
		}
		splitIslandId = fr.splitKey != 0 ? (int)(uint32_t)fr.splitKey : B2_NULL_INDEX;

		for ( int islandIndex = islandCount - 1; islandIndex >= 0; islandIndex -= 1 )
		{
			if ( ( fr.islandBits[islandIndex >> 5] >> ( islandIndex & 31 ) ) & 1u )
			{
				continue;
			}
			int islandId = awakeSet.islandIds[islandIndex];
			TrySleepIsland( islandId );
		}
		profile.sleepIslands = b2GetMilliseconds( sleepTicks );
	}
}

void World::EnableSleeping( bool flag )
{
	if ( locked )
	{
		return;
	}
	if ( flag == enableSleep )
	{
		return;
	}
	enableSleep = flag;
	if ( flag == false )
	{
		int setCount = (int)solverSets.size();
		for ( int i = kFirstSleepingSet; i < setCount; ++i )
		{
			if ( solverSets[i].setIndex != B2_NULL_INDEX && solverSets[i].bodyIds.empty() == false )
			{
				WakeSolverSet( i );
			}
		}
	}
}

b2BodyEvents World::GetBodyEvents()
{
	if ( moveEventsNeedTransforms )
	{
		if ( bodiesStale )
		{
			// only the moved bodies' transforms come back, not the whole body arrays
			moveEventTransforms.resize( moveEvents.size() );
			device->DownloadTransforms( moveEventBodyIds.data(), (int)moveEventBodyIds.size(), moveEventTransforms.data() );
			for ( size_t i = 0; i < moveEvents.size(); ++i )
			{
				moveEvents[i].transform = moveEventTransforms[i];
			}
		}
		else
		{
			for ( size_t i = 0; i < moveEvents.size(); ++i )
			{
				moveEvents[i].transform = bodySims[moveEventBodyIds[i]].transform;
			}
		}
		moveEventsNeedTransforms = false;
	}
	return b2BodyEvents{ moveEvents.data(), (int)moveEvents.size() };
}

b2Counters World::GetCounters()
{
	b2Counters s = {};
	s.bodyCount = bodyIdPool.Count();
	s.shapeCount = shapeIdPool.Count();
	s.contactCount = contactIdPool.Count();
	s.jointCount = jointIdPool.Count();
	s.islandCount = islandIdPool.Count();
	s.staticTreeHeight = broadPhase.trees[b2_staticBody].GetHeight();
	s.treeHeight = b2MaxInt( broadPhase.trees[b2_dynamicBody].GetHeight(), broadPhase.trees[b2_kinematicBody].GetHeight() );
	s.recycledContactCount = recycledContactCount;
	s.awakeContactCount = 0;
	for ( int i = 0; i < kGraphColorCount; ++i )
	{
		s.colorCounts[i] = (int)( colors[i].contactIds.size() + colors[i].jointIds.size() );
		s.awakeContactCount += (int)colors[i].contactIds.size();
	}
	s.awakeContactCount += (int)solverSets[kAwakeSet].contactIds.size();
	return s;
}

// ---------------------------------------------------------------------------------------------
// World registry entry points used by the C API

World* CreateWorldInternal( const b2WorldDef* def, b2WorldId* outId )
{
	int index = -1;
	for ( int i = 0; i < B2_MAX_WORLDS; ++i )
	{
		if ( s_worlds[i] == nullptr )
		{
			index = i;
			break;
		}
	}
	if ( index < 0 )
	{
		*outId = b2_nullWorldId;
		return nullptr;
	}
	uint16_t gen = s_generations[index];
	World* world = new World( def, index, gen );
	s_worlds[index] = world;
	*outId = b2WorldId{ (uint16_t)( index + 1 ), gen };
	return world;
}

void DestroyWorldInternal( World* world )
{
	int index = world->worldId;
	s_generations[index] = (uint16_t)( world->generation + 1 );
	s_worlds[index] = nullptr;
	delete world;
}

void SetLogFcn( b2LogFcn* fcn )
{
	s_logFcn = fcn;
}

void SetAssertFcn( b2AssertFcn* fcn )
{
	s_assertFcn = fcn;
}

} // namespace b2g
