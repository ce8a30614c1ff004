// Host-side world: the bookkeeping half of b2World_Step.
//
// The GPU runs the arithmetic of the step (narrowphase, constraint preparation, the sub-stepped
// soft-step solver and body finalization). The host keeps what decides *order*: id pools, the
// broadphase trees and move buffer, contact creation and destruction, graph coloring, islands and
// solver sets. Every ordering decision follows the reference so that the GPU solves the same
// constraints in the same Gauss-Seidel order and the run reproduces the reference bit for bit.
#pragma once

#include "core.h"
#include "device.h"
#include "layout.h"
#include "tree.h"

#include <unordered_set>
#include <vector>

namespace b2g
{

class DeviceWorld;

enum SolverSetType
{
	kStaticSet = 0,
	kDisabledSet = 1,
	kAwakeSet = 2,
	kFirstSleepingSet = 3,
};

constexpr int kGraphColorCount = B2_GRAPH_COLOR_COUNT;
constexpr int kOverflowIndex = kGraphColorCount - 1;
constexpr int kDynamicColorCount = kGraphColorCount - 4;

#define B2G_PROXY_TYPE( KEY ) ( (b2BodyType)( ( KEY ) & 3 ) )
#define B2G_PROXY_ID( KEY ) ( ( KEY ) >> 2 )
#define B2G_PROXY_KEY( ID, TYPE ) ( ( ( ID ) << 2 ) | ( TYPE ) )
#define B2G_SHAPE_PAIR_KEY( K1, K2 ) ( ( K1 ) < ( K2 ) ? (uint64_t)( K1 ) << 32 | (uint64_t)( K2 ) : (uint64_t)( K2 ) << 32 | (uint64_t)( K1 ) )

struct Body
{
	void* userData = nullptr;
	int setIndex = B2_NULL_INDEX;
	int localIndex = B2_NULL_INDEX;
	int headContactKey = B2_NULL_INDEX;
	int contactCount = 0;
	int headShapeId = B2_NULL_INDEX;
	int shapeCount = 0;
	int headChainId = B2_NULL_INDEX;
	int headJointKey = B2_NULL_INDEX;
	int jointCount = 0;
	int islandId = B2_NULL_INDEX;
	int islandIndex = B2_NULL_INDEX;
	float mass = 0.0f;
	float inertia = 0.0f;
	int bodyMoveIndex = B2_NULL_INDEX;
	int id = B2_NULL_INDEX;
	uint32_t flags = 0;
	b2BodyType type = b2_staticBody;
	uint16_t generation = 0;
	char name[B2_NAME_LENGTH + 1] = {};
};

struct Shape
{
	int id = B2_NULL_INDEX;
	int bodyId = B2_NULL_INDEX;
	int prevShapeId = B2_NULL_INDEX;
	int nextShapeId = B2_NULL_INDEX;
	int sensorIndex = B2_NULL_INDEX;
	b2ShapeType type = b2_circleShape;
	b2SurfaceMaterial material = {};
	float density = 0.0f;
	int proxyKey = B2_NULL_INDEX;
	b2Filter filter = {};
	void* userData = nullptr;
	uint16_t generation = 0;
	bool enableSensorEvents = false;
	bool enableContactEvents = false;
	bool enableCustomFiltering = false;
	bool enableHitEvents = false;
	bool enablePreSolveEvents = false;
};

// A shape a sensor overlaps, with the generation it had then (reference src/sensor.h).
struct SensorVisitor
{
	int shapeId;
	uint16_t generation;
};

// Per sensor shape: last step's and this step's overlaps (sorted by shape id) and the continuous
// collision hits of this step (reference src/sensor.h, src/shape.c:189-197).
struct Sensor
{
	int shapeId = B2_NULL_INDEX;
	std::vector<SensorVisitor> overlaps1;
	std::vector<SensorVisitor> overlaps2;
	std::vector<SensorVisitor> hits;
};

struct ChainShape
{
	int id = B2_NULL_INDEX;
	int bodyId = B2_NULL_INDEX;
	int nextChainId = B2_NULL_INDEX;
	std::vector<int> shapeIndices;
	std::vector<b2SurfaceMaterial> materials;
	uint16_t generation = 0;
};

struct ContactEdge
{
	int bodyId;
	int prevKey;
	int nextKey;
};

struct Contact
{
	ContactEdge edges[2];
	int islandId = B2_NULL_INDEX;
	int islandIndex = B2_NULL_INDEX;
	int setIndex = B2_NULL_INDEX;
	int colorIndex = B2_NULL_INDEX;
	int localIndex = B2_NULL_INDEX;
	int shapeIdA = B2_NULL_INDEX;
	int shapeIdB = B2_NULL_INDEX;
	int contactId = B2_NULL_INDEX;
	uint32_t flags = 0;
	uint32_t generation = 0;
};

struct JointEdge
{
	int bodyId;
	int prevKey;
	int nextKey;
};

struct Joint
{
	void* userData = nullptr;
	int setIndex = B2_NULL_INDEX;
	int colorIndex = B2_NULL_INDEX;
	int localIndex = B2_NULL_INDEX;
	JointEdge edges[2];
	int jointId = B2_NULL_INDEX;
	int islandId = B2_NULL_INDEX;
	int islandIndex = B2_NULL_INDEX;
	b2JointType type = b2_revoluteJoint;
	uint16_t generation = 0;
	bool collideConnected = false;
};

struct ContactLink
{
	int contactId;
	int bodyIdA;
	int bodyIdB;
};

struct JointLink
{
	int jointId;
	int bodyIdA;
	int bodyIdB;
};

struct Island
{
	int setIndex = B2_NULL_INDEX;
	int localIndex = B2_NULL_INDEX;
	int islandId = B2_NULL_INDEX;
	int constraintRemoveCount = 0;
	std::vector<int> bodies;
	std::vector<ContactLink> contacts;
	std::vector<JointLink> joints;
};

// A solver set lists the ids that live in it, in the reference's array order
// (reference src/solver_set.h:17-29). Only the order matters on the host; body and contact data
// stay in the id-indexed arrays.
struct SolverSet
{
	std::vector<int> bodyIds;
	std::vector<int> contactIds;
	std::vector<int> jointIds;
	std::vector<int> islandIds;
	int setIndex = B2_NULL_INDEX;
};

struct GraphColor
{
	BitSet bodySet;
	std::vector<int> contactIds;
	std::vector<int> jointIds;
};

struct MovePair
{
	int shapeIndexA;
	int shapeIndexB;
};

// Set of shape pair keys (never 0: a pair has two distinct shape ids), open addressing with
// linear probing in a power-of-two table kept at most half full; erase shifts the following
// run back so probes never meet tombstones. Holds the pairs that have a contact.
class PairKeySet
{
public:
	bool Contains( uint64_t key ) const
	{
		if ( count == 0 )
		{
			return false;
		}
		for ( size_t i = Slot( key );; i = ( i + 1 ) & mask )
		{
			uint64_t k = slots[i];
			if ( k == key )
			{
				return true;
			}
			if ( k == 0 )
			{
				return false;
			}
		}
	}

	void Insert( uint64_t key )
	{
		if ( 2 * ( count + 1 ) > slots.size() )
		{
			Grow();
		}
		for ( size_t i = Slot( key );; i = ( i + 1 ) & mask )
		{
			if ( slots[i] == key )
			{
				return;
			}
			if ( slots[i] == 0 )
			{
				slots[i] = key;
				MarkDirty( i );
				count += 1;
				version += 1;
				return;
			}
		}
	}

	void Erase( uint64_t key )
	{
		if ( count == 0 )
		{
			return;
		}
		size_t i = Slot( key );
		while ( slots[i] != key )
		{
			if ( slots[i] == 0 )
			{
				return;
			}
			i = ( i + 1 ) & mask;
		}
		// backward shift: move later members of the run into the hole when their home slot
		// does not lie cyclically in (hole, position]
		size_t hole = i;
		for ( size_t j = ( i + 1 ) & mask; slots[j] != 0; j = ( j + 1 ) & mask )
		{
			size_t home = Slot( slots[j] );
			bool stays = hole <= j ? ( hole < home && home <= j ) : ( hole < home || home <= j );
			if ( stays == false )
			{
				slots[hole] = slots[j];
				MarkDirty( hole );
				hole = j;
			}
		}
		slots[hole] = 0;
		MarkDirty( hole );
		count -= 1;
		version += 1;
	}

	// Changes whenever a key is inserted or erased.
	uint64_t Version() const
	{
		return version;
	}

	size_t Size() const
	{
		return count;
	}

private:
	void MarkDirty( size_t i )
	{
		if ( allDirty == false )
		{
			dirty.push_back( (uint32_t)i );
			if ( dirty.size() > slots.size() / 4 )
			{
				dirty.clear();
				allDirty = true;
			}
		}
	}

	size_t Slot( uint64_t key ) const
	{
		return (size_t)PairKeyHash( key ) & mask;
	}

	void Grow()
	{
		std::vector<uint64_t> old;
		old.swap( slots );
		size_t cap = old.empty() ? 1024 : 2 * old.size();
		slots.assign( cap, 0 );
		mask = cap - 1;
		count = 0;
		for ( uint64_t k : old )
		{
			if ( k != 0 )
			{
				Insert( k );
			}
		}
		dirty.clear();
		allDirty = true;
	}

public:
	// Every member in a fixed order, for world snapshots; a restored table is resent whole.
	template <typename A>
	void Transfer( A& a )
	{
		a.Vec( slots );
		a.Pod( mask );
		a.Pod( count );
		a.Pod( version );
		if ( A::kReading )
		{
			dirty.clear();
			allDirty = true;
		}
	}

	// After a restore: a version newer than any a mirror may hold.
	void BumpVersion( uint64_t floor )
	{
		version = ( version > floor ? version : floor ) + 1;
	}

	// slots written since the device mirror was last updated (allDirty: the whole table)
	std::vector<uint64_t> slots;
	std::vector<uint32_t> dirty;
	bool allDirty = true;
	void ClearDirty()
	{
		dirty.clear();
		allDirty = false;
	}

private:
	size_t mask = 0;
	size_t count = 0;
	uint64_t version = 0;
};

class BroadPhase
{
public:
	DynamicTree trees[b2_bodyTypeCount];
	BitSet movedProxies[b2_bodyTypeCount];
	std::vector<int> moveArray;
	PairKeySet pairSet;

	void BufferMove( int queryProxy )
	{
		b2BodyType proxyType = B2G_PROXY_TYPE( queryProxy );
		int proxyId = B2G_PROXY_ID( queryProxy );
		if ( movedProxies[proxyType].GetBit( proxyId ) == false )
		{
			movedProxies[proxyType].SetBitGrow( proxyId );
			moveArray.push_back( queryProxy );
		}
	}

	void UnBufferMove( int proxyKey );
	int CreateProxy( b2AABB aabb, b2BodyType proxyType, uint64_t categoryBits, int shapeIndex, bool forcePairCreation );
	void DestroyProxy( int proxyKey );
	void MoveProxy( int proxyKey, b2AABB aabb );
	void EnlargeProxy( int proxyKey, b2AABB aabb );
};

// Move-buffer size from which the automatic pair query runs on the GPU.
constexpr int kGpuPairThreshold = 128;

class World
{
public:
	explicit World( const b2WorldDef* def, int worldIndex, uint16_t generation );
	~World();

	// Public API helpers
	b2BodyId CreateBody( const b2BodyDef* def );
	b2ShapeId CreateShape( b2BodyId bodyId, const b2ShapeDef* def, const void* geometry, b2ShapeType type );
	int CreateShapeInternal( Body* body, const b2ShapeDef* def, const void* geometry, b2ShapeType type );
	b2ChainId CreateChain( b2BodyId bodyId, const b2ChainDef* def );
	b2JointId CreateRevoluteJoint( const b2RevoluteJointDef* def );
	b2JointId CreateDistanceJoint( const b2DistanceJointDef* def );
	b2JointId CreateFilterJoint( const b2FilterJointDef* def );
	b2JointId CreateMotorJoint( const b2MotorJointDef* def );
	b2JointId CreateMoverJoint( const b2MoverJointDef* def );
	b2JointId CreatePogoJoint( const b2PogoJointDef* def );
	b2JointId CreatePrismaticJoint( const b2PrismaticJointDef* def );
	b2JointId CreateWeldJoint( const b2WeldJointDef* def );
	b2JointId CreateWheelJoint( const b2WheelJointDef* def );
	void Step( float timeStep, int subStepCount );
	void DestroyBody( b2BodyId id );
	void DestroyJoint( b2JointId id );
	void SetBodyTransform( Body* body, b2Vec2 position, b2Rot rotation );
	void ApplyMassFromShapes( Body* body );
	void SetShapeGeometry( Shape* shape, b2ShapeType type, const void* geometry );
	void DestroyShape( Shape* shape, bool updateBodyMass );
	void SetBodyType( Body* body, b2BodyType type );
	void DisableBody( Body* body );
	void EnableBody( Body* body );
	void SyncBodyFlags( Body* body );
	void EnableBodySleep( Body* body, bool enable );
	void SetBodyBullet( Body* body, bool flag );
	void SetBodyAwake( Body* body, bool awake );
	void WakeTouching( Body* body );
	void SetShapeMaterial( Shape* shape, const b2SurfaceMaterial& material );
	void SetShapeFilter( Shape* shape, b2Filter filter );
	void SetShapeDensity( Shape* shape, float density, bool updateBodyMass );

	// The three phases of Step, split at the two points where the host waits for the GPU, so a
	// batch can run each phase for many worlds around one launch per stage.
	bool StepBegin( float timeStep, int subStepCount );
	void StepMiddle();
	void StepEnd();
	void EnableSleeping( bool flag );

	// Host copies of GPU state are refreshed lazily after a step.
	void PullBodies();
	void PushBodiesIfDirty();
	void PullShapes();
	void PullJoints();
	// Contacts' manifolds and materials as the device last left them (read-only host copy).
	const ContactSim& PulledContact( int contactId );
	void PullContacts();

	// Snapshot image of the simulation state (reference src/world_snapshot.c:477-590, 899-970):
	// returns the image size, writing it when it fits in capacity. Restore returns false and
	// leaves the world unchanged for an image it rejects.
	int Snapshot( uint8_t* image, int capacity );
	bool Restore( const uint8_t* image, int size );
	template <typename A>
	void TransferState( A& a );
	uint64_t StateHash();
	std::vector<ContactSim> contactSims;
	bool contactsStale = true;

	b2BodyEvents GetBodyEvents();
	b2Counters GetCounters();
	bool WakeBodyPublic( Body* body )
	{
		return WakeBody( body );
	}

	Body* GetBody( b2BodyId id );

	// identity
	uint16_t worldId;
	uint16_t generation;
	bool locked = false;

	// settings
	b2Vec2 gravity;
	float hitEventThreshold;
	float restitutionThreshold;
	float maxLinearSpeed;
	float contactSpeed;
	float contactHertz;
	float contactDampingRatio;
	float contactRecycleDistance;
	bool enableSleep;
	bool enableWarmStarting = true;
	bool enableContinuous;
	bool enableContactSoftening;
	void* userData = nullptr;
	int workerCount = 1;
	int stepIndex = 0;

	// user callbacks, run on the host (reference src/physics_world.c:272-290, src/broad_phase.c:281-294)
	b2CustomFilterFcn* customFilterFcn = nullptr;
	void* customFilterContext = nullptr;
	b2PreSolveFcn* preSolveFcn = nullptr;
	b2PreContinuousFcn* preContinuousFcn = nullptr;
	void* preSolveContext = nullptr;
	b2FrictionCallback* frictionCallback = nullptr;
	b2RestitutionCallback* restitutionCallback = nullptr;
	// a mixing callback or a shape material changed while a callback is set: every contact's
	// stored values are out of date until its next full narrowphase update
	bool mixingStale = false;
	std::vector<ContactMaterial> mixRefresh;
	int CustomMixing() const
	{
		return ( frictionCallback != nullptr ? 1 : 0 ) | ( restitutionCallback != nullptr ? 2 : 0 );
	}
	float MixFriction( const Shape& a, const Shape& b ) const;
	// (static shape << 32 | fast shape) pairs the custom filter rejects for continuous collision
	std::vector<uint64_t> ccdFilterRejects;
	void BuildContinuousFilterRejects();
	// this step's fast bodies all sweep on the host, from the refit (the pre-continuous callback
	// is in use); fast bullets always do, after the refit
	bool hostContinuousStep = false;
	float MixRestitution( const Shape& a, const Shape& b ) const;
	int splitIslandId = B2_NULL_INDEX;
	float inv_h = 0.0f;
	float inv_dt = 0.0f;

	// id pools and host records
	IdPool bodyIdPool, shapeIdPool, chainIdPool, contactIdPool, jointIdPool, islandIdPool, solverSetIdPool;
	std::vector<Body> bodies;
	std::vector<Shape> shapes;
	std::vector<ChainShape> chains;
	std::vector<Contact> contacts;
	std::vector<Joint> joints;
	std::vector<Island> islands;
	std::vector<SolverSet> solverSets;
	GraphColor colors[kGraphColorCount];
	BroadPhase broadPhase;

	// host mirrors of the id-indexed GPU arrays
	std::vector<BodySim> bodySims;
	std::vector<BodyState> bodyStates;
	std::vector<BodyAux> bodyAux;
	std::vector<ShapeDev> shapeData;
	std::vector<JointSim> jointSims;
	std::vector<ContactInit> pendingContactInits;

	bool bodiesDirty = true;	// host copy newer than device
	bool bodiesStale = false;	// device copy newer than host
	bool shapesDirty = true;
	bool shapesStale = false;
	bool jointsDirty = true;
	bool jointsStale = false;
	bool listsDirty = true; // solver work lists (awake bodies, colors, contacts) changed
	SolverLists solverLists; // the lists last uploaded
	bool islandsDirty = true; // island membership, slots or removed-constraint counts changed
	std::vector<int> islandSlotStaging;
	std::vector<int> islandIdStaging;
	int listsStableSteps = 0; // steps since the lists last changed
	int restitutionShapeCount = 0;

	// events and statistics
	std::vector<b2BodyMoveEvent> moveEvents;
	std::vector<int> moveEventBodyIds;
	std::vector<b2Transform> moveEventTransforms; // gathered from the device for the move events
	bool moveEventsNeedTransforms = false;
	b2Profile profile = {};
	int recycledContactCount = 0;

	DeviceWorld* device = nullptr;

	// Pair query: 0 auto (the GPU chunk scan for move buffers of at least kGpuPairThreshold
	// proxies, outside batched stepping), 1 host tree query, 2 GPU chunk scan, 3 GPU LBVH walk,
	// 4 GPU chunk scan at the pair update only. All give the same pairs in the same order. Modes
	// 0 and 2 also run the next step's query speculatively behind the solve (see
	// SpeculativePairs).
	int pairQuery = 0;
	class PairFinder* pairFinder = nullptr;
	int lastPairQueryPath = 0; // 1 host, 2 GPU at the update, 3 GPU speculative (for tests)

	// The next step's pair query run on the GPU behind this step's solve: its queries are the
	// shapes the solve enlarges (exactly the move buffer the refit leaves), its proxies the trees
	// as rebuilt during the solve with those boxes replaced. The result is used only if, at the
	// next pair update, the move buffer is that set and the trees, the pair set and the shape
	// filters are unchanged since (user calls between steps change one of them); otherwise the
	// update queries as usual.
	struct SpeculativePairs
	{
		bool launched = false;
		bool snapshot = false; // versions taken after the refit
		uint64_t pairSetVersion = 0;
		uint64_t filterVersion = 0;
		uint32_t structureVersion[b2_bodyTypeCount] = {};
		uint32_t boxVersion[b2_bodyTypeCount] = {};
		std::vector<int> shapes;	  // the solve's enlarged shapes in device slot order
		std::vector<int> queryOf;	  // by shape id: slot in shapes, or -1
		std::vector<int> moveIndexOf; // by slot: move-buffer index at the update
	};
	SpeculativePairs speculative;
	int lastEnlargedShapeCount = 0;

	// Contact events of the last step (reference src/physics_world.c:860-866, 975-983,
	// 1478-1505): begin and hit events are cleared when a step starts; end events are double
	// buffered because contacts can also end between steps (shape or body destruction).
	std::vector<b2ContactBeginTouchEvent> contactBeginEvents;
	std::vector<b2ContactEndTouchEvent> contactEndEvents[2];
	std::vector<b2ContactHitEvent> contactHitEvents;
	// joints whose force or torque reached a threshold in the last step (reference
	// src/physics_world.c:867, src/solver.c:1648-1700)
	std::vector<b2JointEvent> jointEvents;
	// sensors and their events (reference src/sensor.c:239-390): begin events are cleared when a
	// step starts, end events are double buffered with the contact end events
	std::vector<Sensor> sensors;
	std::vector<b2SensorBeginTouchEvent> sensorBeginEvents;
	std::vector<b2SensorEndTouchEvent> sensorEndEvents[2];
	b2SensorEvents GetSensorEvents();
	void OverlapSensors();
	void DestroySensor( Shape* shape );
	int jointEventJointCount = 0; // joints with a finite force or torque threshold at the last upload
	int endEventArrayIndex = 0;
	b2ContactEvents GetContactEvents();
	b2ShapeId ShapeIdOf( int shapeId ) const
	{
		return b2ShapeId{ shapeId + 1, worldId, shapes[shapeId].generation };
	}
	b2ContactId ContactIdOf( const Contact& contact ) const
	{
		return b2ContactId{ contact.contactId + 1, worldId, 0, contact.generation };
	}

private:
	friend class DeviceWorld;

	JointSim* CreateJoint( const b2JointDef* base, b2JointType type, b2JointId* id );
	void FinishJoint( b2JointId id );

	// bodies / sets / islands (reference body.c, solver_set.c, island.c)
	void UpdateBodyMassData( Body* body );
	void CreateIslandForBody( int setIndex, Body* body );
	Island* CreateIsland( int setIndex );
	void DestroyIsland( int islandId );
	int MergeIslands( int islandIdA, int islandIdB );
	void AddContactToIsland( int islandId, Contact* contact );
	void LinkContact( Contact* contact );
	void UnlinkContact( Contact* contact );
	void AddJointToIsland( int islandId, Joint* joint );
	void LinkJoint( Joint* joint );
	void SplitIsland( int baseId );
	void WakeSolverSet( int setIndex );
	void TrySleepIsland( int islandId );
	void MergeSolverSets( int setId1, int setId2 );
	void DestroySolverSet( int setIndex );
	bool WakeBody( Body* body );
	void UnlinkJoint( Joint* joint );
	void RemoveBodyFromIsland( Body* body );
	void DestroyJointInternal( Joint* joint );
	void RemoveBodyFromAwakeSet( Body* body );
	bool ShouldBodiesCollide( Body* bodyA, Body* bodyB );

	// shapes / contacts / graph
	void CreateShapeProxy( Shape* shape, bool forcePairCreation );
	void CreateContact( Shape* shapeA, Shape* shapeB );
	void DestroyContact( Contact* contact );
	void AddContactToGraph( Contact* contact );
	void RemoveContactFromGraph( int bodyIdA, int bodyIdB, int colorIndex, int localIndex );
	int AssignJointColor( int bodyIdA, int bodyIdB, b2BodyType typeA, b2BodyType typeB );
	void AddJointToGraph( Joint* joint );
	void RemoveJointFromGraph( int bodyIdA, int bodyIdB, int colorIndex, int localIndex );
	void DestroyContactsBetweenBodies( Body* bodyA, Body* bodyB );
	void DestroyBodyContacts( Body* body );
	void DestroyShapeProxy( Shape* shape );
	void TransferBody( int targetSetIndex, int sourceSetIndex, Body* body );
	void TransferJoint( int targetSetIndex, int sourceSetIndex, Joint* joint );
	void SetBodyTypeFlag( Body* body, b2BodyType type );

	// step stages
	void UpdateBroadPhasePairs();
	void QueryPairsGpuLaunch();
	void QueryPairsGpuFinish( std::vector<std::vector<MovePair>>& results );
	std::vector<int> pairQueryMove; // move-buffer index of each launched query
	bool treeRebuildPending = false; // the pair update's tree rebuild, run while the GPU solves

	// A tree's leaves in query order as last staged for the GPU pair query, valid while the
	// tree's structure version is unchanged. Later box changes come with a move-buffer entry, so
	// staging copies the cache and refreshes the moved proxies' boxes.
	struct ProxyStagingCache
	{
		uint32_t version = 0;
		bool valid = false;
		std::vector<LbvhProxy> proxies;
		std::vector<int> position; // by proxy id: index in proxies
	};
	ProxyStagingCache stagingCache[b2_bodyTypeCount];
	uint64_t shapeFilterVersion = 1; // bumped when a shape id is (re)assigned
	uint64_t syncedFilterVersion = 0; // the version the pair finder holds
	void FillStagingCache( b2BodyType type );
	int StageProxies(); // every proxy in visit order into the pair finder's staging; returns the count
	void SyncPairFinderMirrors();
	void LaunchSpeculativePairs();
	bool TakeSpeculativePairs( std::vector<std::vector<MovePair>>& results );
	bool AcceptPair( int queryProxyKey, int queryShapeIndex, int proxyKey, int shapeId, MovePair* pair );
	void ProcessContactStates( const int* changedIds, int count );
	bool SolveBegin( StepParams& params );
	void SolveEnd();
	StepParams stepParams = {};
	bool solvePending = false;
	uint64_t stepTicks = 0, collideTicks = 0, solveTicks = 0, constraintTicks = 0;
	void FinishSolve( StepParams& params );
	void ResetShapeProxy( Shape* shape, bool destroyProxy = true );
};

World* GetWorld( b2WorldId id );
World* GetWorldFromIndex( int index );

} // namespace b2g
