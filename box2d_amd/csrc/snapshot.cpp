// World snapshots: the simulation state of a world as one byte image, and its restore.
//
// The image is this library's own layout (the reference's holds its internal structs): a header
// (magic, version, a hash of the struct sizes it copies, payload size and an FNV-1a checksum of the
// payload) and the payload, every field in the fixed order of World::TransferState. The device
// copies of bodies, shapes, joints and contacts are pulled into the image; a restore uploads them
// all again and marks every derived device list and mirror out of date. User data pointers and
// callbacks are host wiring: the image holds none, and restored objects read back null user data.

#include "core.h"
#include "device.h"
#include "world.h"

#include <cstring>
#include <type_traits>

namespace b2g
{

namespace
{

constexpr uint32_t kSnapMagic = 0x504E5342u; // "BSNP"
constexpr uint32_t kSnapVersion = 1;

struct SnapHeader
{
	uint32_t magic;
	uint32_t version;
	uint32_t layoutHash;
	uint32_t reserved;
	uint64_t payloadSize;
	uint64_t checksum;
};

uint64_t Fnv64( const uint8_t* p, size_t n )
{
	uint64_t h = 14695981039346656037ull;
	for ( size_t i = 0; i < n; ++i )
	{
		h = ( h ^ p[i] ) * 1099511628211ull;
	}
	return h;
}

// Sizes of everything the image copies byte for byte: a changed struct changes the hash.
uint32_t LayoutHash()
{
	const size_t sizes[] = { sizeof( Body ),	 sizeof( Shape ),	   sizeof( Contact ),	  sizeof( Joint ),
							 sizeof( BodySim ),	 sizeof( BodyState ),  sizeof( BodyAux ),	  sizeof( ShapeDev ),
							 sizeof( JointSim ), sizeof( ContactSim ), sizeof( ContactInit ), sizeof( TreeNode ),
							 sizeof( ContactLink ), sizeof( JointLink ), sizeof( SensorVisitor ), sizeof( b2SurfaceMaterial ),
							 (size_t)kGraphColorCount };
	uint32_t h = 2166136261u;
	for ( size_t s : sizes )
	{
		h = ( h ^ (uint32_t)s ) * 16777619u;
	}
	return h;
}

class Writer
{
public:
	static constexpr bool kReading = false;
	std::vector<uint8_t> bytes;

	bool Fits( uint64_t )
	{
		return true;
	}

	template <typename T>
	void Pod( T& v )
	{
		static_assert( std::is_trivially_copyable<T>::value, "snapshot fields are copied byte for byte" );
		const uint8_t* p = reinterpret_cast<const uint8_t*>( &v );
		bytes.insert( bytes.end(), p, p + sizeof( T ) );
	}

	template <typename T>
	void Vec( std::vector<T>& v )
	{
		static_assert( std::is_trivially_copyable<T>::value, "snapshot arrays are copied byte for byte" );
		uint64_t n = v.size();
		Pod( n );
		const uint8_t* p = reinterpret_cast<const uint8_t*>( v.data() );
		bytes.insert( bytes.end(), p, p + n * sizeof( T ) );
	}
};

class Reader
{
public:
	static constexpr bool kReading = true;
	const uint8_t* p;
	const uint8_t* end;
	bool ok = true;

	// a record count the remaining bytes can hold (every record takes at least one byte)
	bool Fits( uint64_t n )
	{
		ok = ok && n <= (uint64_t)( end - p );
		return ok;
	}

	template <typename T>
	void Pod( T& v )
	{
		static_assert( std::is_trivially_copyable<T>::value, "snapshot fields are copied byte for byte" );
		if ( ok == false || (size_t)( end - p ) < sizeof( T ) )
		{
			ok = false;
			return;
		}
		memcpy( &v, p, sizeof( T ) );
		p += sizeof( T );
	}

	template <typename T>
	void Vec( std::vector<T>& v )
	{
		static_assert( std::is_trivially_copyable<T>::value, "snapshot arrays are copied byte for byte" );
		uint64_t n = 0;
		Pod( n );
		if ( ok == false || n > (uint64_t)( end - p ) / sizeof( T ) )
		{
			ok = false;
			return;
		}
		v.resize( (size_t)n );
		memcpy( v.data(), p, (size_t)n * sizeof( T ) );
		p += n * sizeof( T );
	}
};

template <typename A>
void TransferPool( A& a, IdPool& pool )
{
	a.Vec( pool.freeArray );
	a.Pod( pool.nextIndex );
}

// Arrays of records with a user data pointer: the image holds them with the pointer cleared.
template <typename A, typename T>
void TransferWithUserData( A& a, std::vector<T>& v )
{
	if ( A::kReading )
	{
		a.Vec( v );
		for ( T& item : v )
		{
			item.userData = nullptr;
		}
	}
	else
	{
		std::vector<T> copy = v;
		for ( T& item : copy )
		{
			item.userData = nullptr;
		}
		a.Vec( copy );
	}
}

} // namespace

template <typename A>
void World::TransferState( A& a )
{
	// settings that shape the simulation
	a.Pod( gravity );
	a.Pod( hitEventThreshold );
	a.Pod( restitutionThreshold );
	a.Pod( maxLinearSpeed );
	a.Pod( contactSpeed );
	a.Pod( contactHertz );
	a.Pod( contactDampingRatio );
	a.Pod( contactRecycleDistance );
	a.Pod( enableSleep );
	a.Pod( enableWarmStarting );
	a.Pod( enableContinuous );
	a.Pod( enableContactSoftening );
	a.Pod( stepIndex );
	a.Pod( splitIslandId );
	a.Pod( inv_h );
	a.Pod( inv_dt );
	a.Pod( restitutionShapeCount );
	a.Pod( mixingStale );
	a.Pod( lastEnlargedShapeCount );

	TransferPool( a, bodyIdPool );
	TransferPool( a, shapeIdPool );
	TransferPool( a, chainIdPool );
	TransferPool( a, contactIdPool );
	TransferPool( a, jointIdPool );
	TransferPool( a, islandIdPool );
	TransferPool( a, solverSetIdPool );

	TransferWithUserData( a, bodies );
	TransferWithUserData( a, shapes );
	TransferWithUserData( a, joints );
	a.Vec( contacts );

	uint64_t n = chains.size();
	a.Pod( n );
	if ( A::kReading )
	{
		chains.resize( a.Fits( n ) ? (size_t)n : 0 );
	}
	for ( ChainShape& c : chains )
	{
		a.Pod( c.id );
		a.Pod( c.bodyId );
		a.Pod( c.nextChainId );
		a.Pod( c.generation );
		a.Vec( c.shapeIndices );
		a.Vec( c.materials );
	}

	n = islands.size();
	a.Pod( n );
	if ( A::kReading )
	{
		islands.resize( a.Fits( n ) ? (size_t)n : 0 );
	}
	for ( Island& island : islands )
	{
		a.Pod( island.setIndex );
		a.Pod( island.localIndex );
		a.Pod( island.islandId );
		a.Pod( island.constraintRemoveCount );
		a.Vec( island.bodies );
		a.Vec( island.contacts );
		a.Vec( island.joints );
	}

	n = solverSets.size();
	a.Pod( n );
	if ( A::kReading )
	{
		solverSets.resize( a.Fits( n ) ? (size_t)n : 0 );
	}
	for ( SolverSet& set : solverSets )
	{
		a.Vec( set.bodyIds );
		a.Vec( set.contactIds );
		a.Vec( set.jointIds );
		a.Vec( set.islandIds );
		a.Pod( set.setIndex );
	}

	for ( GraphColor& color : colors )
	{
		a.Vec( color.bodySet.words );
		a.Vec( color.contactIds );
		a.Vec( color.jointIds );
	}

	for ( int t = 0; t < b2_bodyTypeCount; ++t )
	{
		broadPhase.trees[t].Transfer( a );
		a.Vec( broadPhase.movedProxies[t].words );
	}
	a.Vec( broadPhase.moveArray );
	broadPhase.pairSet.Transfer( a );

	n = sensors.size();
	a.Pod( n );
	if ( A::kReading )
	{
		sensors.resize( a.Fits( n ) ? (size_t)n : 0 );
	}
	for ( Sensor& sensor : sensors )
	{
		a.Pod( sensor.shapeId );
		a.Vec( sensor.overlaps1 );
		a.Vec( sensor.overlaps2 );
		a.Vec( sensor.hits );
	}

	// the id-indexed arrays the device steps, pulled current
	a.Vec( bodySims );
	a.Vec( bodyStates );
	a.Vec( bodyAux );
	a.Vec( shapeData );
	a.Vec( jointSims );
	a.Vec( contactSims );
	a.Vec( pendingContactInits );
}

int World::Snapshot( uint8_t* image, int capacity )
{
	if ( locked )
	{
		return 0;
	}
	PullBodies();
	PullShapes();
	PullJoints();
	PullContacts();
	contactSims.resize( contacts.size() );

	Writer w;
	w.bytes.resize( sizeof( SnapHeader ) );
	TransferState( w );
	size_t total = w.bytes.size();
	if ( total > (size_t)INT32_MAX )
	{
		Fatal( "world snapshot of %zu bytes exceeds the API's int size", total );
	}
	if ( image == nullptr || (size_t)capacity < total )
	{
		return (int)total;
	}
	SnapHeader header = {};
	header.magic = kSnapMagic;
	header.version = kSnapVersion;
	header.layoutHash = LayoutHash();
	header.payloadSize = total - sizeof( SnapHeader );
	header.checksum = Fnv64( w.bytes.data() + sizeof( SnapHeader ), (size_t)header.payloadSize );
	memcpy( w.bytes.data(), &header, sizeof( header ) );
	memcpy( image, w.bytes.data(), total );
	return (int)total;
}

bool World::Restore( const uint8_t* image, int size )
{
	if ( locked || image == nullptr || size < (int)sizeof( SnapHeader ) )
	{
		return false;
	}
	SnapHeader header;
	memcpy( &header, image, sizeof( header ) );
	if ( header.magic != kSnapMagic || header.version != kSnapVersion || header.layoutHash != LayoutHash() ||
		 header.payloadSize != (uint64_t)size - sizeof( SnapHeader ) ||
		 header.checksum != Fnv64( image + sizeof( SnapHeader ), (size_t)header.payloadSize ) )
	{
		return false;
	}

	// versions the live caches and mirrors may hold
	uint32_t structureFloor[b2_bodyTypeCount], boxFloor[b2_bodyTypeCount];
	for ( int t = 0; t < b2_bodyTypeCount; ++t )
	{
		structureFloor[t] = broadPhase.trees[t].StructureVersion();
		boxFloor[t] = broadPhase.trees[t].BoxVersion();
	}
	uint64_t pairFloor = broadPhase.pairSet.Version();

	Reader r;
	r.p = image + sizeof( SnapHeader );
	r.end = image + size;
	TransferState( r );
	if ( r.ok == false || r.p != r.end )
	{
		// the checksum matched, so the image was written by a build with another field order
		Fatal( "world snapshot payload does not match this library's layout" );
	}

	for ( int t = 0; t < b2_bodyTypeCount; ++t )
	{
		broadPhase.trees[t].BumpVersions( structureFloor[t], boxFloor[t] );
		stagingCache[t].valid = false;
	}
	broadPhase.pairSet.BumpVersion( pairFloor );
	shapeFilterVersion += 1;
	syncedFilterVersion = 0;
	speculative = SpeculativePairs{};
	treeRebuildPending = false;

	// events belong to the steps that made them
	moveEvents.clear();
	moveEventBodyIds.clear();
	moveEventsNeedTransforms = false;
	contactBeginEvents.clear();
	contactEndEvents[0].clear();
	contactEndEvents[1].clear();
	contactHitEvents.clear();
	jointEvents.clear();
	sensorBeginEvents.clear();
	sensorEndEvents[0].clear();
	sensorEndEvents[1].clear();
	mixRefresh.clear();
	ccdFilterRejects.clear();

	// every device copy is resent before the next step
	bodiesDirty = shapesDirty = jointsDirty = listsDirty = islandsDirty = true;
	bodiesStale = shapesStale = jointsStale = false;
	solverLists = SolverLists{};
	listsStableSteps = 0;
	device->UploadContacts( contactSims.data(), (int)contacts.size() );
	contactsStale = false;
	return true;
}

} // namespace b2g
