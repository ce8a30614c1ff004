// GPU half of a world: device arrays, host<->device synchronization and the stage launches of
// one b2World_Step. Only plain types cross this interface so the host bookkeeping (world.cpp)
// compiles without HIP headers.
#pragma once

#include "layout.h"

#include <cstdint>
#include <mutex>
#include <vector>

namespace b2g
{

class World;
class TapeStore;

// Work lists the host rebuilds when the set of awake bodies or the graph coloring changes.
struct SolverLists
{
	std::vector<int> awakeBodies;		  // body ids in awake-set order
	std::vector<int> collideContacts;	  // every awake contact (graph colors + non-touching)
	std::vector<int> constraintContacts; // colored contacts by active color, then overflow contacts
	std::vector<int> colorContactStart;	  // per active color: first constraint index (+ sentinel)
	std::vector<int> jointList;			  // colored joints by active color, then overflow joints
	std::vector<int> colorJointStart;	  // per active color: first joint index (+ sentinel)
	int activeColorCount = 0;
	int overflowContactStart = 0;
	int overflowContactCount = 0;
	int overflowJointStart = 0;
	int overflowJointCount = 0;

	// Warm start by body: for awake index i, entries [warmStart[i], warmStart[i + 1]) of warmList
	// are the body's constraints in the warm-start stage's order (overflow joints, overflow
	// contacts, then per color its joints and contacts). Entry: index << 3 | owner << 2 | joint << 1
	// | side (1: the body is the constraint's body B). index is a constraint index for contacts and
	// a jointList index for joints; owner marks the one entry per contact that also accumulates
	// the contact's total normal impulse. Only dynamic bodies have entries.
	std::vector<int> warmStart;
	std::vector<int> warmList;

	// Level schedule of the overflow contacts (empty when there are none). The overflow joints
	// come first in every stage and are walked in order before the contacts. The overflow sequence is cut into chunks of at most
	// kOverflowChunkContacts consecutive contacts; within a chunk each contact's level is one more
	// than the last level that touched either of its dynamic bodies, so contacts of one level share
	// no dynamic body and every body still sees its contacts in sequence order. Layout (ints):
	// [0] chunk count, [1] bodies offset, [2] slots offset, [3] entries offset, [4] level bounds
	// offset, then per chunk {contact begin, contact end, body begin, body end, level begin, level
	// end}; bodies: each chunk's distinct body ids; slots: per contact the chunk-local slots of its
	// bodies A and B; entries: per chunk its contacts (chunk-local index) sorted by level; level
	// bounds: per chunk the entry positions where each level starts, plus the end.
	std::vector<int> overflowPlan;
};

constexpr int kOverflowChunkContacts = 512;
constexpr int kOverflowChunkBodies = 2 * kOverflowChunkContacts;

// A pair of ids: (sensor, visitor) shape ids, or (active sensor index, shape id).
struct IdPair
{
	int first;
	int second;
};

// A contact's friction and restitution from the user's mixing callbacks.
struct ContactMaterial
{
	int contactId;
	float friction;
	float restitution;
};

// Results the host needs after the collide stage.
struct CollideResult
{
	// contact id * 4 + kind (0 mixing values out of date, 1 disjoint, 2 started touching,
	// 3 stopped touching), ascending
	std::vector<int> changes;
	int recycledCount = 0;
	// contacts left to the host's pre-solve callback (ascending id), untouched by the device
	std::vector<int> deferred;
};

// Results the host needs after body finalization.
// A contact hit event found by the device (contact and shape ids, world point, normal, speed).
struct HitRecord
{
	int contactId;
	int shapeIdA;
	int shapeIdB;
	b2Vec2 point;
	b2Vec2 normal;
	float approachSpeed;
};

struct FinalizeResult
{
	std::vector<HitRecord> hits; // ascending contact id
	std::vector<int> enlargedBodies; // awake indices of bodies with an enlarged shape, ascending
	std::vector<int> enlargedShapes; // shape ids whose fat AABB grew
	std::vector<b2AABB> enlargedShapeAABBs;
	std::vector<uint32_t> islandBits; // awake islands by slot in the awake set (sleep enabled)
	uint64_t splitKey = 0;			  // sleepiest split candidate: sleep time bits << 32 | island id
	int fastBodyCount = 0;
	std::vector<int> jointEvents; // ids of joints that reached a force or torque threshold, ascending
};

// Device results of a launched solve that the next step's speculative pair query reads: the
// shapes whose fat box grew (finalize and continuous collision, slot order) with their boxes, and
// an event on the world's stream after which they are complete.
struct SolvePairInputs
{
	const int* enlargedShapes;
	const b2AABB* enlargedShapeAABBs;
	const int* enlargedShapeCount; // kept by the pack kernel after it clears its own counter
	int capacity;				   // entries the enlarged-shape list can hold
	void* ready;				   // hipEvent_t
};

class DeviceWorld
{
public:
	DeviceWorld();
	~DeviceWorld();

	// The inputs of a speculative pair query behind the last SolveLaunch (false when it was
	// recorded for a batch instead of launched).
	bool GetSolvePairInputs( SolvePairInputs* out ) const;

	// Host -> device uploads of id-indexed arrays (whole-array copies; used when the host edited them).
	void UploadBodies( const World& world );
	void UploadShapes( const World& world );
	void UploadJoints( const World& world );
	// Per awake body: its island's slot in the awake set (negative when the island has removed
	// constraints) and the island id, read by the finalize kernel's island sleep reduction.
	void UploadIslands( const std::vector<int>& slots, const std::vector<int>& ids, int islandCount );
	// The whole contact array from the host copies (a restored world).
	void UploadContacts( const ContactSim* sims, int count );
	// Overlaps of the given sensor shapes with every shape in a broadphase tree, after the step's
	// solve: (index into sensorShapes, shape id), unordered.
	void SensorOverlaps( const std::vector<int>& sensorShapes, std::vector<IdPair>* overlaps );
	// The last solve's continuous-collision hits on static sensors: (sensor shape, fast shape).
	void ContinuousSensorHits( std::vector<IdPair>* hits );
	// The narrowphase of contacts the device deferred for the pre-solve callback, run on the
	// host copies (bodies, shapes and contacts must be current there) with the world's callback;
	// the updated contacts go back to the device and their change records are appended.
	void CollideOnHost( World& world, const std::vector<int>& contactIds, const StepParams& params,
						std::vector<int>* changes );
	// Continuous collision of one fast body on the host copies of bodies and shapes (current there)
	// with the world's custom filter and pre-continuous callbacks: against the static tree, and for
	// a bullet also against the kinematic and dynamic trees (flushed by the caller). Enlarged shapes
	// get enlargedAABB and the body kEnlargeBounds; sensor hits are appended as (sensor shape, fast
	// shape). The device copies are not touched.
	static void ContinuousOnHost( World& world, int bodyId, std::vector<IdPair>* sensorHits );
	// The given bodies' sims and all their shapes, from the host copies.
	void UploadBodyShapes( const World& world, const std::vector<int>& bodyIds );
	// Sorted (static shape << 32 | fast shape) keys the custom filter rejects for continuous collision.
	void UploadContinuousFilterRejects( const std::vector<uint64_t>& keys );
	void UploadContactInits( const std::vector<ContactInit>& inits, int contactCapacity );
	// User mixing callbacks: flag every contact's friction and restitution as out of date (the
	// narrowphase then reports each at its next full update), and upload new values.
	void MarkContactMixingStale();
	void UploadContactMaterials( const std::vector<ContactMaterial>& list );
	void UploadLists( const SolverLists& lists );
	// Upload the per-body warm-start lists of the lists last uploaded (they must match them).
	void UploadWarmLists( const SolverLists& lists );

	// Device -> host downloads (after a step, on demand).
	void DownloadBodies( World& world );
	void DownloadShapes( World& world );
	// The transforms of the given bodies (ids below the uploaded body count), in the given order.
	void DownloadTransforms( const int* bodyIds, int count, b2Transform* out );
	void DownloadJoints( World& world );
	void DownloadContact( int contactId, ContactSim* out );
	void DownloadContacts( int count, ContactSim* out ); // the first count contacts (by id)

	// Stages. Each runs as a launch half (enqueue, or record when recording for a batch) and a
	// finish half (wait for the stream, read the results back).
	void Collide( const StepParams& params, CollideResult* result );
	void Solve( const StepParams& params, bool anyRestitution, bool needSleepTimes, FinalizeResult* result );
	void CollideLaunch( const StepParams& params );
	void CollideFinish( CollideResult* result );
	void SolveLaunch( const StepParams& params, bool anyRestitution, bool needSleepTimes );
	// SolveLaunch on the process's helper thread when it is free (else here), so that the caller's
	// host work overlaps the launches; every later device call of this world joins it first.
	void SolveLaunchAsync( const StepParams& params, bool anyRestitution, bool needSleepTimes );
	void JoinLaunch() const;
	// Wake the helper thread ahead of a step's solve launch.
	void ArmLaunchHelper();
	void SolveFinish( FinalizeResult* result );

	// Record instead of launching (the batch that steps many worlds launches the records).
	void SetRecording( bool flag );
	// A recorded world whose awake bodies exceed the batch's persistent solver records no solve
	// stages; the batch launches its solve here, on the world's own stream, after its own.
	void LaunchDeferredSolve();
	bool Recording() const;
	void* Record();

	// Per-stage GPU timings of the last step in milliseconds (collide, solve, finalize).
	float lastCollideMs = 0.0f;
	float lastSolveMs = 0.0f;

	// Optional per-kernel timing (HIP events around every launch).
	void EnableKernelTiming( bool flag );
	int GetKernelStats( void* stats, int capacity ); // b2amd_KernelStats array
	void ResetKernelStats();

	// Solver path: 0 auto, 1 color kernels, 2 fused when the world fits.
	void SetSolverPath( int path );
	// Whether the next solve of a world with these list sizes takes the color-kernel path with the
	// warm start body by body (the host then builds the per-body warm-start lists).
	bool WantsWarmLists( int bodyCount, int constraintCount, int activeColorCount, int overflowCount ) const;

	// Step recording for the adjoint: while a tape is set, every solve appends one record (body
	// state at the step's start, constraints in solve order) to the tape. index is the world's
	// slot on the tape. The solve uses the fused path.
	void SetTape( TapeStore* tape, int index );
	TapeStore* Tape() const;
	// Write the final body state record of a finished tape (bodies only) and wait for it.
	void CaptureFinalState( char* dst, int bodyCount );
	// Device addresses of the id-indexed shape array (ShapeDev) and body sims (BodySim).
	const void* ShapesDevice() const;
	const void* SimsDevice() const;
	int AwakeBodyCount() const;

	void* stream = nullptr;

	struct Impl;

private:
	Impl* impl;
};

// Steps many worlds together: the worlds record their stage arguments (DeviceWorld recording
// mode) and the batch launches one kernel per stage over all of them on its own stream. The
// substep loop runs as one persistent launch whose workgroups take worlds from a queue.
class Batch
{
public:
	Batch();
	~Batch();

	// Launch the recorded collide stage of the worlds and wait for it.
	void Collide( DeviceWorld** worlds, int count );
	// Launch the recorded solve stages (prepare .. finalize, continuous, readback) and wait.
	void Solve( DeviceWorld** worlds, int count );

	float lastFusedMs = 0.0f; // all buckets' persistent launches, first start to last end
	double lastFusedBytes = 0.0;
	// per world-size bucket of the last batch step: worlds and their warm-start form, and the
	// workgroup size, workgroups and GPU time of the launch that ran them (zeros: no worlds)
	int lastBucketWorlds[4] = {};
	int lastBucketThreads[4] = {};
	int lastBucketWarmByBody[4] = {};
	int lastBucketGrid[4] = {};
	float lastBucketMs[4] = {};

	struct Impl;

private:
	Impl* impl;
};

struct BatchStats
{
	float solveKernelMs; // persistent substep-loop launches of the last batch step
	int solveThreads;	 // workgroup size of the bucket with the most worlds
	int solveGrid;		 // workgroups over all buckets
	int worlds;			 // worlds stepped
	float phaseMs[5];	 // wall time: host phase 1, GPU collide, host phase 2, GPU solve, host phase 3
	double solveBytes;	 // algorithmic bytes of the substep-loop launches
	int bucketWorlds[4];
	int bucketThreads[4];
	int bucketWarmByBody[4];
	int bucketGrid[4];
	float bucketMs[4];
};

// The batched persistent solver's configuration for worlds of up to maxBodies awake bodies: the
// workgroup size that steps one world (a batch's single launch takes the size of the bucket with
// the most worlds) and the warm-start form of the bucket's worlds (1 body by body, 0 color by
// color). Neither changes any result. Four buckets by awake body count: 1-64, 65-256, 257-1024,
// 1025 up to the persistent solver's LDS capacity.
constexpr int kBatchBuckets = 4;
struct BatchTuning
{
	int maxBodies;
	int threads;
	int warmByBody;
};
int BatchBucketOf( int bodyCount );
// The configuration in force for a bucket: the offline-tuned table unless set otherwise.
BatchTuning BatchTuningFor( int bucket );
// Override a bucket's configuration (threads 64, 128, 256, 512 or 1024; warmByBody 0 or 1), or
// restore the table's with threads 0. False for an invalid bucket or value.
bool SetBatchTuning( int bucket, int threads, int warmByBody );
// The table's configuration for a bucket (what the offline tuner chose).
BatchTuning BatchTuningTable( int bucket );

// Step count worlds as one batch (defined in batch.cpp). Returns the number of worlds stepped.
int StepWorlds( World** worlds, int count, float timeStep, int subStepCount, BatchStats* stats );

// GPU broadphase pair query (broadphase.hip): every proxy of a world in the reference trees' visit
// order, queried with the fat boxes of the moved proxies through 64-proxy chunk boxes (or a
// Morton-ordered LBVH).
struct LbvhProxy
{
	int proxyKey; // proxy id << 2 | body type of its tree
	int shapeId;
	b2AABB aabb; // the tree's fat box
};

// What the pair callback's shape tests read of a shape (b2Filter and the owning body).
struct PairShapeFilter
{
	uint64_t categoryBits;
	uint64_t maskBits;
	int groupIndex;
	int bodyId;
};

// Home-slot hash of the existing-pair set: the 64-bit MurmurHash3 finalizer. The host set and the
// device query's mirror of it probe the same slots.
constexpr uint64_t PairKeyHash( uint64_t key )
{
	key ^= key >> 33;
	key *= 0xff51afd7ed558ccdull;
	key ^= key >> 33;
	key *= 0xc4ceb9fe1a85ec53ull;
	key ^= key >> 33;
	return key;
}

// A shape's proxy in a staged pair query: its key and its position in the staged proxies
// (-1, -1 for a shape without a proxy).
struct ShapeSlot
{
	int proxyKey;
	int position;
};

class PairFinder
{
public:
	PairFinder();
	~PairFinder();

	// Device mirror of the world's existing-pair set (open addressing, linear probing, 0 = empty),
	// updated before a Launch: the whole table when `all` (or its capacity changed), otherwise
	// only the listed slots, which take their current values. The query then drops candidates
	// whose shape pair is already in the set.
	void SyncPairSet( const uint64_t* slots, size_t capacity, const uint32_t* dirty, size_t dirtyCount, bool all );

	// Pinned staging of the per-shape filters (indexed by shape id) and their upload; the query
	// then drops candidates of one body or whose filters reject each other.
	PairShapeFilter* FilterStaging( int count );
	void SyncShapeFilters( int count );

	// Pinned host staging the caller fills before Find: every proxy of the world in the order
	// pairs are to be listed, the moved proxies' queries, and 1 per shape id whose proxy is in the
	// move buffer.
	LbvhProxy* ProxyStaging( int count );
	LbvhProxy* QueryStaging( int count );
	uint8_t* MovedStaging( int count );

	// Search structure of the following launches: 0 a scan of 64-proxy chunk boxes, one wavefront
	// per query (the default); 1 a Morton-ordered LBVH walked by one lane per query.
	void SetMode( int mode );

	// Every (query, proxy) whose fat boxes overlap, that passes the pair callback's self, tree,
	// moved-proxy, same-body and filter rules and whose shape pair is not in the mirrored pair set,
	// as (query index << 32 | proxy index) keys, each query's keys in ascending proxy order (the
	// LBVH sorts all keys; the chunk scan may interleave different queries' keys). Launch
	// uploads the staging and starts the device work without waiting (the staging may be refilled
	// only after Finish); Finish waits and returns the keys, valid until the next Launch.
	void Launch( int proxyCount, int queryCount, int movedCount );
	const uint64_t* Finish( int* count );

	// Speculative query for the next step, enqueued behind a launched solve: the queries are the
	// solve's enlarged shapes (with their new boxes, which also replace their staged proxies'
	// boxes), taken on the device once in.ready has fired. The caller fills ProxyStaging with the
	// trees' proxies as they will stand at the next pair update (except for those boxes) and
	// ShapeSlotStaging with (proxy key, staged position) per shape id, and syncs the pair set and
	// filters first. FinishSpeculative waits and returns the keys (query index = the enlarged
	// shape's slot in the solve's list), the number of queries and whether the result is complete
	// (false when the candidates overflowed: the caller then queries again).
	ShapeSlot* ShapeSlotStaging( int count );
	void LaunchSpeculative( int proxyCount, int shapeCount, const SolvePairInputs& in );
	const uint64_t* FinishSpeculative( int* count, int* queryCount, bool* complete );

	struct Impl;

private:
	Impl* impl;
};

// Device storage of recorded steps (adjoint.hip). Records are carved out of large device chunks
// and stay valid until the store is destroyed.
class TapeStore
{
public:
	struct WorldTape
	{
		World* world = nullptr;
		std::vector<char*> steps; // device address of each step's record
		std::vector<int> bodyCounts;
		std::vector<int> constraintCounts;
		std::vector<int> subStepCounts;
		char* final = nullptr; // final state record (bodies only)
		int finalBodyCount = 0;
	};

	~TapeStore();
	char* Alloc( size_t bytes );

	std::vector<WorldTape> worlds;
	bool finished = false;

private:
	std::mutex mutex;
	std::vector<char*> chunks;
	size_t chunkUsed = 0;
	size_t chunkSize = 0;
};

// Record layout: header at 0, bodies at kTapeBodiesOffset, constraints after the bodies.
constexpr size_t kTapeBodiesOffset = 256;
size_t TapeConstraintsOffset( int bodyCount );
size_t TapeRecordBytes( int bodyCount, int constraintCount );

// Last HIP error text (empty if none) for diagnostics.
const char* LastDeviceError();

} // namespace b2g
