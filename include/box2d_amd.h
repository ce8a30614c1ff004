// box2d_amd extensions to the Box2D C API.
//
// These entry points have no counterpart in the reference library. They expose what a GPU step
// adds: per-kernel timing of the step for roofline measurement, and batched GPU evaluation of the
// narrowphase manifold functions (the GPU form of the reference's b2Collide* routines,
// reference include/box2d/collision.h:615-680) for callers that collide many shape pairs at once.
#pragma once

#include "box2d/box2d.h"

#ifdef __cplusplus
extern "C" {
#endif

/// Aggregated GPU time of one kernel class of the step.
typedef struct b2amd_KernelStats
{
	char name[32];
	/// number of launches recorded
	int launches;
	/// summed launch durations in milliseconds (HIP events on the world's stream)
	double totalMs;
	/// summed algorithmic bytes the launches had to move (fields read and written once)
	double algorithmicBytes;
} b2amd_KernelStats;

/// Record per-kernel HIP event timings during b2World_Step (adds a small overhead per launch).
B2_API void b2amd_World_EnableKernelTiming( b2WorldId worldId, bool flag );

/// Copy the accumulated kernel statistics. Returns the number of kernel classes written.
B2_API int b2amd_World_GetKernelStats( b2WorldId worldId, b2amd_KernelStats* stats, int capacity );

/// Clear the accumulated kernel statistics.
B2_API void b2amd_World_ResetKernelStats( b2WorldId worldId );

/// Solver path of a world's step. The fused path runs the whole substep loop in one workgroup
/// with the awake body states in LDS (up to 5400 awake bodies); the color-kernel path launches one
/// kernel per graph color and stage; the XCD path runs the whole substep loop in one launch on the
/// 32 compute units of one XCD, with a barrier among its workgroups between colors instead of
/// kernel boundaries (MI355X only; worlds without overflow constraints). Auto picks the fused path
/// for small worlds (up to 4096 contact constraints), else the color kernels. Results are
/// bit-identical across paths; the switch exists for testing and measurement. Fused falls back to
/// color kernels for worlds beyond its LDS budget, XCD to color kernels where it does not apply.
typedef enum b2amd_SolverPath
{
	b2amd_solverAuto = 0,
	b2amd_solverColorKernels = 1,
	b2amd_solverFused = 2,
	b2amd_solverXcd = 3,
} b2amd_SolverPath;

B2_API void b2amd_World_SetSolverPath( b2WorldId worldId, b2amd_SolverPath path );

/// Broadphase pair query of a world's step (the move-buffer query of reference
/// src/broad_phase.c:412-529). The GPU query stages every proxy in the reference trees' visit
/// order and tests each moved proxy's box against 64-proxy chunk boxes of that order, one
/// wavefront per query, or walks a Morton-ordered linear BVH with one lane per query
/// (b2amd_pairQueryGpuLbvh). The chunk scan also runs speculatively: behind each step's solve it
/// queries the proxies that step enlarges, and the next step uses that result when nothing that
/// it depends on (move buffer, trees, pair set, filters) changed in between (Auto and
/// b2amd_pairQueryGpu; b2amd_pairQueryGpuAtUpdate only queries at the pair update). All give the
/// same pairs in the same order; Auto uses the GPU when at least 128 proxies moved (outside
/// b2amd_StepWorlds, whose host phase uses the tree query).
typedef enum b2amd_PairQuery
{
	b2amd_pairQueryAuto = 0,
	b2amd_pairQueryHostTree = 1,
	b2amd_pairQueryGpu = 2,
	b2amd_pairQueryGpuLbvh = 3,
	b2amd_pairQueryGpuAtUpdate = 4,
} b2amd_PairQuery;

B2_API void b2amd_World_SetPairQuery( b2WorldId worldId, b2amd_PairQuery mode );

/// Path the last step's pair query took: 0 no query (empty move buffer), 1 host tree, 2 GPU at the
/// pair update, 3 GPU speculative (run behind the previous step's solve).
B2_API int b2amd_World_GetLastPairQuery( b2WorldId worldId );

/// Block until all GPU work issued by this library on the current device has completed
/// (a device-wide synchronize; b2World_Step already returns with its own work complete).
B2_API void b2amd_Synchronize( void );

/// 1 when the current device waits by spinning (hipDeviceScheduleSpin), which the library asks
/// for when it makes the process's first HIP call (unless B2AMD_SCHEDULE_YIELD is set); 0 when
/// the runtime was already set up by the host application with another wait mode.
B2_API int b2amd_SpinWaitActive( void );

/// A shape for batched narrowphase evaluation. type is a b2ShapeType; the union holds the
/// geometry in the shape's local frame.
typedef struct b2amd_Shape
{
	int type;
	int padding;
	union
	{
		b2Circle circle;
		b2Capsule capsule;
		b2Polygon polygon;
		b2Segment segment;
	};
} b2amd_Shape;

/// Evaluate count manifolds on the GPU: manifold i collides shapesA[i] with shapesB[i], where
/// transforms[i] is the pose of B in A's frame (as for the reference b2Collide* functions).
/// Pairs must be in the reference's primary order (A's type >= B's type in
/// circle < capsule < segment < polygon, with segment-vs-polygon taking the segment as A).
/// Host arrays in, host array out; synchronous. Returns 0 on success, -1 for an unsupported pair.
B2_API int b2amd_CollideBatch( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
							   b2LocalManifold* manifolds );

/// b2amd_CollideBatch for callers that cannot allocate device memory (code under HIP graph
/// capture, tight serving loops, hosts that forbid hipMalloc). No counterpart in the reference.
/// The pairs and manifolds are staged in a fixed device array that each GPU holds for the life of
/// the process; no call allocates or frees device memory. At most 4096 pairs per call: a larger
/// count returns -2 and logs the limit. Same arguments, results and -1 as b2amd_CollideBatch;
/// -3 when the staging array cannot be located on the current device. Calls that overlap on one
/// GPU share its staging array and race: serialize them.
B2_API int b2amd_CollideBatchNoAlloc( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms,
									  int count, b2LocalManifold* manifolds );

/// Evaluate count time-of-impact queries on the GPU (the continuous-collision kernel's code;
/// reference src/distance.c:1148-1441). Host arrays in, host array out; synchronous. Returns 0.
B2_API int b2amd_TimeOfImpactBatch( const b2TOIInput* inputs, int count, b2TOIOutput* outputs );

/// Evaluate count GJK distance queries on the GPU, each starting from and updating caches[i]
/// (reference src/distance.c:424-611). Host arrays in and out; synchronous. Returns 0.
B2_API int b2amd_ShapeDistanceBatch( const b2DistanceInput* inputs, b2SimplexCache* caches, int count, b2DistanceOutput* outputs );

/// Timing of the last b2amd_StepWorlds call's persistent solver launch.
typedef struct b2amd_BatchStats
{
	/// GPU time of the substep-loop launch over all worlds, milliseconds (HIP events)
	float solveKernelMs;
	/// workgroup size the autotuner chose (256, 512 or 1024)
	int solveThreads;
	/// number of persistent workgroups launched
	int solveGrid;
	/// worlds stepped
	int worlds;
	/// wall-clock milliseconds of the step's phases: host phase 1 (pairs, uploads), GPU collide,
	/// host phase 2 (touch processing, coloring), GPU solve, host phase 3 (refit, sleep)
	float phaseMs[5];
	/// algorithmic bytes of the substep-loop launch: per world, substeps x (380 B per contact
	/// constraint + 32 B per awake body + 256 B per joint) + 64 B per awake body
	double solveAlgorithmicBytes;
	/// per world-size bucket (awake bodies 1-64, 65-256, 257-1024, more): worlds in the bucket and
	/// their warm-start form (1 body by body); the workgroup size, workgroup count and GPU
	/// milliseconds of the persistent launch that ran them (one launch runs every bucket, with the
	/// workgroup size of the bucket that has the most worlds); zeros for an empty bucket
	int bucketWorlds[4];
	int bucketThreads[4];
	int bucketWarmByBody[4];
	int bucketWorkgroups[4];
	float bucketMs[4];
} b2amd_BatchStats;

/// Step count independent worlds by timeStep with subStepCount substeps, as one batch: host
/// bookkeeping runs on host threads and every GPU stage is one launch over all the worlds, the
/// substep loop a persistent launch over a world queue. Each world ends exactly as if stepped
/// alone with b2World_Step. A world whose awake bodies do not fit the persistent solver's LDS
/// (more than 5400) has its substep loop run on its own, by the per-world path, after the batch's.
/// Returns the number of worlds stepped; stats may be null.
B2_API int b2amd_StepWorlds( const b2WorldId* worldIds, int count, float timeStep, int subStepCount, b2amd_BatchStats* stats );

/// Workgroup size of b2amd_StepWorlds' persistent solver for every world-size bucket: 64, 128, 256,
/// 512 or 1024 threads, or 0 for each bucket's own (the default). The results do not depend on it.
/// Returns 0, or -1 for another value. No counterpart in the reference (its workers are threads).
B2_API int b2amd_SetBatchSolverThreads( int threads );

/// The persistent solver's configuration of one world-size bucket (0: 1-64 awake bodies, 1: 65-256,
/// 2: 257-1024, 3: more): workgroup size (64, 128, 256, 512 or 1024; a batch's launch uses the size
/// of the bucket with the most worlds) and warm-start form of the bucket's worlds (1 body by body,
/// 0 color by color), overriding the offline-tuned table; threads 0 restores the table's. The form
/// applies to worlds whose solver lists are built afterwards. Results do not depend on either.
/// Returns 0, or -1 for an invalid bucket or value. No counterpart in the reference.
B2_API int b2amd_SetBatchTuning( int bucket, int threads, int warmByBody );

/// The configuration in force for a bucket and the offline table's (any pointer may be null).
/// Returns 0, or -1 for an invalid bucket. No counterpart in the reference.
B2_API int b2amd_GetBatchTuning( int bucket, int* maxBodies, int* threads, int* warmByBody, int* tableThreads,
								 int* tableWarmByBody );

// ---------------------------------------------------------------------------------------------
// Differentiation of recorded steps.
//
// A tape records, for each step of a set of worlds (b2World_Step or b2amd_StepWorlds), the body
// state at the step's start and the step's contact constraints in solve order with the discrete
// decisions the step took (manifold recycled or recomputed, warm-start point matches). Each record
// is a checkpoint: the backward pass recomputes the step's substeps from it on the GPU and runs
// the hand-written adjoint of the step in reverse, step by step. The active contact set, the
// feature ids and every clamp pattern are those of the recorded forward (held fixed). Supported:
// dynamic bodies with polygon shapes against static or dynamic polygons, no joints, no overflow
// constraints, no restitution or rolling resistance, no speed caps, locks or continuous events.

/// Body state (or its gradient) in the form the adjoint uses: center of mass, rotation,
/// velocities.
typedef struct b2amd_BodyState
{
	b2Vec2 center;
	b2Rot rotation;
	b2Vec2 linearVelocity;
	float angularVelocity;
	float padding;
} b2amd_BodyState;

/// Step header of a record.
typedef struct b2amd_TapeStep
{
	int bodyCount;		  // awake bodies, slots 1..bodyCount (slot 0 is the static identity)
	int constraintCount;  // contact constraints in solve order
	int jointCount;		  // joints in the step (not differentiable)
	int activeColorCount; // graph colors in solve order
	int colorStart[25];	  // first constraint of each active color (+ sentinel)
	int overflowStart;	  // overflow constraints [overflowStart, overflowStart + overflowCount)
	int overflowCount;
	int subStepCount;
	float timeStep;
	float h;	 // substep
	float inv_h;
	float contactSpeed; // maximum contact push-out speed (the solver uses -contactSpeed)
	float maxLinearSpeed;
	float maxAngularSpeed; // 0.25 pi / timeStep
	b2Vec2 gravity;
	int enableWarmStarting;
	float inv_dt;
	int padding[2];
} b2amd_TapeStep;

/// One awake body of a record, by slot.
typedef struct b2amd_TapeBody
{
	int bodyId;		// index1 - 1 of the body's b2BodyId
	uint32_t flags; // solver flags (locks, speed cap, continuous)
	b2Vec2 center;
	b2Vec2 origin; // transform position
	b2Rot rotation;
	b2Vec2 linearVelocity;
	float angularVelocity;
	float invMass;
	float invInertia;
	float gravityScale;
	float linearDamping;
	float angularDamping;
	b2Vec2 localCenter;
	b2Vec2 force;
	float torque;
	float padding;
} b2amd_TapeBody;

/// One contact constraint of a record, after the step's narrowphase and constraint preparation.
typedef struct b2amd_TapeConstraint
{
	int contactId;
	int shapeIdA, shapeIdB;
	int bodyIdA, bodyIdB;
	int slotA, slotB; // body slots of the step (0: static body)
	int pointCount;
	/// bit 0: manifold recycled this step (anchors, normal and base separations kept, separations
	/// updated); bits 2-3 and 4-5: warm-start source of point 0 and 1 (0 none, 1 + index of the
	/// point of the previous step's manifold of the same contact)
	uint32_t info;
	uint16_t id[2];
	b2Vec2 normal;
	b2Vec2 anchorA[2]; // relative to body A's center of mass
	b2Vec2 anchorB[2];
	float separation[2];
	float baseSeparation[2]; // separations at the manifold's last full update
	float normalImpulse[2];	 // warm-start impulses
	float tangentImpulse[2];
	float friction;
	float restitution;
	float rollingResistance;
	float tangentSpeed;
	float biasRate, massScale, impulseScale; // constraint softness
	b2Rot cachedRotationA, cachedRotationB;	 // body rotations at the manifold's last full update
	float padding;
} b2amd_TapeConstraint;

typedef struct b2amd_Tape b2amd_Tape;

/// Start recording the steps of count worlds. Every later step of these worlds (b2World_Step or
/// b2amd_StepWorlds) appends a record until b2amd_Tape_Finish. A world is on at most one tape.
/// Recorded worlds use the fused solver. Returns null if a world is already on a tape.
B2_API b2amd_Tape* b2amd_Tape_Create( const b2WorldId* worldIds, int count );

/// Stop recording and capture each world's final body state. Returns the number of steps.
B2_API int b2amd_Tape_Finish( b2amd_Tape* tape );

/// Release the tape's device memory.
B2_API void b2amd_Tape_Destroy( b2amd_Tape* tape );

/// Steps recorded (equal for all worlds of a tape stepped together).
B2_API int b2amd_Tape_GetStepCount( b2amd_Tape* tape, int worldIndex );

/// Array strides of the gradient buffers: body and shape id capacities over the tape's worlds.
B2_API int b2amd_Tape_GetBodyStride( b2amd_Tape* tape );
B2_API int b2amd_Tape_GetShapeStride( b2amd_Tape* tape );

/// Copy one record to the host: step < step count gives the step's header, bodies and
/// constraints; step == step count gives the final state (bodies only). Capacities bound the
/// copies. Returns 0, or -1 for an invalid world or step.
B2_API int b2amd_Tape_ReadStep( b2amd_Tape* tape, int worldIndex, int step, b2amd_TapeStep* header, b2amd_TapeBody* bodies,
								int bodyCapacity, b2amd_TapeConstraint* constraints, int constraintCapacity );

/// Final body states of all worlds: states[world * bodyStride + bodyId] (bodies not awake at the
/// end are left untouched).
B2_API int b2amd_Tape_GetFinalStates( b2amd_Tape* tape, b2amd_BodyState* states );

/// Diagnostics of a backward pass.
typedef struct b2amd_AdjointStats
{
	/// GPU time of the backward launch, milliseconds (HIP events)
	float backwardMs;
	/// recomputed values that differ from the recorded forward (0: the recompute is bit-exact)
	int replayMismatches;
	/// union over worlds of unsupported features found (0: the gradient is complete); bits:
	/// 1 overflow constraints, 2 joints, 4 restitution or rolling resistance, 8 non-polygon
	/// manifold, 16 speed cap, 32 awake set changed, 64 motion locks, 128 continuous collision
	int unsupported;
	int worlds;
	int steps;
} b2amd_AdjointStats;

/// Reverse pass over the whole tape. finalGrad[world * bodyStride + bodyId] is the gradient of
/// a scalar loss with respect to the final body states; the pass writes the gradient with respect
/// to the initial body states to initialGrad (same layout) and the gradient with respect to each
/// shape's friction to shapeFrictionGrad[world * shapeStride + shapeId] (contact friction is
/// sqrt(frictionA * frictionB), shape frictions held constant over the run). Host arrays;
/// synchronous. Returns 0.
B2_API int b2amd_Tape_Backward( b2amd_Tape* tape, const b2amd_BodyState* finalGrad, b2amd_BodyState* initialGrad,
								float* shapeFrictionGrad, b2amd_AdjointStats* stats );

/// Current state of any body of a tape's world (static bodies included), in record form.
B2_API int b2amd_Tape_GetBody( b2amd_Tape* tape, int worldIndex, int bodyId, b2amd_TapeBody* body );

/// A polygon shape of a tape's world, with its body id and friction (for checkers).
B2_API int b2amd_Tape_GetShape( b2amd_Tape* tape, int worldIndex, int shapeId, b2amd_Shape* shape, int* bodyId, float* friction );

#ifdef __cplusplus
}
#endif
