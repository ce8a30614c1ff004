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
/// kernel per graph color and stage. Auto picks the fused path for small worlds (up to 4096
/// contact constraints). Results are bit-identical across paths; the switch exists for testing
/// and measurement. Fused falls back to color kernels for worlds beyond its LDS budget.
typedef enum b2amd_SolverPath
{
	b2amd_solverAuto = 0,
	b2amd_solverColorKernels = 1,
	b2amd_solverFused = 2,
} b2amd_SolverPath;

B2_API void b2amd_World_SetSolverPath( b2WorldId worldId, b2amd_SolverPath path );

/// Block until all GPU work issued by this library on the current device has completed
/// (a device-wide synchronize; b2World_Step already returns with its own work complete).
B2_API void b2amd_Synchronize( void );

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
} b2amd_BatchStats;

/// Step count independent worlds by timeStep with subStepCount substeps, as one batch: host
/// bookkeeping runs on host threads and every GPU stage is one launch over all the worlds, the
/// substep loop a persistent launch over a world queue. Each world ends exactly as if stepped
/// alone with b2World_Step. Every world's awake bodies must fit the fused solver (5400).
/// Returns the number of worlds stepped; stats may be null.
B2_API int b2amd_StepWorlds( const b2WorldId* worldIds, int count, float timeStep, int subStepCount, b2amd_BatchStats* stats );

#ifdef __cplusplus
}
#endif
