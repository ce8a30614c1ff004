// Test-side driver for the reference's own benchmark scenes. Nothing here is part of the product.
//
// The scene builders are the reference's: shared/benchmarks.c with its helpers shared/human.c
// and shared/utils.c, compiled from the reference's shared/ directory by oracle/Makefile. This file only
// runs them, the way the reference's benchmark app does (benchmark/main.c:150-159, 318-351:
// default world def, continuous on, the scene's step function before each b2World_Step of
// 1/60 s with 4 substeps), and reduces each step to numbers a test can compare.
//
// It is compiled twice, against two libraries with the same C API:
//   oracle/_ref/libb2ref_bench.so   the reference library, which writes the golden fixture;
//   oracle/_ref/libb2amd_bench.so   box2d_amd's library (GPU step), which the GPU tests check.
//
// A step reduces to the move-event count, the awake body count and an FNV-1a hash of the move
// events sorted by body index (index, generation, transform bits, fell-asleep flag). Sorting
// makes the hash independent of event order, so it pins the state of every body that moved.

#include "benchmarks.h"

#include "box2d/box2d.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef void CreateFcn( b2WorldId worldId );
typedef float StepFcn( b2WorldId worldId, int stepCount );

typedef struct
{
	const char* name;
	CreateFcn* createFcn;
	StepFcn* stepFcn;
	int totalStepCount;
} BenchScene;

// Same scenes, step functions and step counts as the reference benchmark app.
static const BenchScene s_scenes[] = {
	{ "compounds", CreateCompounds, NULL, 500 },
	{ "joint_grid", CreateJointGrid, NULL, 500 },
	{ "junkyard", CreateJunkyard, StepJunkyard, 800 },
	{ "large_pyramid", CreateLargePyramid, NULL, 500 },
	{ "many_pyramids", CreateManyPyramids, NULL, 200 },
	{ "rain", CreateRain, StepRain, 1000 },
	{ "smash", CreateSmash, NULL, 300 },
	{ "spinner", CreateSpinner, StepSpinner, 500 },
	{ "tumbler", CreateTumbler, NULL, 750 },
	{ "washer", CreateWasher, NULL, 500 },
};

#define BS_SCENE_COUNT ( (int)( sizeof( s_scenes ) / sizeof( s_scenes[0] ) ) )

// The scene builders keep their state in globals, so one scene runs at a time.
static b2WorldId s_worldId;
static int s_scene = -1;
static int s_stepIndex;
static b2BodyMoveEvent* s_sorted;
static int s_sortedCapacity;

static uint32_t FloatBits( float x )
{
	uint32_t u;
	memcpy( &u, &x, sizeof( u ) );
	return u;
}

static uint32_t Fnv1a( uint32_t hash, uint32_t word )
{
	for ( int b = 0; b < 4; ++b )
	{
		hash ^= ( word >> ( 8 * b ) ) & 0xFFu;
		hash *= 0x01000193u;
	}
	return hash;
}

static int CompareMoves( const void* a, const void* b )
{
	int ia = ( (const b2BodyMoveEvent*)a )->bodyId.index1;
	int ib = ( (const b2BodyMoveEvent*)b )->bodyId.index1;
	return ( ia > ib ) - ( ia < ib );
}

int bs_scene_count( void )
{
	return BS_SCENE_COUNT;
}

const char* bs_scene_name( int index )
{
	return index >= 0 && index < BS_SCENE_COUNT ? s_scenes[index].name : NULL;
}

int bs_scene_steps( int index )
{
	return index >= 0 && index < BS_SCENE_COUNT ? s_scenes[index].totalStepCount : 0;
}

void bs_destroy( void )
{
	if ( s_scene >= 0 )
	{
		b2DestroyWorld( s_worldId );
		s_scene = -1;
	}
}

// Builds scene `index` in a fresh world stepped by `workerCount` threads (the reference's own
// scheduler; box2d_amd ignores it). Returns the world id packed as index1 | generation << 16.
uint32_t bs_create_workers( int index, int workerCount )
{
	bs_destroy();
	if ( index < 0 || index >= BS_SCENE_COUNT )
	{
		return 0;
	}
	b2WorldDef worldDef = b2DefaultWorldDef();
	worldDef.enableContinuous = true;
	worldDef.workerCount = workerCount;
	s_worldId = b2CreateWorld( &worldDef );
	s_scenes[index].createFcn( s_worldId );
	s_scene = index;
	s_stepIndex = 0;
	return (uint32_t)s_worldId.index1 | (uint32_t)s_worldId.generation << 16;
}

uint32_t bs_create( int index )
{
	return bs_create_workers( index, 1 );
}

// Runs `count` benchmark steps with no reduction (the timed loop of benchmark/main.c:343-352).
void bs_run( int count )
{
	if ( s_scene < 0 )
	{
		return;
	}
	const BenchScene* scene = s_scenes + s_scene;
	for ( int i = 0; i < count; ++i )
	{
		if ( scene->stepFcn != NULL )
		{
			scene->stepFcn( s_worldId, s_stepIndex );
		}
		b2World_Step( s_worldId, 1.0f / 60.0f, 4 );
		s_stepIndex += 1;
	}
}

// Runs one benchmark step and writes {move count, awake body count, sorted move-event hash}.
int bs_step( uint32_t* out )
{
	if ( s_scene < 0 )
	{
		return 0;
	}
	const BenchScene* scene = s_scenes + s_scene;
	if ( scene->stepFcn != NULL )
	{
		scene->stepFcn( s_worldId, s_stepIndex );
	}
	b2World_Step( s_worldId, 1.0f / 60.0f, 4 );
	s_stepIndex += 1;

	b2BodyEvents events = b2World_GetBodyEvents( s_worldId );
	if ( events.moveCount > s_sortedCapacity )
	{
		s_sortedCapacity = events.moveCount + events.moveCount / 2;
		s_sorted = (b2BodyMoveEvent*)realloc( s_sorted, (size_t)s_sortedCapacity * sizeof( b2BodyMoveEvent ) );
	}
	if ( events.moveCount > 0 )
	{
		memcpy( s_sorted, events.moveEvents, (size_t)events.moveCount * sizeof( b2BodyMoveEvent ) );
		qsort( s_sorted, (size_t)events.moveCount, sizeof( b2BodyMoveEvent ), CompareMoves );
	}

	uint32_t hash = 0x811C9DC5u;
	for ( int i = 0; i < events.moveCount; ++i )
	{
		const b2BodyMoveEvent* e = s_sorted + i;
		hash = Fnv1a( hash, (uint32_t)e->bodyId.index1 );
		hash = Fnv1a( hash, e->bodyId.generation );
		hash = Fnv1a( hash, FloatBits( e->transform.p.x ) );
		hash = Fnv1a( hash, FloatBits( e->transform.p.y ) );
		hash = Fnv1a( hash, FloatBits( e->transform.q.c ) );
		hash = Fnv1a( hash, FloatBits( e->transform.q.s ) );
		hash = Fnv1a( hash, e->fellAsleep ? 1u : 0u );
	}
	out[0] = (uint32_t)events.moveCount;
	out[1] = (uint32_t)b2World_GetAwakeBodyCount( s_worldId );
	out[2] = hash;
	return 1;
}

// The last step's move events sorted by body index, 7 words each (index, generation, transform
// bits, fell-asleep flag), for reporting the first body that differs.
int bs_moves( uint32_t* out, int capacity )
{
	if ( s_scene < 0 )
	{
		return 0;
	}
	b2BodyEvents events = b2World_GetBodyEvents( s_worldId );
	int n = events.moveCount < capacity / 7 ? events.moveCount : capacity / 7;
	for ( int i = 0; i < n; ++i )
	{
		const b2BodyMoveEvent* e = s_sorted + i;
		uint32_t* r = out + 7 * i;
		r[0] = (uint32_t)e->bodyId.index1;
		r[1] = e->bodyId.generation;
		r[2] = FloatBits( e->transform.p.x );
		r[3] = FloatBits( e->transform.p.y );
		r[4] = FloatBits( e->transform.q.c );
		r[5] = FloatBits( e->transform.q.s );
		r[6] = e->fellAsleep ? 1u : 0u;
	}
	return n;
}
