// Stepping many worlds together.
//
// Each world keeps its own host bookkeeping and device arrays, so its results are exactly those
// of stepping it alone. What changes is the schedule: the three host phases of a step run for all
// worlds in parallel on host threads, and between them each GPU stage is one launch over every
// world (recorded arguments, see DeviceWorld's recording mode), with the substep loop as one
// persistent launch. A step of N worlds costs two host-GPU round trips instead of 2N.

#include "device.h"
#include "host_threads.h"
#include "world.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace b2g
{

static std::mutex s_batchMutex;
static ThreadPool* s_pool = nullptr;
static Batch* s_batch = nullptr;

int StepWorlds( World** worlds, int count, float timeStep, int subStepCount, BatchStats* stats )
{
	std::lock_guard<std::mutex> guard( s_batchMutex );
	if ( s_pool == nullptr )
	{
		s_pool = new ThreadPool( HostThreads() - 1 );
		s_batch = new Batch();
	}

	float phase[5] = {};
	uint64_t t = b2GetTicks();
	auto lap = [&]( int k ) {
		phase[k] = b2GetMilliseconds( t );
		t = b2GetTicks();
	};

	std::vector<char> began( count, 0 );
	s_pool->For( count, [&]( int i ) {
		worlds[i]->device->SetRecording( true );
		began[i] = worlds[i]->StepBegin( timeStep, subStepCount ) ? 1 : 0;
	} );

	std::vector<World*> active;
	std::vector<DeviceWorld*> devices;
	for ( int i = 0; i < count; ++i )
	{
		if ( began[i] )
		{
			active.push_back( worlds[i] );
			devices.push_back( worlds[i]->device );
		}
		else
		{
			worlds[i]->device->SetRecording( false );
		}
	}
	int n = (int)active.size();
	lap( 0 );

	s_batch->Collide( devices.data(), n );
	lap( 1 );
	s_pool->For( n, [&]( int i ) { active[i]->StepMiddle(); } );
	lap( 2 );
	s_batch->Solve( devices.data(), n );
	lap( 3 );
	s_pool->For( n, [&]( int i ) {
		active[i]->StepEnd();
		active[i]->device->SetRecording( false );
	} );
	lap( 4 );

	if ( stats != nullptr )
	{
		stats->solveKernelMs = s_batch->lastFusedMs;
		stats->solveThreads = 0;
		stats->solveGrid = 0;
		int most = 0;
		for ( int k = 0; k < kBatchBuckets; ++k )
		{
			stats->bucketWorlds[k] = s_batch->lastBucketWorlds[k];
			stats->bucketThreads[k] = s_batch->lastBucketThreads[k];
			stats->bucketWarmByBody[k] = s_batch->lastBucketWarmByBody[k];
			stats->bucketGrid[k] = s_batch->lastBucketGrid[k];
			stats->bucketMs[k] = s_batch->lastBucketMs[k];
			stats->solveGrid += s_batch->lastBucketGrid[k];
			if ( s_batch->lastBucketWorlds[k] > most )
			{
				most = s_batch->lastBucketWorlds[k];
				stats->solveThreads = s_batch->lastBucketThreads[k];
			}
		}
		stats->worlds = n;
		stats->solveBytes = s_batch->lastFusedBytes;
		for ( int k = 0; k < 5; ++k )
		{
			stats->phaseMs[k] = phase[k];
		}
	}
	return n;
}

} // namespace b2g
