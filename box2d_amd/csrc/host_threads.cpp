// Host worker threads (see host_threads.h).

#include "host_threads.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <immintrin.h>

namespace b2g
{

int HostThreads()
{
	const char* v = getenv( "B2AMD_HOST_THREADS" );
	int n = v != nullptr ? atoi( v ) : (int)std::thread::hardware_concurrency();
	return std::max( 1, std::min( n, 16 ) );
}

ThreadPool::ThreadPool( int workerCount )
{
	for ( int i = 0; i < workerCount; ++i )
	{
		threads.emplace_back( [this]() { Worker(); } );
	}
}

ThreadPool::~ThreadPool()
{
	{
		std::lock_guard<std::mutex> lock( mutex );
		stop = true;
	}
	wake.notify_all();
	for ( std::thread& t : threads )
	{
		t.join();
	}
}

void ThreadPool::For( int count, const std::function<void( int )>& fn )
{
	{
		std::lock_guard<std::mutex> lock( mutex );
		job = &fn;
		jobCount = count;
		next.store( 0 );
		pending = (int)threads.size();
		generation += 1;
	}
	wake.notify_all();
	Drain();
	std::unique_lock<std::mutex> lock( mutex );
	finished.wait( lock, [this]() { return pending == 0; } );
	job = nullptr;
}

void ThreadPool::Drain()
{
	for ( ;; )
	{
		int i = next.fetch_add( 1 );
		if ( i >= jobCount )
		{
			return;
		}
		( *job )( i );
	}
}

void ThreadPool::Worker()
{
	uint64_t seen = 0;
	for ( ;; )
	{
		{
			std::unique_lock<std::mutex> lock( mutex );
			wake.wait( lock, [&]() { return stop || generation != seen; } );
			if ( stop )
			{
				return;
			}
			seen = generation;
		}
		Drain();
		{
			std::lock_guard<std::mutex> lock( mutex );
			pending -= 1;
		}
		finished.notify_one();
	}
}

StepHelper::StepHelper()
{
	thread = std::thread( [this]() { Run(); } );
}

StepHelper::~StepHelper()
{
	{
		std::lock_guard<std::mutex> lock( mutex );
		stop = true;
	}
	wake.notify_all();
	thread.join();
}

void StepHelper::Arm()
{
	if ( busy.load( std::memory_order_acquire ) )
	{
		return;
	}
	int expected = kIdle;
	if ( state.compare_exchange_strong( expected, kArmed ) )
	{
		{
			std::lock_guard<std::mutex> lock( mutex );
		}
		wake.notify_one();
	}
}

bool StepHelper::TryPost( std::function<void()> fn )
{
	if ( busy.exchange( true, std::memory_order_acq_rel ) )
	{
		return false;
	}
	job = std::move( fn );
	{
		std::lock_guard<std::mutex> lock( mutex );
		// idle or armed (the thread may drop from armed to idle meanwhile; either way it sees this)
		state.store( kPosted, std::memory_order_release );
	}
	wake.notify_one();
	return true;
}

void StepHelper::Wait()
{
	while ( state.load( std::memory_order_acquire ) != kDone )
	{
		_mm_pause();
	}
	job = nullptr;
	state.store( kIdle, std::memory_order_release );
	busy.store( false, std::memory_order_release );
}

void StepHelper::Run()
{
	// how long an armed thread spins for its job before it sleeps again
	const auto spinLimit = std::chrono::milliseconds( 4 );
	for ( ;; )
	{
		{
			std::unique_lock<std::mutex> lock( mutex );
			wake.wait( lock, [this]() {
				int s = state.load( std::memory_order_acquire );
				return stop || s == kArmed || s == kPosted;
			} );
			if ( stop )
			{
				return;
			}
		}
		auto start = std::chrono::steady_clock::now();
		int spins = 0;
		while ( state.load( std::memory_order_acquire ) == kArmed )
		{
			_mm_pause();
			if ( ( ++spins & 1023 ) == 0 && std::chrono::steady_clock::now() - start > spinLimit )
			{
				int expected = kArmed;
				state.compare_exchange_strong( expected, kIdle );
			}
		}
		if ( state.load( std::memory_order_acquire ) == kPosted )
		{
			job();
			state.store( kDone, std::memory_order_release );
		}
	}
}

StepHelper* SharedStepHelper()
{
	static StepHelper* helper = []() -> StepHelper* {
		const char* v = getenv( "B2AMD_LAUNCH_THREAD" );
		if ( v != nullptr && atoi( v ) == 0 )
		{
			return nullptr;
		}
		// never destroyed: its thread may outlive the worlds of a process that exits without
		// destroying them
		return new StepHelper();
	}();
	return helper;
}

} // namespace b2g
