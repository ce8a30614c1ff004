// Host worker threads (see host_threads.h).

#include "host_threads.h"

#include <algorithm>
#include <cstdlib>

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

} // namespace b2g
