// Host worker threads: a fixed pool that runs index ranges.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace b2g
{

// Host threads to use: B2AMD_HOST_THREADS, else the hardware concurrency, clamped to [1, 16].
int HostThreads();

// Fixed worker threads running index ranges; the calling thread works too.
class ThreadPool
{
public:
	explicit ThreadPool( int workerCount );
	~ThreadPool();

	// Run fn(i) for i in [0, count) and return when every call has finished. Not reentrant: one
	// caller at a time.
	void For( int count, const std::function<void( int )>& fn );

	int WorkerCount() const
	{
		return (int)threads.size();
	}

private:
	void Drain();
	void Worker();

	std::vector<std::thread> threads;
	std::mutex mutex;
	std::condition_variable wake, finished;
	const std::function<void( int )>* job = nullptr;
	int jobCount = 0;
	std::atomic<int> next{ 0 };
	int pending = 0;
	uint64_t generation = 0;
	bool stop = false;
};

} // namespace b2g
