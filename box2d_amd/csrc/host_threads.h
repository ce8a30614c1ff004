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

// One helper thread that runs one job at a time for a caller that waits for it later: a world's
// step hands it the solve's kernel launches and does its own host work meanwhile. Arm() wakes the
// thread ahead of the job so that it spins for a while and a posted job starts without a wake-up.
// TryPost() fails while another caller's job is outstanding (the caller then runs the job itself);
// the caller that posted must call Wait() before it posts again.
class StepHelper
{
public:
	StepHelper();
	~StepHelper();

	void Arm();
	bool TryPost( std::function<void()> fn );
	void Wait();

private:
	enum State : int
	{
		kIdle = 0,
		kArmed = 1,
		kPosted = 2,
		kDone = 3,
	};
	void Run();

	std::thread thread;
	std::mutex mutex;
	std::condition_variable wake;
	std::atomic<int> state{ kIdle };
	std::atomic<bool> busy{ false };
	std::function<void()> job;
	bool stop = false;
};

// The process's step helper, created on first use; null when B2AMD_LAUNCH_THREAD=0.
StepHelper* SharedStepHelper();

} // namespace b2g
