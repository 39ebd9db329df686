// The request loop's host logic, independent of the device: admission of
// queued requests by a byte budget, flights (requests launched together on
// one stream slot), early completion of single requests through flags the
// device raises, completion of whole flights, and bounded waits.
//
// The executor X does the device work:
//   void X::worker_start()                   first call on the worker thread
//   bool X::launch(Flight &F, int slot)      start a flight on stream slot
//                                            `slot`; sets R->flag (0..n-1)
//                                            for requests that complete on
//                                            their own; false: nothing runs
//   bool X::flag_raised(Flight &F, int k)    request flag k is up
//   bool X::flight_done(Flight &F)           all of the flight's work ended
//   bool X::finish(Flight &F, Req &R, bool failed)
//                                            completes a request from its
//                                            output (host work); its result
//   void X::release(Flight &F)               the flight's device and host
//                                            memory go back to their pools
// The loop calls these on its own worker thread only, without its lock
// held, except flag_raised and flight_done (cheap polls, under the lock).
//
// Accounting: a request's bytes are fixed when it is queued and held from
// admission to the end of its flight, so the bytes in use return to 0
// whenever no flight runs.  An idle device admits one request even above
// the budget, so every request is admitted eventually.
#ifndef IW_SERVER_H
#define IW_SERVER_H

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace iwx {

template <class X>
class RequestLoop {
public:
	struct Req {
		int id = -1;
		size_t bytes = 0;                   // device bytes held while it runs
		double t_submit = 0, t_start = 0, t_done = 0;
		int flag = -1;                      // its completion flag in the flight; -1: none
		bool done = false, ok = false;      // under the loop's lock
		bool finished = false;              // worker thread only
		typename X::ReqData data;
	};
	struct Flight {
		std::vector<Req *> reqs;
		size_t held = 0;                    // bytes admitted with it
		int slot = -1;
		bool failed = false;
		double t_launch = 0;
		typename X::FlightData data;
	};

	RequestLoop(X &x, size_t budget, int slots) : x_(x), budget_(budget)
	{
		for (int k = slots - 1; k >= 0; k--) free_slots_.push_back(k);
	}
	~RequestLoop() { stop(); }

	void start() { worker_ = std::thread([this] { loop(); }); }

	// Ends the worker once every queued request is done.
	void stop()
	{
		{
			std::lock_guard<std::mutex> lock(mu_);
			stop_ = true;
		}
		cv_work_.notify_all();
		if (worker_.joinable()) worker_.join();
	}

	// Queues R (the loop owns it from now on); returns its id.
	int submit(Req *R)
	{
		std::lock_guard<std::mutex> lock(mu_);
		R->id = next_id_++;
		R->t_submit = now();
		reqs_[R->id] = R;
		queue_.push_back(R);
		cv_work_.notify_one();
		return R->id;
	}

	// 1: done and ok, 0: done and failed or unknown id, -1: not done within
	// timeout_ms (<= 0: no limit); `why` then names what is pending.
	int wait(int id, double timeout_ms, std::string *why = nullptr)
	{
		std::unique_lock<std::mutex> lock(mu_);
		auto it = reqs_.find(id);
		if (it == reqs_.end()) return 0;
		Req *R = it->second;
		if (!wait_until(lock, timeout_ms, [&] { return R->done; })) {
			if (why) *why = pending_locked("request " + std::to_string(id));
			return -1;
		}
		return R->ok ? 1 : 0;
	}

	// As wait, for every request queued so far.
	int drain(double timeout_ms, std::string *why = nullptr)
	{
		std::unique_lock<std::mutex> lock(mu_);
		if (!wait_until(lock, timeout_ms, [&] { return queue_.empty() && flights_.empty(); })) {
			if (why) *why = pending_locked("drain");
			return -1;
		}
		int ok = 1;
		for (auto &kv : reqs_) ok &= kv.second->ok ? 1 : 0;
		return ok;
	}

	bool times(int id, double *submit, double *start, double *done)
	{
		std::lock_guard<std::mutex> lock(mu_);
		auto it = reqs_.find(id);
		if (it == reqs_.end() || !it->second->done) return false;
		if (submit) *submit = it->second->t_submit;
		if (start) *start = it->second->t_start;
		if (done) *done = it->second->t_done;
		return true;
	}

	size_t used()
	{
		std::lock_guard<std::mutex> lock(mu_);
		return used_;
	}
	size_t max_used()
	{
		std::lock_guard<std::mutex> lock(mu_);
		return max_used_;
	}
	int flights_launched()
	{
		std::lock_guard<std::mutex> lock(mu_);
		return launched_;
	}

	template <class F>
	void for_each_req(F f)
	{
		std::lock_guard<std::mutex> lock(mu_);
		for (auto &kv : reqs_) f(*kv.second);
	}

	double now() const
	{
		return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
	}

private:
	template <class P>
	bool wait_until(std::unique_lock<std::mutex> &lock, double timeout_ms, P pred)
	{
		if (timeout_ms <= 0) {
			cv_done_.wait(lock, pred);
			return true;
		}
		return cv_done_.wait_for(lock, std::chrono::duration<double, std::milli>(timeout_ms), pred);
	}

	// What is still pending, for an error message (lock held).
	std::string pending_locked(const std::string &what)
	{
		std::string s = what + " timed out: " + std::to_string(queue_.size()) + " queued, " +
			std::to_string(flights_.size()) + " flights running, " + std::to_string(used_) + " bytes held";
		const double t = now();
		for (Flight *F : flights_) {
			int open = 0;
			for (Req *R : F->reqs) open += R->done ? 0 : 1;
			char buf[160];
			std::snprintf(buf, sizeof buf, "; flight on slot %d: %d requests, %d not done, launched %.0f ms ago",
				F->slot, (int)F->reqs.size(), open, t - F->t_launch);
			s += buf;
		}
		return s;
	}

	void complete_req(Flight &F, Req &R)
	{
		const bool ok = x_.finish(F, R, F.failed);
		R.finished = true;
		const double t = now();
		std::lock_guard<std::mutex> lock(mu_);
		R.ok = ok;
		R.done = true;
		R.t_done = t;
		cv_done_.notify_all();
	}

	void loop()
	{
		x_.worker_start();
		std::unique_lock<std::mutex> lock(mu_);
		for (;;) {
			bool progress = false;
			// Requests whose own work has ended complete at once.
			for (size_t f = 0; f < flights_.size(); f++) {
				Flight *F = flights_[f];
				if (F->failed) continue;
				for (Req *R : F->reqs) {
					if (R->finished || R->flag < 0 || !x_.flag_raised(*F, R->flag)) continue;
					lock.unlock();
					complete_req(*F, *R);
					lock.lock();
					progress = true;
				}
			}
			// Ended flights complete their requests, then release their bytes,
			// memory and stream slot.  A flight stays listed until all of
			// that is done, so drain() cannot return while its requests are
			// still completing or its bytes are still counted.  Only this
			// thread changes flights_, so index k stays valid while the lock
			// is released.
			for (size_t k = 0; k < flights_.size();) {
				Flight *F = flights_[k];
				if (!x_.flight_done(*F)) {
					k++;
					continue;
				}
				lock.unlock();
				for (Req *R : F->reqs)
					if (!R->finished) complete_req(*F, *R);
				x_.release(*F);
				lock.lock();
				flights_.erase(flights_.begin() + k);
				used_ -= F->held;
				free_slots_.push_back(F->slot);
				delete F;
				progress = true;
				cv_done_.notify_all();
			}
			// Admit what fits, as one flight on one free slot.
			if (!queue_.empty() && !free_slots_.empty()) {
				Flight *F = new Flight();
				while (!queue_.empty()) {
					Req *R = queue_.front();
					const bool idle = used_ == 0 && F->reqs.empty();
					if (used_ + R->bytes > budget_ && !idle) break;
					queue_.pop_front();
					used_ += R->bytes;
					F->held += R->bytes;
					R->t_start = now();
					F->reqs.push_back(R);
				}
				if (F->reqs.empty()) {
					delete F;
				} else {
					if (used_ > max_used_) max_used_ = used_;
					F->slot = free_slots_.back();
					free_slots_.pop_back();
					F->t_launch = now();
					flights_.push_back(F);
					launched_++;
					lock.unlock();
					F->failed = !x_.launch(*F, F->slot);
					lock.lock();
					continue;
				}
			}
			if (stop_ && queue_.empty() && flights_.empty()) break;
			if (flights_.empty()) {
				// Nothing runs, so nothing is held: wait for a request.
				cv_work_.wait(lock, [&] { return stop_ || !queue_.empty(); });
			} else if (!progress) {
				cv_work_.wait_for(lock, std::chrono::microseconds(100));
			}
		}
	}

	X &x_;
	std::mutex mu_;
	std::condition_variable cv_work_, cv_done_;
	std::thread worker_;
	bool stop_ = false;
	size_t budget_ = 0, used_ = 0, max_used_ = 0;
	int launched_ = 0;
	std::vector<int> free_slots_;
	std::deque<Req *> queue_;
	std::vector<Flight *> flights_;
	std::map<int, Req *> reqs_;
	int next_id_ = 0;
	std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();

public:
	// Requests are owned by the loop; the executor frees its own data.
	void delete_requests()
	{
		std::lock_guard<std::mutex> lock(mu_);
		for (auto &kv : reqs_) delete kv.second;
		reqs_.clear();
	}
};

}  // namespace iwx

#endif
