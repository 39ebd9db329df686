// The request loop (imageworsener_amd/csrc/iw_server.h) driven by a mock
// executor on the CPU, built with ThreadSanitizer by tests/test_server_loop.py.
// A mock device plays the GPU: one thread per flight ends the flight's
// requests after a delay, raising their flags (atomic stores, as the GPU
// writes coherent host memory) and then the flight's completion, while
// client threads submit, wait and drain.  Checks:
//   - every request completes, with its own result, in every scenario;
//   - a drain returns only once every request is done and every flight has
//     released its bytes;
//   - the bytes held return to 0 after each drain, and never exceed the
//     budget unless one request alone is larger (admitted on an idle device);
//   - requests whose total bytes are many times the budget all complete (the
//     held bytes are exact, so admission never starves);
//   - a flight that never ends makes wait() and drain() time out with a
//     report naming it, and the loop completes it once it ends;
//   - a flight whose launch fails, or whose completion query reports an
//     error, completes its requests as failed instead of hanging.
// Exit status 0: all checks passed.
#include "iw_server.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>

namespace {

int g_fail = 0;
#define CHECK(c)                                                                 \
	do {                                                                         \
		if (!(c)) {                                                              \
			std::fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #c); \
			g_fail++;                                                            \
		}                                                                        \
	} while (0)

struct MockExec;

// Every request the loop holds is done (what a successful drain promises).
template <class Loop>
bool all_done(Loop &loop)
{
	bool all = true;
	loop.for_each_req([&](typename Loop::Req &R) { all = all && R.done; });
	return all;
}

struct MockReq {
	int tag = 0;          // the value the request's result must carry
	int work_us = 0;      // device time of the request
	bool early = false;   // completes on its own (a flag), like error diffusion
	int *result = nullptr;
};

struct FlightState {
	std::vector<std::atomic<int>> flags;
	std::atomic<int> done{0};
	std::vector<int> out;  // "device output": written before the flags
	explicit FlightState(size_t n) : flags(n), out(n, 0)
	{
		for (auto &f : flags) f.store(0);
	}
};

struct MockFlight {
	std::shared_ptr<FlightState> st;
};

// The mock GPU: every flight runs on its own thread, so flights on different
// slots overlap, as streams do.
struct Device {
	struct Work {
		std::shared_ptr<FlightState> st;
		std::vector<int> tags, work_us;
		std::vector<bool> early;
		bool hang;
	};
	std::mutex mu;
	std::condition_variable cv;
	bool release_hung = false;
	std::vector<std::thread> threads;

	void submit(Work w)
	{
		std::lock_guard<std::mutex> lock(mu);
		threads.emplace_back([this, w] { run(w); });
	}
	void run(const Work &w)
	{
		if (w.hang) {
			// A flight that does not end until told to.
			std::unique_lock<std::mutex> lock(mu);
			cv.wait(lock, [&] { return release_hung; });
		}
		for (size_t k = 0; k < w.tags.size(); k++) {
			std::this_thread::sleep_for(std::chrono::microseconds(w.work_us[k]));
			w.st->out[k] = w.tags[k] * 7 + 1;
			if (w.early[k]) w.st->flags[k].store(1, std::memory_order_release);
		}
		w.st->done.store(1, std::memory_order_release);
	}
	void release()
	{
		{
			std::lock_guard<std::mutex> lock(mu);
			release_hung = true;
		}
		cv.notify_all();
	}
	void join()
	{
		release();
		std::vector<std::thread> t;
		{
			std::lock_guard<std::mutex> lock(mu);
			t.swap(threads);
		}
		for (auto &x : t) x.join();
	}
};

struct MockExec {
	typedef MockReq ReqData;
	typedef MockFlight FlightData;
	typedef iwx::RequestLoop<MockExec> Loop;
	Device *dev = nullptr;
	std::atomic<int> fail_next_launch{0};
	std::atomic<int> error_next_query{0};
	std::atomic<int> hang_next{0};
	std::atomic<int> in_flight{0}, max_in_flight{0};
	int slots = 0;
	std::vector<std::atomic<int>> slot_busy;
	explicit MockExec(int n) : slots(n), slot_busy(n)
	{
		for (auto &s : slot_busy) s.store(0);
	}

	void worker_start() {}
	bool launch(Loop::Flight &F, int slot)
	{
		CHECK(slot >= 0 && slot < slots);
		CHECK(slot_busy[slot].exchange(1) == 0);  // one flight per slot
		const int n = ++in_flight;
		int m = max_in_flight.load();
		while (n > m && !max_in_flight.compare_exchange_weak(m, n)) {
		}
		F.data.st = std::make_shared<FlightState>(F.reqs.size());
		if (fail_next_launch.exchange(0)) {
			F.data.st->done.store(1);
			return false;
		}
		Device::Work w;
		w.st = F.data.st;
		w.hang = hang_next.exchange(0) != 0;
		for (size_t k = 0; k < F.reqs.size(); k++) {
			w.tags.push_back(F.reqs[k]->data.tag);
			w.work_us.push_back(F.reqs[k]->data.work_us);
			w.early.push_back(F.reqs[k]->data.early);
			if (F.reqs[k]->data.early) F.reqs[k]->flag = (int)k;
		}
		dev->submit(std::move(w));
		return true;
	}
	bool flag_raised(Loop::Flight &F, int k) { return F.data.st->flags[k].load(std::memory_order_acquire) == 1; }
	bool flight_done(Loop::Flight &F)
	{
		if (F.data.st->done.load(std::memory_order_acquire) != 1) return false;
		if (error_next_query.exchange(0)) F.failed = true;
		return true;
	}
	bool finish(Loop::Flight &F, Loop::Req &R, bool failed)
	{
		if (failed) return false;
		size_t k = 0;
		while (k < F.reqs.size() && F.reqs[k] != &R) k++;
		CHECK(k < F.reqs.size());
		*R.data.result = F.data.st->out[k];
		return *R.data.result == R.data.tag * 7 + 1;
	}
	void release(Loop::Flight &F)
	{
		CHECK(F.data.st->done.load() == 1);
		CHECK(slot_busy[F.slot].exchange(0) == 1);
		--in_flight;
	}
};

struct Setup {
	Device dev;
	MockExec x;
	MockExec::Loop loop;
	std::vector<int> results;
	Setup(size_t budget, int slots, size_t nreq) : x(slots), loop(x, budget, slots), results(nreq, -1)
	{
		x.dev = &dev;
		loop.start();
	}
	~Setup()
	{
		dev.release();
		loop.stop();
		loop.delete_requests();
		dev.join();
	}
	MockExec::Loop::Req *req(int tag, size_t bytes, int work_us, bool early)
	{
		auto *R = new MockExec::Loop::Req();
		R->bytes = bytes;
		R->data.tag = tag;
		R->data.work_us = work_us;
		R->data.early = early;
		R->data.result = &results[tag];
		return R;
	}
};

// Many requests of random sizes from several client threads; some larger
// than the budget on their own.
void scenario_mixed()
{
	const size_t budget = 4000;
	const int n = 400, clients = 3;
	Setup s(budget, 4, n);
	std::vector<int> ids(n, -1);
	std::vector<std::thread> th;
	for (int c = 0; c < clients; c++) {
		th.emplace_back([&, c] {
			std::mt19937 rng(100 + c);
			for (int i = c; i < n; i += clients) {
				const size_t bytes = (i % 97 == 0) ? 6000 : 1 + rng() % 400;
				ids[i] = s.loop.submit(s.req(i, bytes, (int)(rng() % 300), (rng() % 2) == 0));
				if (rng() % 8 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 500));
			}
		});
	}
	for (auto &t : th) t.join();
	std::string why;
	for (int i = 0; i < n; i++) CHECK(s.loop.wait(ids[i], 60000, &why) == 1);
	CHECK(s.loop.drain(60000, &why) == 1); CHECK(all_done(s.loop));
	for (int i = 0; i < n; i++) CHECK(s.results[i] == i * 7 + 1);
	CHECK(s.loop.used() == 0);
	CHECK(s.loop.max_used() <= 6000 + budget);
	CHECK(s.x.max_in_flight.load() <= 4);
	std::printf("mixed: %d requests, %d flights, max held %zu of %zu, max %d flights at once\n", n,
		s.loop.flights_launched(), s.loop.max_used(), budget, s.x.max_in_flight.load());
}

// Requests arriving while others run start on free slots: flights overlap.
void scenario_overlap()
{
	const int n = 40;
	Setup s(1 << 20, 4, n);
	for (int i = 0; i < n; i++) {
		s.loop.submit(s.req(i, 100, 3000, i % 2 == 0));
		std::this_thread::sleep_for(std::chrono::microseconds(300));
	}
	std::string why;
	CHECK(s.loop.drain(60000, &why) == 1); CHECK(all_done(s.loop));
	for (int i = 0; i < n; i++) CHECK(s.results[i] == i * 7 + 1);
	CHECK(s.x.max_in_flight.load() >= 2);
	CHECK(s.x.max_in_flight.load() <= 4);
	CHECK(s.loop.used() == 0);
	std::printf("overlap: %d requests, %d flights, max %d flights at once\n", n, s.loop.flights_launched(),
		s.x.max_in_flight.load());
}

// Total bytes far above the budget, in several waves of submit + drain on one
// loop: the held bytes must return to 0 after every wave.
void scenario_budget_turnover()
{
	const size_t budget = 100;
	const int waves = 20, per = 30;
	Setup s(budget, 2, waves * per);
	std::string why;
	for (int w = 0; w < waves; w++) {
		for (int i = 0; i < per; i++) s.loop.submit(s.req(w * per + i, 60, 20, i % 3 == 0));
		CHECK(s.loop.drain(30000, &why) == 1); CHECK(all_done(s.loop));
		CHECK(s.loop.used() == 0);
	}
	for (int i = 0; i < waves * per; i++) CHECK(s.results[i] == i * 7 + 1);
	std::printf("turnover: %d requests of 60 bytes through a budget of 100, %d flights\n", waves * per,
		s.loop.flights_launched());
}

// A flight that never ends: waits time out and name it; once the device
// finishes, everything completes.
void scenario_timeout()
{
	Setup s(1000, 2, 4);
	s.x.hang_next = 1;
	const int a = s.loop.submit(s.req(0, 10, 10, true));
	std::string why;
	CHECK(s.loop.wait(a, 200, &why) == -1);
	CHECK(why.find("flight on slot") != std::string::npos);
	CHECK(s.loop.drain(100, &why) == -1);
	std::printf("timeout report: %s\n", why.c_str());
	const int b = s.loop.submit(s.req(1, 10, 10, false));  // runs on the other slot
	CHECK(s.loop.wait(b, 30000, &why) == 1);
	s.dev.release();
	CHECK(s.loop.wait(a, 30000, &why) == 1);
	CHECK(s.loop.drain(30000, &why) == 1); CHECK(all_done(s.loop));
	CHECK(s.loop.used() == 0);
}

// A failed launch and a failed completion query end their requests as
// failed; later requests still run.
void scenario_failures()
{
	Setup s(1000, 1, 6);
	std::string why;
	s.x.fail_next_launch = 1;
	const int a = s.loop.submit(s.req(0, 10, 10, true));
	CHECK(s.loop.wait(a, 30000, &why) == 0);
	s.x.error_next_query = 1;
	const int b = s.loop.submit(s.req(1, 10, 10, false));
	CHECK(s.loop.wait(b, 30000, &why) == 0);
	const int c = s.loop.submit(s.req(2, 10, 10, true));
	const int d = s.loop.submit(s.req(3, 10, 10, false));
	CHECK(s.loop.wait(c, 30000, &why) == 1);
	CHECK(s.loop.wait(d, 30000, &why) == 1);
	CHECK(s.loop.drain(30000, &why) == 0);  // a and b failed
	CHECK(s.loop.used() == 0);
}

}  // namespace

int main()
{
	scenario_mixed();
	scenario_overlap();
	scenario_budget_turnover();
	scenario_timeout();
	scenario_failures();
	if (g_fail) {
		std::fprintf(stderr, "%d checks failed\n", g_fail);
		return 1;
	}
	std::printf("request loop: all checks passed\n");
	return 0;
}
