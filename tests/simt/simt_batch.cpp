// Runs the batched kernel's device code (simplify_patch) on the CPU: one fiber per GPU thread,
// rendezvous at every barrier and wavefront operation (shim/hip/hip_runtime.h).  A deadlock (a
// barrier or wavefront operation that not every thread of its group reaches) is reported with the
// source line each fiber waits at.  Test infrastructure only.
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

#include <memory>
#include <vector>

#include "batch_device.h"

namespace simt
{

enum State
{
	kRunnable,
	kAtBarrier,
	kAtWave,
	kDone
};

struct Fiber
{
	ucontext_t ctx;
	Dim idx;
	State state;
	int line; // source line of the last rendezvous
	unsigned long long syncs;
	std::vector<char> stack;
};

constexpr int kMaxGroups = 8;

struct Group
{
	int id = 0;
	int threads = 0;
	std::vector<Fiber> fibers;
	int current = -1;
	ucontext_t scheduler;
	// workgroup barrier
	int barrier_count = 0;
	// wavefront rendezvous, double-buffered by generation
	std::vector<int> wave_count;
	std::vector<unsigned long long> wave_gen;
	std::vector<uint64_t> wave_buf; // [waves][2][64]
};

// the workgroup whose fibers run now (several may be in flight, interleaved sweep by sweep)
static Group g_groups[kMaxGroups];
static Group* gp = &g_groups[0];
#define g (*gp)
static volatile sig_atomic_t g_watchdog = 0;

const Dim& thread_idx()
{
	return g.fibers[g.current].idx;
}

unsigned int lane()
{
	return g.fibers[g.current].idx.x & 63;
}

static void park(State st, int line)
{
	Fiber& f = g.fibers[g.current];
	f.state = st;
	f.line = line;
	f.syncs++;
	swapcontext(&f.ctx, &g.scheduler);
}

void block_barrier(int line)
{
	g.barrier_count++;
	if (g.barrier_count == g.threads)
	{
		g.barrier_count = 0;
		for (auto& f : g.fibers)
			if (f.state == kAtBarrier)
				f.state = kRunnable;
		Fiber& me = g.fibers[g.current];
		me.line = line;
		me.syncs++;
		return;
	}
	park(kAtBarrier, line);
}

const uint64_t* wave_sync(uint64_t v, int line)
{
	const int w = (int)(g.fibers[g.current].idx.x >> 6);
	const int members = (g.threads - w * 64) < 64 ? (g.threads - w * 64) : 64;
	const unsigned long long gen = g.wave_gen[w];
	uint64_t* buf = &g.wave_buf[((size_t)w * 2 + (gen & 1)) * 64];
	buf[lane()] = v;
	g.wave_count[w]++;
	if (g.wave_count[w] == members)
	{
		g.wave_count[w] = 0;
		g.wave_gen[w]++;
		for (int l = 0; l < members; ++l)
		{
			Fiber& f = g.fibers[w * 64 + l];
			if (f.state == kAtWave)
				f.state = kRunnable;
		}
		Fiber& me = g.fibers[g.current];
		me.line = line;
		me.syncs++;
	}
	else
		park(kAtWave, line);
	return buf;
}

// what a workgroup runs: one mesh (simplify_patch), or meshes from a shared cursor
// (simplify_from_cursor, the loop of the largest class's kernel)
struct Task
{
	const mhip::BatchArgs* a;
	const mhip::PatchLayout* L;
	uint32_t patch; // one-mesh task
	const uint32_t* ids; // cursor task: the list, its length and the shared cursor
	uint32_t n;
	uint32_t* cursor;
	uint8_t* smem;
	uint32_t* part;
	uint32_t* sh;
	uint64_t* box;
	float* shf;
	uint32_t* next;
};

static Task g_tasks[kMaxGroups];

static void fiber_main()
{
	const Task& t = g_tasks[g.id];
	const bool attr = t.a->attribute_count != 0;
	if (t.ids)
	{
		if (attr)
			mhip::simplify_from_cursor<true>(*t.a, *t.L, t.ids, t.n, t.cursor, t.smem, t.part, t.sh, t.box, t.shf, t.next);
		else
			mhip::simplify_from_cursor<false>(*t.a, *t.L, t.ids, t.n, t.cursor, t.smem, t.part, t.sh, t.box, t.shf, t.next);
	}
	else if (attr)
		mhip::simplify_patch<true>(*t.a, *t.L, t.patch, t.smem, t.part, t.sh, t.box, t.shf);
	else
		mhip::simplify_patch<false>(*t.a, *t.L, t.patch, t.smem, t.part, t.sh, t.box, t.shf);
	g.fibers[g.current].state = kDone;
	swapcontext(&g.fibers[g.current].ctx, &g.scheduler);
}

static void report(const char* why)
{
	fprintf(stderr, "simt: %s (workgroup %d, patch %u)\n", why, g.id, g_tasks[g.id].patch);
	for (int t = 0; t < g.threads; ++t)
	{
		const Fiber& f = g.fibers[t];
		if (t < 4 || t % 64 == 0 || f.state != g.fibers[0].state || f.line != g.fibers[0].line)
			fprintf(stderr, "  thread %3d state %d line %d syncs %llu%s\n", t, (int)f.state, f.line, f.syncs, t == g.current ? " (running)" : "");
	}
}

static void on_alarm(int)
{
	g_watchdog = 1;
	report("watchdog: no completion within the time limit");
	_exit(3);
}

static void on_segv(int)
{
	report("access fault (past the LDS image or elsewhere)");
	_exit(4);
}

// the LDS image ends exactly at an inaccessible page, so any access past the layout faults
struct GuardedImage
{
	uint8_t* base = nullptr;
	size_t mapped = 0;
	uint8_t* data = nullptr;

	explicit GuardedImage(size_t bytes)
	{
		const size_t page = (size_t)sysconf(_SC_PAGESIZE);
		const size_t body = (bytes + page - 1) / page * page;
		mapped = body + page;
		base = (uint8_t*)mmap(nullptr, mapped, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
		if (base == MAP_FAILED)
			abort();
		mprotect(base + body, page, PROT_NONE);
		data = base + body - bytes;
	}
	~GuardedImage()
	{
		munmap(base, mapped);
	}
};

static uint64_t g_rng = 0;
static unsigned int g_retries = 0;

static uint64_t next_random()
{
	g_rng ^= g_rng << 13;
	g_rng ^= g_rng >> 7;
	g_rng ^= g_rng << 17;
	return g_rng;
}

static void init_group(Group& grp, int id, int threads)
{
	grp.id = id;
	grp.threads = threads;
	grp.fibers.resize(threads);
	const int waves = (threads + 63) / 64;
	grp.wave_count.assign(waves, 0);
	grp.wave_gen.assign(waves, 0);
	grp.wave_buf.assign((size_t)waves * 2 * 64, 0);
	grp.barrier_count = 0;
	for (int t = 0; t < threads; ++t)
	{
		Fiber& f = grp.fibers[t];
		f.idx = Dim{(unsigned int)t, 0, 0};
		f.state = kRunnable;
		f.line = 0;
		f.syncs = 0;
		f.stack.resize(256 * 1024);
		getcontext(&f.ctx);
		f.ctx.uc_stack.ss_sp = f.stack.data();
		f.ctx.uc_stack.ss_size = f.stack.size();
		f.ctx.uc_link = nullptr;
		makecontext(&f.ctx, fiber_main, 0);
	}
}

// one scheduling sweep over the current group's runnable fibers (in a fresh random order with
// shuffle set); sets done when every fiber has finished, returns whether any fiber ran
static bool sweep(std::vector<int>& order, bool shuffle, bool& done)
{
	const int threads = g.threads;
	bool progressed = false, all_done = true;
	if (shuffle)
		for (int t = threads - 1; t > 0; --t)
		{
			int u = (int)(next_random() % (uint64_t)(t + 1));
			int x = order[t];
			order[t] = order[u];
			order[u] = x;
		}
	for (int i = 0; i < threads; ++i)
	{
		const int t = order[i];
		Fiber& f = g.fibers[t];
		if (f.state != kDone)
			all_done = false;
		if (f.state != kRunnable)
			continue;
		g.current = t;
		swapcontext(&g.scheduler, &f.ctx);
		progressed = true;
	}
	done = all_done;
	return progressed;
}

// runs `groups` workgroups of `threads` fibers each (their tasks in g_tasks), interleaved: every
// round visits the unfinished groups in a random order (with shuffle set) and gives each one
// sweep, so the groups' atomics on shared words interleave in varying orders.  Returns 0, or -1
// on a deadlock within a group.
static int run_groups(int groups, int threads, bool shuffle)
{
	std::vector<std::vector<int>> orders(groups, std::vector<int>(threads));
	std::vector<char> finished(groups, 0);
	for (int k = 0; k < groups; ++k)
	{
		init_group(g_groups[k], k, threads);
		for (int t = 0; t < threads; ++t)
			orders[k][t] = t;
	}
	std::vector<int> visit(groups);
	for (int k = 0; k < groups; ++k)
		visit[k] = k;
	for (;;)
	{
		if (shuffle)
			for (int k = groups - 1; k > 0; --k)
			{
				int u = (int)(next_random() % (uint64_t)(k + 1));
				int x = visit[k];
				visit[k] = visit[u];
				visit[u] = x;
			}
		bool any_left = false;
		for (int k : visit)
		{
			if (finished[k])
				continue;
			gp = &g_groups[k];
			bool done = false;
			const bool progressed = sweep(orders[k], shuffle, done);
			if (done)
			{
				finished[k] = 1;
				continue;
			}
			any_left = true;
			if (!progressed)
			{
				report("deadlock: no thread can proceed");
				return -1;
			}
		}
		if (!any_left)
			return 0;
	}
}

static int run_group(int threads, bool shuffle)
{
	return run_groups(1, threads, shuffle);
}

} // namespace simt

// Simplifies every mesh of a packed batch with the kernel's device code, mesh by mesh, using the
// launch's LDS layout for (vmax, imax).  status[p] receives the kernel's status word.
// vertex_lock of the next simt_simplify_batch call (one byte per packed vertex, or null)
static const unsigned char* g_vertex_lock = nullptr;

extern "C" void simt_set_vertex_lock(const unsigned char* lock)
{
	g_vertex_lock = lock;
}

// attributes of the next calls (packed per vertex like the positions, stride in bytes, one weight
// per attribute; count 0: none)
static const float* g_attributes = nullptr;
static size_t g_attribute_stride = 0;
static std::vector<float> g_weights;

extern "C" void simt_set_attributes(const float* attributes, size_t stride, const float* weights, size_t count)
{
	g_attributes = count ? attributes : nullptr;
	g_attribute_stride = stride;
	g_weights.assign(weights, weights + count);
}

// the attribute fields of a batch's arguments, as the host sets them (the attributes with weight
// > 0), with their global scratch in `store` (undefined at the start, like a fresh allocation)
struct AttrStore
{
	std::vector<float> values;
	std::vector<mhip::Quad> quadrics;
	std::vector<mhip::Grad> grads;
};

static void set_attributes(mhip::BatchArgs& a, AttrStore& store)
{
	uint32_t used = 0;
	for (size_t k = 0; k < g_weights.size(); ++k)
		if (g_weights[k] > 0)
		{
			a.attr_map.index[used] = (uint32_t)k;
			a.attr_map.weight[used] = g_weights[k];
			used++;
		}
	a.attribute_count = g_attributes ? used : 0;
	if (!a.attribute_count)
		return;
	a.attributes = g_attributes;
	a.attribute_stride_floats = g_attribute_stride / sizeof(float);
	const size_t n = a.vertex_count + 1;
	store.values.resize(n * used);
	store.quadrics.resize(n);
	store.grads.resize(n * used);
	for (auto& v : store.values)
		v = (float)(simt::next_random() % 1000) - 500.f;
	for (auto& q : store.quadrics)
		q.a00 = q.a11 = q.a22 = q.a10 = q.a20 = q.a21 = q.b0 = q.b1 = q.b2 = q.c = q.w = q.pad = (float)(simt::next_random() % 1000) - 500.f;
	for (auto& gr : store.grads)
		gr.gx = gr.gy = gr.gz = gr.gw = (float)(simt::next_random() % 1000) - 500.f;
	a.attr_values = store.values.data();
	a.attr_quadrics = store.quadrics.data();
	a.attr_grads = store.grads.data();
}

extern "C" int simt_simplify_batch(unsigned int* destination, unsigned int* result_counts, float* result_errors, unsigned int* status,
                                   unsigned int* stats, const unsigned int* indices, const unsigned int* index_offsets, const float* positions,
                                   const unsigned int* vertex_offsets, size_t stride, const unsigned int* targets, size_t count,
                                   float target_error, unsigned int options, unsigned int vmax, unsigned int imax, int watchdog_s,
                                   unsigned int mode, unsigned long long seed)
{
	// mode bit 0: fill the LDS image with random bytes before every mesh (LDS starts undefined);
	// mode bit 1: random fiber order in every scheduling sweep
	simt::g_rng = seed * 0x9E3779B97F4A7C15ull + 1;
	mhip::BatchArgs a = {};
	a.destination = destination;
	a.result_counts = result_counts;
	a.result_errors = result_errors;
	a.status = status;
	a.stats = stats;
	a.indices = indices;
	a.index_offsets = index_offsets;
	a.positions = positions;
	a.vertex_offsets = vertex_offsets;
	a.vertex_lock = g_vertex_lock;
	a.vertex_count = count ? vertex_offsets[count] : 0;
	// the quadric buffer in global memory, undefined at the start like a fresh device allocation
	std::vector<mhip::Quad> quadrics(a.vertex_count + 1);
	for (auto& q : quadrics)
		q.a00 = q.a11 = q.a22 = q.a10 = q.a20 = q.a21 = q.b0 = q.b1 = q.b2 = q.c = q.w = q.pad = (float)(simt::next_random() % 1000) - 500.f;
	a.quadrics = quadrics.data();
	a.position_stride_floats = stride / sizeof(float);
	a.target_index_counts = targets;
	a.patch_count = count;
	a.target_error = target_error;
	a.options = options;
	AttrStore attr_store;
	set_attributes(a, attr_store);
	// the host's two layouts for this size: capped candidate arrays first, full arrays for a
	// mesh that leaves with status 5; each LDS image ends at its own guard page
	const mhip::PatchLayout layouts[2] = {mhip::patch_layout(vmax, imax, mhip::batch_candidate_cap(imax)), mhip::patch_layout(vmax, imax, imax + 4)};
	simt::GuardedImage image0(layouts[0].total), image1(layouts[1].total);
	uint8_t* images[2] = {image0.data, image1.data};
	uint32_t part[mhip::kBatchThreads / 64];
	uint32_t sh[16];
	uint64_t box[6];
	float shf[8];
	signal(SIGSEGV, simt::on_segv);
	signal(SIGBUS, simt::on_segv);
	if (watchdog_s > 0)
	{
		signal(SIGALRM, simt::on_alarm);
		alarm((unsigned int)watchdog_s);
	}
	int rc = 0;
	simt::g_retries = 0;
	for (size_t p = 0; p < count && rc == 0; ++p)
	{
		for (int attempt = 0; attempt < 2 && rc == 0; ++attempt)
		{
			const mhip::PatchLayout& L = layouts[attempt];
			uint8_t* smem = images[attempt];
			if (mode & 1)
			{
				for (size_t b = 0; b < L.total; ++b)
					smem[b] = (uint8_t)simt::next_random();
				for (auto& w : sh)
					w = (uint32_t)simt::next_random();
				for (auto& w : part)
					w = (uint32_t)simt::next_random();
			}
			simt::g_tasks[0] = simt::Task{&a, &L, (uint32_t)p, nullptr, 0, nullptr, smem, part, sh, box, shf, nullptr};
			rc = simt::run_group(mhip::kBatchThreads, (mode & 2) != 0);
			if (status[p] != 5)
				break;
			simt::g_retries++;
		}
	}
	if (watchdog_s > 0)
		alarm(0);
	return rc;
}

extern "C" unsigned int simt_layout_bytes(unsigned int vmax, unsigned int imax, unsigned int cmax)
{
	return mhip::patch_layout(vmax, imax, cmax).total;
}

// meshes the last simt_simplify_batch call ran again with full candidate arrays
extern "C" unsigned int simt_retries()
{
	return simt::g_retries;
}

// The largest class's wrapper (simplify_from_cursor): `groups` workgroups, each with its own LDS
// image (full candidate arrays, as that class holds them; random bytes at the start, then whatever
// the workgroup's previous mesh left), take the meshes ids[0..count) from one shared cursor,
// interleaved sweep by sweep in random orders.  Every mesh must be taken exactly once (status
// words start at 99) and the cursor must end at count + groups (each workgroup's last take is the
// one that finds the list empty).  Returns 0, -1 on a deadlock, -2 if the cursor is off.
extern "C" int simt_simplify_batch_cursor(unsigned int* destination, unsigned int* result_counts, float* result_errors, unsigned int* status,
                                          const unsigned int* ids, const unsigned int* indices, const unsigned int* index_offsets,
                                          const float* positions, const unsigned int* vertex_offsets, size_t stride, const unsigned int* targets,
                                          size_t count, float target_error, unsigned int options, unsigned int vmax, unsigned int imax,
                                          int groups, int watchdog_s, unsigned long long seed)
{
	if (groups < 1 || groups > simt::kMaxGroups)
		return -3;
	simt::g_rng = seed * 0x9E3779B97F4A7C15ull + 1;
	mhip::BatchArgs a = {};
	a.destination = destination;
	a.result_counts = result_counts;
	a.result_errors = result_errors;
	a.status = status;
	a.stats = nullptr;
	a.indices = indices;
	a.index_offsets = index_offsets;
	a.positions = positions;
	a.vertex_offsets = vertex_offsets;
	a.vertex_lock = g_vertex_lock;
	a.vertex_count = count ? vertex_offsets[count] : 0;
	std::vector<mhip::Quad> quadrics(a.vertex_count + 1);
	for (auto& q : quadrics)
		q.a00 = q.a11 = q.a22 = q.a10 = q.a20 = q.a21 = q.b0 = q.b1 = q.b2 = q.c = q.w = q.pad = (float)(simt::next_random() % 1000) - 500.f;
	a.quadrics = quadrics.data();
	a.position_stride_floats = stride / sizeof(float);
	a.target_index_counts = targets;
	a.patch_count = count;
	a.target_error = target_error;
	a.options = options;
	AttrStore attr_store;
	set_attributes(a, attr_store);
	const mhip::PatchLayout L = mhip::patch_layout(vmax, imax, imax + 4);
	std::vector<std::unique_ptr<simt::GuardedImage>> images;
	struct Shared
	{
		uint32_t part[mhip::kBatchThreads / 64];
		uint32_t sh[16];
		uint64_t box[6];
		float shf[8];
		uint32_t next;
	};
	std::vector<Shared> shared(groups);
	uint32_t cursor = 0;
	for (int k = 0; k < groups; ++k)
	{
		images.emplace_back(new simt::GuardedImage(L.total));
		uint8_t* smem = images.back()->data;
		for (size_t b = 0; b < L.total; ++b)
			smem[b] = (uint8_t)simt::next_random();
		Shared& sd = shared[k];
		for (auto& w : sd.sh)
			w = (uint32_t)simt::next_random();
		for (auto& w : sd.part)
			w = (uint32_t)simt::next_random();
		sd.next = (uint32_t)simt::next_random();
		simt::g_tasks[k] = simt::Task{&a, &L, 0xffffffffu, ids, (uint32_t)count, &cursor, smem, sd.part, sd.sh, sd.box, sd.shf, &sd.next};
	}
	signal(SIGSEGV, simt::on_segv);
	signal(SIGBUS, simt::on_segv);
	if (watchdog_s > 0)
	{
		signal(SIGALRM, simt::on_alarm);
		alarm((unsigned int)watchdog_s);
	}
	int rc = simt::run_groups(groups, mhip::kBatchThreads, true);
	if (watchdog_s > 0)
		alarm(0);
	if (rc == 0 && cursor != (uint32_t)count + (uint32_t)groups)
		rc = -2;
	return rc;
}
