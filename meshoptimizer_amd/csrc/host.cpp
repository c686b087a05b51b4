// Host orchestrator and C ABI of libmeshopt_hip.so.
//
// One simplify call = the setup kernels, then the pass loop of src/simplifier.cpp:2493-2541
// run as kernel sequences on one HIP stream, with the per-pass loop conditions read back
// through a pinned scalar block.  All state lives in one device arena per context.
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/meshopt_hip.h"
#include "../../include/meshoptimizer.h"
#include "batch.h"
#include "clod.h"
#include "sloppy.h"
#include "common.h"
#include "kernels.h"
#include "prims.h"

namespace mhip
{

struct HipError : std::runtime_error
{
	using std::runtime_error::runtime_error;
};

void fatal_hip(hipError_t err, const char* expr, const char* file, int line)
{
	char buf[512];
	snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", int(err), hipGetErrorString(err), expr, file, line);
	throw HipError(buf);
}

// ------------------------------------------------------------------------------------------
// arena: bump allocation of 256-byte aligned sub-buffers; a dry run measures the footprint

struct Arena
{
	char* base = nullptr;
	size_t size = 0;
	size_t used = 0;

	template <typename T>
	T* take(size_t count)
	{
		size_t off = (used + 255) & ~size_t(255);
		used = off + count * sizeof(T) + 16;
		return base ? reinterpret_cast<T*>(base + off) : nullptr;
	}
};

struct PassRecord
{
	meshopt_hip_pass_stats s;
};

// device words of a batched call's lists: the status of every mesh, one count word per class,
// the class id lists and their ordering's scratch, the largest class's cursor
static size_t batch_list_words(size_t meshes)
{
	return meshes + 4 + 2 * (size_t)kBatchClasses * meshes + kBatchClasses + 1;
}

// host words of a batched call: the status read-back (every mesh, then the class counts), the list
// of meshes run again in the retry class and its length
static size_t batch_host_words(size_t meshes)
{
	return 2 * meshes + kBatchClasses + 1;
}

// device buffer for the host entry points' input/output staging (grows; released explicitly)
struct StagingCache
{
	void* buf = nullptr;
	size_t size = 0;
	void* get(size_t bytes)
	{
		if (bytes > size)
		{
			if (buf)
				HIP_CHECK(hipFree(buf));
			buf = nullptr;
			size = 0;
			HIP_CHECK(hipMalloc(&buf, bytes + bytes / 4 + 256));
			size = bytes + bytes / 4 + 256;
		}
		return buf;
	}
	void release()
	{
		if (buf)
			(void)hipFree(buf);
		buf = nullptr;
		size = 0;
	}
};

struct Context
{
	std::mutex mutex;
	int device = -1;
	hipStream_t stream = nullptr;
	char* arena = nullptr;
	size_t arena_size = 0;
	StagingCache staging;
	StagingCache batch_attr; // batched calls with attributes: attribute quadrics, gradients and values
	CallScalars* sc_host = nullptr; // pinned mirror of the device scalar block
	std::vector<meshopt_hip_pass_stats> passes;
	meshopt_hip_call_stats call{};
	// round-kernel (k_round_x) timing of this context's calls (meshopt_hip_kernel_timing), and the events of
	// the batched calls
	bool timing = false;
	std::vector<hipEvent_t> ev_begin, ev_end;
	hipEvent_t batch_ev[2] = {nullptr, nullptr};
	double window_fraction = 0; // evaluated / candidates of the previous pass (0: none yet)
	uint32_t* batch_dev = nullptr;    // batched calls: status, class counts and class id lists, then 4 statistics words per mesh
	size_t batch_dev_words = 0;
	uint32_t* batch_pinned = nullptr; // batched calls: status read-back, then the retry list
	uint32_t* batch_pinned_dev = nullptr; // the same memory's device address (kernels write the status there)
	size_t batch_pinned_words = 0;
	size_t batch_last_meshes = 0;     // meshes of the last batched call (their statistics follow the lists)
	Quad* batch_quadrics = nullptr; // batched calls: one vertex quadric per vertex of the batch
	size_t batch_quadric_count = 0;
	uint32_t* sloppy_words = nullptr; // sloppy calls: pinned read-back words

	uint32_t* sloppy_pinned()
	{
		if (!sloppy_words)
			HIP_CHECK(hipHostMalloc((void**)&sloppy_words, 8 * sizeof(uint32_t), hipHostMallocDefault));
		return sloppy_words;
	}
	// the batched calls' size classes: the launch shape (set through meshopt_hip_batch_set_shape, or
	// read from the tuning table at the first batched call) and the caps it gives on this device
	BatchShape batch_shape = batch_default_shape();
	bool batch_shape_set = false;
	BatchClasses batch_cls{};
	bool batch_cls_ready = false;
	size_t cu_count = 0;
	hipStream_t aux[kBatchClasses - 1] = {}; // batch size classes beyond the first
	hipEvent_t fork_ev = nullptr, join_ev[kBatchClasses - 1] = {};

	void ensure_batch_buffers(size_t meshes)
	{
		// status, class counts, id lists, their ordering's scratch, the largest class's cursor, and
		// the per-mesh statistics
		size_t words = batch_list_words(meshes) + 4 * meshes;
		if (words > batch_dev_words)
		{
			if (batch_dev)
				HIP_CHECK(hipFree(batch_dev));
			batch_dev = nullptr;
			batch_dev_words = 0;
			HIP_CHECK(hipMalloc((void**)&batch_dev, words * sizeof(uint32_t)));
			batch_dev_words = words;
		}
		// the status of every mesh, then the class counts (they follow it on the device too), then
		// the retry list
		if (batch_host_words(meshes) > batch_pinned_words)
		{
			if (batch_pinned)
				HIP_CHECK(hipHostFree(batch_pinned));
			batch_pinned = nullptr;
			batch_pinned_words = 0;
			HIP_CHECK(hipHostMalloc((void**)&batch_pinned, batch_host_words(meshes) * sizeof(uint32_t), hipHostMallocDefault));
			HIP_CHECK(hipHostGetDevicePointer((void**)&batch_pinned_dev, batch_pinned, 0));
			batch_pinned_words = batch_host_words(meshes);
		}
	}

	void ensure_batch_quadrics(size_t vertices)
	{
		if (vertices <= batch_quadric_count)
			return;
		if (batch_quadrics)
			HIP_CHECK(hipFree(batch_quadrics));
		batch_quadrics = nullptr;
		batch_quadric_count = 0;
		size_t grow = vertices + vertices / 4;
		HIP_CHECK(hipMalloc((void**)&batch_quadrics, grow * sizeof(Quad)));
		batch_quadric_count = grow;
	}

	// one clusterlod level's packing buffers (meshopt_hip_clod_level_device)
	void* clod_buf = nullptr;
	size_t clod_bytes = 0;
	void* ensure_clod(size_t bytes)
	{
		if (bytes > clod_bytes)
		{
			if (clod_buf)
				HIP_CHECK(hipFree(clod_buf));
			clod_buf = nullptr;
			clod_bytes = 0;
			size_t grow = bytes + bytes / 4;
			HIP_CHECK(hipMalloc(&clod_buf, grow));
			clod_bytes = grow;
		}
		return clod_buf;
	}

	void ensure_aux_streams()
	{
		if (fork_ev)
			return;
		for (int k = 0; k < kBatchClasses - 1; ++k)
		{
			HIP_CHECK(hipStreamCreateWithFlags(&aux[k], hipStreamNonBlocking));
			HIP_CHECK(hipEventCreateWithFlags(&join_ev[k], hipEventDisableTiming));
		}
		HIP_CHECK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
	}

	// the context's device and stream (the allocation-free entry needs nothing more of it)
	void ensure_stream()
	{
		if (stream)
			return;
		HIP_CHECK(hipGetDevice(&device));
		HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
	}

	void ensure_init()
	{
		ensure_stream();
		if (!sc_host)
			HIP_CHECK(hipHostMalloc((void**)&sc_host, sizeof(CallScalars), hipHostMallocDefault));
	}

	void ensure_arena(size_t bytes)
	{
		if (bytes <= arena_size)
			return;
		if (arena)
			HIP_CHECK(hipFree(arena));
		arena = nullptr;
		arena_size = 0;
		size_t grow = bytes + bytes / 4;
		HIP_CHECK(hipMalloc((void**)&arena, grow));
		arena_size = grow;
	}

	// device memory held between calls (arena and staging)
	size_t held_bytes() const
	{
		return arena_size + staging.size;
	}

	void release_memory()
	{
		if (arena)
			(void)hipFree(arena);
		arena = nullptr;
		arena_size = 0;
		staging.release();
	}

	void release()
	{
		if (batch_dev)
			(void)hipFree(batch_dev);
		if (batch_pinned)
			(void)hipHostFree(batch_pinned);
		if (batch_quadrics)
			(void)hipFree(batch_quadrics);
		if (sloppy_words)
			(void)hipHostFree(sloppy_words);
		sloppy_words = nullptr;
		batch_dev = nullptr;
		batch_pinned = nullptr;
		batch_quadrics = nullptr;
		batch_dev_words = batch_pinned_words = batch_quadric_count = 0;
		for (int k = 0; k < kBatchClasses - 1; ++k)
		{
			if (aux[k])
				(void)hipStreamDestroy(aux[k]);
			if (join_ev[k])
				(void)hipEventDestroy(join_ev[k]);
			aux[k] = nullptr;
			join_ev[k] = nullptr;
		}
		if (fork_ev)
			(void)hipEventDestroy(fork_ev);
		fork_ev = nullptr;
		staging.release();
		batch_attr.release();
		if (arena)
			(void)hipFree(arena);
		if (sc_host)
			(void)hipHostFree(sc_host);
		if (stream)
			(void)hipStreamDestroy(stream);
		arena = nullptr;
		arena_size = 0;
		sc_host = nullptr;
		stream = nullptr;
	}
};

static Context& context()
{
	static Context ctx;
	return ctx;
}

// Contexts of the single-mesh entry points: a call takes an idle context (a new one when every
// context is busy) and hands it back when it returns, so calls from several threads run on their
// own streams and arenas and overlap on the device, as the reference's calls may run concurrently.
struct ContextPool
{
	std::mutex m;
	std::vector<std::unique_ptr<Context>> all;
	std::vector<Context*> idle;

	Context* acquire()
	{
		std::lock_guard<std::mutex> guard(m);
		if (idle.empty())
		{
			all.emplace_back(new Context());
			return all.back().get();
		}
		Context* c = idle.back();
		idle.pop_back();
		return c;
	}

	void release(Context* c)
	{
		std::lock_guard<std::mutex> guard(m);
		idle.push_back(c);
	}

	// frees the device memory and streams of the idle contexts
	void release_idle()
	{
		std::lock_guard<std::mutex> guard(m);
		for (Context* c : idle)
		{
			std::lock_guard<std::mutex> g2(c->mutex);
			c->release();
		}
	}
};

static ContextPool& pool()
{
	static ContextPool p;
	return p;
}

struct PooledContext
{
	Context* ctx;
	PooledContext()
	    : ctx(pool().acquire())
	{
	}
	~PooledContext()
	{
		pool().release(ctx);
	}
	PooledContext(const PooledContext&) = delete;
	PooledContext& operator=(const PooledContext&) = delete;
};

// kernel timing (meshopt_hip_kernel_timing, meshopt_hip_batch_kernel_timing): on or off for
// every context, and the totals gathered since the last query
struct TimingTotals
{
	std::mutex m;
	bool on = false;
	double x_ms = 0;
	size_t x_launches = 0, x_visits = 0;
	double batch_ms = 0;
	size_t batch_launches = 0, batch_meshes = 0;
};

static TimingTotals& timing_totals()
{
	static TimingTotals t;
	return t;
}

static bool timing_on()
{
	TimingTotals& t = timing_totals();
	std::lock_guard<std::mutex> guard(t.m);
	return t.on;
}

static thread_local std::string t_last_error;
static thread_local std::vector<meshopt_hip_pass_stats> t_passes;
static thread_local meshopt_hip_call_stats t_call{};

// ------------------------------------------------------------------------------------------
// buffer plan of one call

struct Plan
{
	Mesh m;
	uint32_t* sp_first;
	uint32_t* sp_flags;
	uint32_t* sp_dense;
	uint32_t* sparse_remap;
	uint32_t* table;
	size_t table_size;
	uint32_t* v_flags;
	uint32_t* v_a;
	uint32_t* v_b;
	uint32_t* v_c;
	uint32_t* v_d;
	uint32_t* raw_off;
	uint32_t* raw_keys;
	uint32_t* raw_corners;
	uint32_t* w_off;
	uint32_t* w_keys;
	uint32_t* w_corners;
	uint32_t* i_alt0;
	uint32_t* i_alt1;
	uint32_t* kind_counts;
	uint32_t* comp; // Prune: component per vertex
	float* cerr;    // Prune: error per component
	SolveBuffers solve; // meshopt_simplifyWithUpdate only
};

static size_t pow2_at_least(size_t n)
{
	size_t p = 1;
	while (p < n)
		p <<= 1;
	return p;
}

static void carve(Arena& a, Plan& p, size_t V_in, size_t V, size_t I, size_t A, bool sparse, bool prune, bool solve)
{
	Mesh& m = p.m;
	size_t T = I / 3 + 1;
	size_t cap = I + 4;
	size_t Vp = V + 1;

	m.ib[0] = a.take<uint32_t>(I + 4);
	m.ib[1] = a.take<uint32_t>(I + 4);
	p.sp_first = sparse ? a.take<uint32_t>(V_in + 1) : nullptr;
	p.sp_flags = sparse ? a.take<uint32_t>(I + 1) : nullptr;
	p.sp_dense = sparse ? a.take<uint32_t>(I + 1) : nullptr;
	p.sparse_remap = sparse ? a.take<uint32_t>(I + 1) : nullptr;

	m.praw = a.take<float4>(Vp);
	m.pos = a.take<float4>(Vp);
	m.remap = a.take<uint32_t>(Vp);
	m.wedge = a.take<uint32_t>(Vp);
	m.kind = a.take<uint8_t>(Vp);
	for (int k = 0; k < 2; ++k)
	{
		m.loop[k] = a.take<uint32_t>(Vp);
		m.loopback[k] = a.take<uint32_t>(Vp);
	}
	m.vq = a.take<Quad>(Vp);
	m.vg = (A && solve) ? a.take<Grad>(Vp) : nullptr;
	m.aq = A ? a.take<Quad>(Vp) : nullptr;
	m.ag = A ? a.take<Grad>(Vp * A) : nullptr;
	m.va = A ? a.take<float>(Vp * A) : nullptr;
	m.cremap = a.take<uint32_t>(Vp);
	m.src_off = a.take<uint32_t>(Vp + 1);
	m.lock_rank = a.take<uint32_t>(Vp);
	for (int k = 0; k < 2; ++k)
	{
		m.pend[k] = a.take<uint64_t>(2 * Vp);
	}
	m.adj_off = a.take<uint32_t>(Vp + 1);
	m.adj = a.take<uint2>(I + 4);

	m.tri_cnt = a.take<uint32_t>(T);
	m.tri_off = a.take<uint32_t>(T);

	m.c_v0 = a.take<uint32_t>(cap);
	m.c_v1 = a.take<uint32_t>(cap);
	m.c_err = a.take<float>(cap);
	m.keys = a.take<uint32_t>(cap);
	m.vals = a.take<uint32_t>(cap);
	m.keys_alt = a.take<uint32_t>(cap);
	m.vals_alt = a.take<uint32_t>(cap);
	m.r_v0 = a.take<uint32_t>(cap);
	m.r_v1 = a.take<uint32_t>(cap);
	m.r_r0 = a.take<uint32_t>(cap);
	m.r_r1 = a.take<uint32_t>(cap);
	m.r_err = a.take<float>(cap);
	m.src_rank = a.take<uint32_t>(cap);
	m.src_val = a.take<uint32_t>(cap);
	m.status = a.take<uint8_t>(cap);
	m.list[0] = a.take<uint4>(cap);
	m.list[1] = a.take<uint4>(cap);
	m.stop_tile = a.take<uint32_t>(3 * (cap / kScanTile + 2));

	// setup-only buffers alias the candidate arrays (not live at the same time)
	p.table_size = pow2_at_least(2 * V + 2);
	p.table = a.take<uint32_t>(p.table_size);
	p.v_flags = a.take<uint32_t>(Vp);
	p.v_a = a.take<uint32_t>(Vp);
	p.v_b = a.take<uint32_t>(Vp);
	p.v_c = a.take<uint32_t>(Vp);
	p.v_d = a.take<uint32_t>(Vp);
	p.raw_off = a.take<uint32_t>(Vp + 1);
	p.raw_keys = m.r_v0;
	p.raw_corners = m.r_v1;
	p.w_off = a.take<uint32_t>(Vp + 1);
	p.w_keys = m.r_r0;
	p.w_corners = m.r_r1;
	p.i_alt0 = m.keys_alt;
	p.i_alt1 = m.vals_alt;
	p.kind_counts = a.take<uint32_t>(8);
	p.comp = prune ? a.take<uint32_t>(Vp) : nullptr;
	p.cerr = prune ? a.take<float>(Vp) : nullptr;
	p.solve.update = solve ? a.take<uint8_t>(Vp + 8) : nullptr;
	p.solve.listed = solve ? a.take<uint32_t>(Vp) : nullptr;
	p.solve.waiting = solve ? a.take<uint32_t>(Vp) : nullptr;
	p.solve.queue[0] = solve ? a.take<uint32_t>(Vp) : nullptr;
	p.solve.queue[1] = solve ? a.take<uint32_t>(Vp) : nullptr;
	p.solve.cand = solve ? a.take<float4>(Vp) : nullptr;
	p.solve.state = solve ? a.take<uint8_t>(Vp + 8) : nullptr;

	size_t biggest = std::max(std::max(I, cap), Vp + 1);
	size_t temp_words = std::max(sort_temp_words(biggest), scan_temp_words(biggest + 16));
	m.temp = a.take<uint32_t>(temp_words + 64);
	m.sc = a.take<CallScalars>(1);
}

// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_count_kinds(const uint32_t* remap, const uint8_t* kind, size_t n, uint32_t* counts)
{
	__shared__ uint32_t h[6];
	if (threadIdx.x < 6)
		h[threadIdx.x] = 0;
	__syncthreads();
	uint32_t local[6] = {0, 0, 0, 0, 0, 0};
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		if (remap[i] == (uint32_t)i)
		{
			uint8_t k = kind[i];
			local[k < 5 ? k : 4]++;
			local[5]++;
		}
	for (int q = 0; q < 6; ++q)
	{
		uint32_t v = local[q];
		for (int o = 32; o > 0; o >>= 1)
			v += __shfl_xor(v, o, 64);
		if (__lane_id() == 0 && v)
			atomicAdd(&h[q], v);
	}
	__syncthreads();
	if (threadIdx.x < 6 && h[threadIdx.x])
		atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_scalars_init(CallScalars* sc)
{
	sc->result_error = 0.f;
	sc->vertex_error = 0.f;
	sc->vertex_error_bits = 0u;
	sc->candidates = 0;
	sc->dual_count = 0;
	sc->work[2] = 0;
}

static float host_float_bits(uint32_t u)
{
	float f;
	memcpy(&f, &u, sizeof(f));
	return f;
}

// candidate lists up to this length are decided in one rank window (no extension phase): a
// longer list's window holds the ranks its stop is expected within, which spares the rounds the
// ranks past it (grid 256: 2.76 ms per call with 131072 here, 2.56 ms with 65536)
constexpr size_t kWholeWindow = 65536;
// list length the single-workgroup tail of the commit rounds takes in one sweep
constexpr uint32_t kTailList = 1024;
// list length at which the tail takes over from the head rounds: a tail round costs about 5 us
// plus 22 ns per listed rank (one workgroup), a head round 5-8 us at any short length
constexpr uint32_t kTailStart = 256;
// head rounds of a window seen for the first time
constexpr int kFirstHeads = 16;
// a window whose learned head-round count exceeds g_deep_heads has long chains of dependent ranks
// (as on large regular grids) and runs deep rounds: its head rounds from g_deep_from_head on, and
// its tail, examine the alternatives of moving corners (meshopt_hip_set_round_policy sets both)
constexpr int kDeepHeads = 64;
constexpr int kDeepFromHead = 8;
static std::atomic<int> g_deep_heads{kDeepHeads}, g_deep_from_head{kDeepFromHead};
// head rounds of the commit, learned per window size (by its power of two) from every call in the
// process: the rounds a list needs grow with its length, and the contexts of concurrent calls and
// serving slots see calls of very different sizes (0: not seen yet)
static std::atomic<int> g_head_rounds[33];
// the same for the extension window [W, C) of a pass, learned only from extensions that opened
// (most never do: each head of an extension that does not open is an empty launch)
static std::atomic<int> g_ext_head_rounds[33];

static int size_bucket(size_t n)
{
	int bucket = 0;
	while (bucket < 32 && (size_t(1) << bucket) < n)
		bucket++;
	return bucket;
}

// read-backs of the calling thread's current call (meshopt_hip_call_stats::read_backs)
static thread_local unsigned int t_read_backs;

static void sync_scalars(Context& ctx, const Mesh& m)
{
	HIP_CHECK(hipMemcpyAsync(ctx.sc_host, m.sc, sizeof(CallScalars), hipMemcpyDeviceToHost, ctx.stream));
	HIP_CHECK(hipStreamSynchronize(ctx.stream));
	t_read_backs++;
}

struct CallArgs
{
	unsigned int* destination; // device
	const unsigned int* indices; // device
	size_t index_count;
	const float* positions; // device
	size_t vertex_count;
	size_t position_stride;
	const float* attributes; // device
	size_t attribute_stride;
	const float* weights; // host
	size_t attribute_count;
	float* out_positions;  // device, meshopt_simplifyWithUpdate: positions written back (same layout as positions)
	float* out_attributes; // device, the same for the attributes
	const unsigned char* lock; // device
	size_t target_index_count;
	float target_error;
	unsigned int options;
};

// the whole simplify call on device buffers; returns the index count, sets *result_error_sq
static size_t simplify_on_device(Context& ctx, const CallArgs& c, float* result_error_out)
{
	hipStream_t s = ctx.stream;
	const bool sparse = (c.options & kOptSparse) != 0;
	const bool prune = (c.options & kOptPrune) != 0;
	const bool solve = (c.options & kOptInternalSolve) != 0;
	const size_t I = c.index_count;
	if (solve && (!c.out_positions || (c.attribute_count && !c.out_attributes)))
		throw std::invalid_argument("the solve needs writable positions (and attributes)");
	const size_t V_in = c.vertex_count;

	uint32_t amap[kMaxAttributes];
	float aw[kMaxAttributes];
	size_t A = 0;
	for (size_t k = 0; k < c.attribute_count && k < (size_t)kMaxAttributes; ++k)
		if (c.weights[k] > 0)
		{
			amap[A] = (uint32_t)k;
			aw[A] = c.weights[k];
			A++;
		}

	size_t V_alloc = sparse ? std::min(V_in, I) : V_in;

	Arena dry;
	Plan plan;
	carve(dry, plan, V_in, V_alloc, I, A, sparse, prune, solve);
	ctx.ensure_arena(dry.used + 4096);
	Arena arena;
	arena.base = ctx.arena;
	arena.size = ctx.arena_size;
	carve(arena, plan, V_in, V_alloc, I, A, sparse, prune, solve);
	Mesh& m = plan.m;

	m.index_count = (uint32_t)I;
	m.attr_count = (uint32_t)A;
	KernelTimer timer{};
	ctx.timing = timing_on();
	if (ctx.timing)
	{
		if (ctx.ev_begin.empty())
		{
			ctx.ev_begin.resize(4096);
			ctx.ev_end.resize(4096);
			for (size_t i = 0; i < ctx.ev_begin.size(); ++i)
			{
				HIP_CHECK(hipEventCreate(&ctx.ev_begin[i]));
				HIP_CHECK(hipEventCreate(&ctx.ev_end[i]));
			}
		}
		timer.begin = ctx.ev_begin.data();
		timer.end = ctx.ev_end.data();
		timer.capacity = (int)ctx.ev_begin.size();
		timer.count = 0;
		m.timer = &timer;
	}
	else
		m.timer = nullptr;
	m.ib_cur = 0;
	m.loop_cur = 0;
	m.epoch = 0;

	t_read_backs = 0;
	HIP_CHECK(hipMemsetAsync(m.sc, 0, sizeof(CallScalars), s));
	k_scalars_init<<<1, 1, 0, s>>>(m.sc);
	if (I)
	{
		HIP_CHECK(hipMemcpyAsync(m.ib[0], c.indices, I * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
		// every later gather trusts the indices: reject out-of-range ones before any kernel uses them
		launch_check_indices(s, m.ib[0], I, V_in, &m.sc->work[2]);
		sync_scalars(ctx, m);
		if (ctx.sc_host->work[2])
			throw std::invalid_argument("index buffer references a vertex >= vertex_count");
	}

	size_t V = V_in;
	const uint32_t* sparse_remap = nullptr;
	if (sparse)
	{
		launch_sparse_remap(s, m.ib[0], m.ib[0], I, V_in, plan.sp_first, plan.sp_flags, plan.sp_dense, plan.sparse_remap, &m.sc->unique_count, m.temp);
		sync_scalars(ctx, m);
		V = I ? ctx.sc_host->unique_count : 0;
		sparse_remap = plan.sparse_remap;
	}
	m.vertex_count = (uint32_t)V;

	launch_gather_positions(s, c.positions, c.position_stride / sizeof(float), sparse_remap, m.praw, V);
	launch_position_remap(s, m, plan.table, pow2_at_least(2 * V + 2));
	launch_wedges(s, m, plan.v_flags, plan.v_a, plan.v_b, plan.v_c, plan.v_d);

	launch_corner_csr(s, m.ib[0], I, nullptr, V, plan.raw_off, plan.raw_keys, plan.raw_corners, plan.i_alt0, plan.i_alt1, m.temp);
	launch_classify(s, m, plan.raw_off, plan.raw_corners, c.lock, sparse_remap, c.options);

	launch_rescale(s, m);
	if (A)
		launch_rescale_attributes(s, m, c.attributes, c.attribute_stride / sizeof(float), sparse_remap, amap, aw);
	if (prune)
	{
		// buildComponents / measureComponents (src/simplifier.cpp:2457-2472) on the input
		// index buffer and the rescaled positions; the wedge scratch is free again
		uint32_t* scratch[5] = {plan.v_flags, plan.v_a, plan.v_b, plan.v_c, plan.v_d};
		launch_components(s, m, I, plan.comp, scratch);
		launch_measure_components(s, m, plan.comp, plan.cerr, scratch);
	}

	launch_corner_csr(s, m.ib[0], I, m.remap, V, plan.w_off, plan.w_keys, plan.w_corners, plan.i_alt0, plan.i_alt1, m.temp);
	HIP_CHECK(hipMemsetAsync(m.vq, 0, (V + 1) * sizeof(Quad), s));
	launch_vertex_quadrics(s, m, plan.w_off, plan.w_corners, c.lock, sparse_remap, c.options);
	if (A)
		launch_attribute_quadrics(s, m, plan.raw_off, plan.raw_corners);

	// attribute quadrics read the raw corner list; everything after may reuse its space
	launch_dual_count(s, m, plan.raw_off);
	HIP_CHECK(hipMemsetAsync(plan.kind_counts, 0, 8 * sizeof(uint32_t), s));
	if (V)
		k_count_kinds<<<std::min(div_up(V, 256 * 8), 512u), 256, 0, s>>>(m.remap, m.kind, V, plan.kind_counts);
	uint32_t kind_counts[8];
	HIP_CHECK(hipMemcpyAsync(kind_counts, plan.kind_counts, sizeof(kind_counts), hipMemcpyDeviceToHost, s));

	// round minima start as "no entry" (tag of epoch 0, rank all-ones)
	for (int b = 0; b < 2; ++b)
	{
		HIP_CHECK(hipMemsetAsync(m.pend[b], 0xff, 2 * (V + 1) * sizeof(uint64_t), s));
	}

	// the first pass's adjacency and pick go before the set-up read-back, which returns their count
	launch_pick_capacity(s, m, I);
	if (!prune && I)
	{
		launch_adjacency(s, m, I / 3);
		launch_pick(s, m, I / 3);
	}

	sync_scalars(ctx, m);
	const CallScalars& h = *ctx.sc_host;

	ctx.window_fraction = 0;
	t_call = meshopt_hip_call_stats{};
	t_call.vertex_count = (unsigned int)V;
	t_call.unique_positions = kind_counts[5];
	for (int k = 0; k < 5; ++k)
		t_call.kinds[k] = kind_counts[k];
	t_passes.clear();

	m.capacity = h.pick_capacity;

	float vertex_scale = h.extent;
	float error_scale = (c.options & kOptErrorAbsolute) ? vertex_scale : 1.f;
	float error_limit = (c.target_error * c.target_error) / (error_scale * error_scale);

	size_t result_count = I;
	float result_error_sq = 0.f;
	// Prune: smallest component error still present (src/simplifier.cpp:2466-2470)
	float prune_next = prune ? host_float_bits(h.prune_min_bits) : 0.f;

	// Outside Prune a pass makes one read-back: when the stop scan of the first rank window finds
	// no break, the window is extended on the device to the whole list (its rounds and stop scan
	// are enqueued and leave at once otherwise), and the next pass's adjacency and pick follow the
	// index remap, reading the new triangle count on the device.  Prune reads the pass's error back
	// before its component pruning, and picks after it.
	bool picked = !prune;
	// set once a window of this call needed more rounds than the cheap ones drain (long chains of
	// dependent ranks, as on regular grids): later windows of the call run deep rounds
	bool long_chains = false;
	while (result_count > c.target_index_count)
	{
		size_t T = result_count / 3;
		if (!picked)
		{
			launch_adjacency(s, m, T);
			launch_pick(s, m, T);
			sync_scalars(ctx, m);
		}
		picked = false;
		size_t C = h.candidates;
		if (C == 0)
			break;

		launch_rank_sort(s, m, C);
		uint32_t goal = (uint32_t)((result_count - c.target_index_count) / 3);

		// rank window: the first pass guesses from the triangle goal, later passes from the
		// fraction of candidates the previous pass evaluated; a short list is one window
		size_t W = ctx.window_fraction > 0 ? (size_t)(ctx.window_fraction * (double)C * 1.25) : (size_t)goal * 3 / 4;
		W = C <= kWholeWindow ? C : std::min(C, std::max<size_t>(W, 32768));

		// a window the tail takes in one sweep per round needs no head rounds
		const int bucket = size_bucket(W);
		const int learned = g_head_rounds[bucket].load(std::memory_order_relaxed);
		const int heads = W <= kTailStart ? 0 : (learned ? learned : kFirstHeads);
		// the flip test of the first rounds waits for every moving corner (cheap, and enough for
		// most meshes); a window known to need long chains of rounds examines the moving corners'
		// alternatives from round kDeepFromHead on, which decides ranks far earlier on such meshes
		const int deep_heads = g_deep_heads.load(std::memory_order_relaxed), deep_from = g_deep_from_head.load(std::memory_order_relaxed);
		bool deep = long_chains || learned > deep_heads || deep_heads < 0;
		launch_perform_init(s, m, C, deep);
		// head rounds with one thread per rank of the window (the list only shrinks; workgroups
		// past its current length leave at once, while a grid smaller than the list would make
		// each thread walk several ranks' dependent gathers in turn), then the tail and the
		// flip-failure resolution; their grid strides over any list length, so a grid capped at
		// `max_threads` stays correct
		auto window_rounds = [&](size_t lo, size_t hi, int head_base, int head_count, size_t max_threads) {
			for (int hd = 0; hd < head_count; ++hd)
				launch_round(s, m, std::min(hi - lo, max_threads), head_base + hd, hd, deep && hd >= deep_from);
			launch_round_tail(s, m, head_count, (uint32_t)(hi - lo + 2), deep);
			launch_resolve(s, m, lo, hi);
		};
		launch_window_extend(s, m, 0, W, false);
		int heads_run = heads;
		bool switched_deep = false;
		if (!learned && W > kWholeWindow && heads)
		{
			// a window size not seen yet: head rounds in doubling batches, each followed by a
			// read-back of the list length, until the list is short enough for the tail (a first
			// guess too small would hand the tail a list of hundreds of thousands of ranks).  A list
			// still long after deep_heads rounds has long chains: deep rounds from then on, with the
			// per-source lists they read built first
			int done = 0, batch = heads;
			for (;;)
			{
				for (int k = 0; k < batch && done < kMaxHeadRounds; ++k, ++done)
					launch_round(s, m, W, done, done, deep && done >= deep_from);
				sync_scalars(ctx, m);
				uint32_t left = 0;
				for (int q = 0; q < kListSegments; ++q)
					left += h.list_count[(h.round + done) % 3][q];
				if (left < kTailStart || done >= kMaxHeadRounds)
					break;
				if (!deep && deep_heads >= 0 && done >= deep_heads && left >= kTailList)
				{
					launch_source_lists(s, m, C);
					deep = switched_deep = true;
				}
				batch = done;
			}
			heads_run = done;
			launch_round_tail(s, m, done, (uint32_t)(W + 2), deep);
			launch_resolve(s, m, 0, W);
		}
		else
			window_rounds(0, W, 0, heads, (size_t)-1);
		launch_stop_scan(s, m, W, C, goal, error_limit, false);
		const int ext_bucket = size_bucket(C - W);
		const int ext_learned = g_ext_head_rounds[ext_bucket].load(std::memory_order_relaxed);
		const int ext_heads = std::max(2, std::min(ext_learned ? ext_learned : 2, kHeadSizeSlots - heads_run));
		if (W < C)
		{
			// the rest of the list, decided only when the window held no break (otherwise every
			// kernel below leaves at once: a few small grids)
			launch_window_extend(s, m, W, C, true);
			window_rounds(W, C, heads_run, ext_heads, (size_t)64 * 1024);
			launch_stop_scan(s, m, C, C, goal, error_limit, true);
		}
		launch_stop_commit(s, m, C, goal, error_limit);
		// a pass without commits changes nothing below: the loop then ends on the read-back
		launch_apply(s, m, C);
		launch_remap_triangles(s, m, T);
		if (!prune)
		{
			launch_adjacency(s, m, T, &m.sc->next_triangles);
			launch_pick(s, m, T, &m.sc->next_triangles);
			picked = true;
		}
		sync_scalars(ctx, m);
		if (h.work[3])
			throw std::runtime_error("priority rounds did not converge (internal error)");

		// head rounds for the next window of this size: enough to bring the list under kTailStart
		// (when it was not reached, extrapolated from the last two head rounds' decay, at most
		// twice as many), an overshoot given back an eighth at a time, so that windows of one size
		// with different needs (passes of one mesh, meshes of a serving stream) settle near the
		// larger need rather than alternate: a head too many costs about what the tail round it
		// replaces costs, a head too few hands the tail a list it sweeps one workgroup at a time.
		// `base`: the first head's slot in head_sizes
		auto learn_heads = [&](int base, int count) {
			int need = -1;
			for (int hd = 0; hd < count; ++hd)
				if (h.head_sizes[base + hd] < kTailStart)
				{
					need = hd;
					break;
				}
			if (need < 0)
			{
				need = count * 2;
				const double a = h.head_sizes[base + count - 2], b = h.head_sizes[base + count - 1];
				if (b > 0 && b < a)
					need = std::min(need, count + 1 + (int)std::ceil(std::log((double)kTailStart / b) / std::log(b / a)));
			}
			const int next = std::max(need, count - count / 8);
			return std::max(2, std::min(kMaxHeadRounds, next));
		};
		if (heads_run)
			g_head_rounds[bucket].store(learn_heads(0, heads_run), std::memory_order_relaxed);
		if (W < C)
		{
			if (h.extended)
				g_ext_head_rounds[ext_bucket].store(learn_heads(heads_run, ext_heads), std::memory_order_relaxed);
			else if (ext_learned > 2)
				g_ext_head_rounds[ext_bucket].store(ext_learned - ext_learned / 8, std::memory_order_relaxed);
		}
		long_chains |= switched_deep || (heads_run >= deep_heads && h.head_sizes[heads_run - 1] >= kTailList);
		ctx.window_fraction = std::min(1.0, (double)(h.stop + 1) / (double)C);

		meshopt_hip_pass_stats ps{};
		ps.live_triangles = (unsigned int)T;
		ps.candidates = (unsigned int)C;
		ps.stop = h.stop_reason;
		ps.evaluated = h.stop_reason ? h.stop + 1 : h.stop;
		ps.committed = h.committed;
		ps.skipped = h.skipped;
		ps.rejected = h.rejected;
		ps.rounds = h.round;
		ps.tail_rounds = h.tail_rounds;
		t_passes.push_back(ps);
		t_call.device_extensions += h.extended;

		if (h.committed == 0)
		{
			// the index remap and loop update ran before the read-back; without a commit the
			// reference stops before them (src/simplifier.cpp:2523-2524), and the buffers they read
			// are untouched (both are double-buffered), so those stay current
			m.ib_cur ^= 1;
			m.loop_cur ^= 1;
			break;
		}
		result_count = (size_t)h.next_triangles * 3;

		// components whose error fell under the pass's error (src/simplifier.cpp:2539-2541)
		if (prune && result_count > c.target_index_count && prune_next <= h.vertex_error)
		{
			launch_prune_triangles(s, m, result_count / 3, plan.comp, plan.cerr, h.vertex_error);
			sync_scalars(ctx, m);
			result_count = (size_t)h.next_triangles * 3;
			prune_next = host_float_bits(h.prune_min_bits);
		}
	}

	// the last read-back (every pass ends with one, after its stop scan) holds the result error
	result_error_sq = h.result_error;
	float vertex_error_sq = h.vertex_error;

	// remaining components up to the error limit (src/simplifier.cpp:2545-2569)
	bool prune_stale = true;
	while (prune && result_count > c.target_index_count && prune_next <= error_limit)
	{
		float cutoff = prune_next * 1.5f < error_limit ? prune_next * 1.5f : error_limit;
		launch_component_max(s, m, plan.cerr, cutoff);
		launch_prune_triangles(s, m, result_count / 3, plan.comp, plan.cerr, cutoff);
		sync_scalars(ctx, m);
		float cmax = host_float_bits(h.prune_max_bits);
		size_t new_count = (size_t)h.next_triangles * 3;
		prune_next = host_float_bits(h.prune_min_bits);
		if (new_count == result_count && !prune_stale)
			break;
		prune_stale = false;
		result_count = new_count;
		result_error_sq = result_error_sq < cmax ? cmax : result_error_sq;
		vertex_error_sq = vertex_error_sq < cmax ? cmax : vertex_error_sq;
	}

	// meshopt_simplifyWithUpdate: move the surviving vertices and write them back
	// (src/simplifier.cpp:2575-2599)
	if (solve)
	{
		launch_solve_setup(s, m, plan.solve, result_count);
		sync_scalars(ctx, m);
		const uint32_t listed = h.work[4];
		uint32_t round = 0, batch = 4;
		// Kahn's order: the queue empties only once every listed primary is solved, and every
		// round solves at least the lowest unsolved one
		while (h.work[5 + round % 3] != 0)
		{
			if (round > listed)
				throw std::runtime_error("position solve did not converge (internal error)");
			launch_solve_rounds(s, m, plan.solve, listed, round, batch);
			round += batch;
			batch = std::min<uint32_t>(batch * 2, 64);
			sync_scalars(ctx, m);
		}
		t_call.solve_vertices = listed;
		t_call.solve_rounds = round;

		SolveOutput o;
		o.positions = c.out_positions;
		o.position_stride = c.position_stride / sizeof(float);
		o.attributes = c.out_attributes;
		o.attribute_stride = c.attribute_stride / sizeof(float);
		o.vertex_lock = c.lock;
		o.sparse_remap = sparse_remap;
		for (size_t k = 0; k < A; ++k)
		{
			o.map.index[k] = amap[k];
			o.map.weight[k] = aw[k];
		}
		launch_solve_finish(s, m, plan.solve, o);
	}

	if ((c.options & kOptInternalDebug) && !sparse)
		launch_debug_bits(s, m, result_count);
	if (sparse)
		launch_sparse_backmap(s, m.ib[m.ib_cur], result_count, sparse_remap);
	if (result_count)
		HIP_CHECK(hipMemcpyAsync(c.destination, m.ib[m.ib_cur], result_count * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
	sync_scalars(ctx, m);

	if (ctx.timing)
	{
		double total = 0;
		for (int i = 0; i < timer.count; ++i)
		{
			float ms = 0;
			HIP_CHECK(hipEventElapsedTime(&ms, timer.begin[i], timer.end[i]));
			total += ms;
		}
		TimingTotals& t = timing_totals();
		std::lock_guard<std::mutex> guard(t.m);
		t.x_ms += total;
		t.x_launches += (size_t)timer.count;
		t.x_visits += h.x_visits;
	}

	t_call.passes = (unsigned int)t_passes.size();
	t_call.read_backs = t_read_backs;
	t_call.result_error_sq = result_error_sq;
	t_call.vertex_error_sq = vertex_error_sq;

	if (result_error_out)
		*result_error_out = sqrtf(result_error_sq) * error_scale;
	return result_count;
}

// ------------------------------------------------------------------------------------------
// host-pointer entry: stage the used spans of the caller's buffers in device memory

struct DeviceStaging
{
	std::vector<void*> ptrs;
	~DeviceStaging()
	{
		for (void* p : ptrs)
			(void)hipFree(p);
	}
	void* alloc(size_t bytes)
	{
		void* p = nullptr;
		HIP_CHECK(hipMalloc(&p, bytes ? bytes : 16));
		ptrs.push_back(p);
		return p;
	}
};


// Byte offsets of the staged streams of one host call inside the context's staging buffer.
struct HostStaging
{
	bool attr_inside; // attributes interleaved in the position stream: staged with it
	size_t pos_bytes, attr_span, lock_bytes;
	size_t off_idx, off_dst, off_pos, off_attr, off_lock, total;
};

static HostStaging plan_staging(size_t index_count, const float* positions, size_t vertex_count, size_t position_stride, const float* attributes,
                                size_t attribute_stride, size_t attribute_count, const unsigned char* lock)
{
	HostStaging h;
	size_t pos_span = vertex_count ? (vertex_count - 1) * position_stride + 12 : 0;
	const char* pos_lo = reinterpret_cast<const char*>(positions);
	h.attr_inside = attributes && attribute_count && reinterpret_cast<const char*>(attributes) >= pos_lo &&
	                reinterpret_cast<const char*>(attributes) + (vertex_count ? (vertex_count - 1) * attribute_stride + attribute_count * 4 : 0) <=
	                    pos_lo + (vertex_count ? (vertex_count - 1) * position_stride + position_stride : 0);
	h.pos_bytes = h.attr_inside ? vertex_count * position_stride : pos_span;
	h.attr_span = (attributes && attribute_count && vertex_count && !h.attr_inside) ? (vertex_count - 1) * attribute_stride + attribute_count * 4 : 0;
	h.lock_bytes = lock ? vertex_count : 0;
	h.off_idx = 0;
	h.off_dst = (h.off_idx + index_count * 4 + 255) & ~size_t(255);
	h.off_pos = (h.off_dst + index_count * 4 + 255) & ~size_t(255);
	h.off_attr = (h.off_pos + h.pos_bytes + 255) & ~size_t(255);
	h.off_lock = (h.off_attr + h.attr_span + 255) & ~size_t(255);
	h.total = h.off_lock + h.lock_bytes + 256;
	return h;
}

// One host call on the given context (the caller holds whatever lock guards it): stage the used
// spans of the inputs, run the device path, copy the result back.  With the solve bit the
// staged position and attribute spans are updated on the device and copied back over the
// caller's (whose buffers are then writable; bytes the solve leaves alone return unchanged).
static size_t simplify_host_on(Context& ctx, const unsigned int* indices, unsigned int* destination, size_t index_count, const float* positions,
                               size_t vertex_count, size_t position_stride, const float* attributes, size_t attribute_stride, const float* weights,
                               size_t attribute_count, const unsigned char* lock, size_t target_index_count, float target_error, unsigned int options,
                               float* result_error)
{
	const bool update = (options & kOptInternalSolve) != 0;
	ctx.ensure_init();
	hipStream_t s = ctx.stream;

	const HostStaging h = plan_staging(index_count, positions, vertex_count, position_stride, attributes, attribute_stride, attribute_count, lock);
	const size_t off_idx = h.off_idx, off_dst = h.off_dst, off_pos = h.off_pos, off_attr = h.off_attr, off_lock = h.off_lock;
	const size_t lock_bytes = h.lock_bytes;
	char* d = static_cast<char*>(ctx.staging.get(h.total));
	if (index_count)
		HIP_CHECK(hipMemcpyAsync(d + off_idx, indices, index_count * 4, hipMemcpyHostToDevice, s));
	if (h.pos_bytes)
		HIP_CHECK(hipMemcpyAsync(d + off_pos, positions, h.pos_bytes, hipMemcpyHostToDevice, s));
	const float* d_attr = nullptr;
	if (h.attr_inside)
		d_attr = reinterpret_cast<const float*>(d + off_pos + (reinterpret_cast<const char*>(attributes) - reinterpret_cast<const char*>(positions)));
	else if (h.attr_span)
	{
		HIP_CHECK(hipMemcpyAsync(d + off_attr, attributes, h.attr_span, hipMemcpyHostToDevice, s));
		d_attr = reinterpret_cast<const float*>(d + off_attr);
	}
	if (lock_bytes)
		HIP_CHECK(hipMemcpyAsync(d + off_lock, lock, lock_bytes, hipMemcpyHostToDevice, s));

	CallArgs c{};
	c.destination = reinterpret_cast<unsigned int*>(d + off_dst);
	c.indices = reinterpret_cast<const unsigned int*>(d + off_idx);
	c.index_count = index_count;
	c.positions = reinterpret_cast<const float*>(d + off_pos);
	c.vertex_count = vertex_count;
	c.position_stride = position_stride;
	c.attributes = d_attr;
	c.attribute_stride = attribute_stride;
	c.weights = weights;
	c.attribute_count = attribute_count;
	c.lock = lock_bytes ? reinterpret_cast<const unsigned char*>(d + off_lock) : nullptr;
	c.target_index_count = target_index_count;
	c.target_error = target_error;
	c.options = options;
	c.out_positions = update ? reinterpret_cast<float*>(d + off_pos) : nullptr;
	c.out_attributes = update ? const_cast<float*>(d_attr) : nullptr;

	size_t n = simplify_on_device(ctx, c, result_error);
	if (n)
		HIP_CHECK(hipMemcpyAsync(destination, c.destination, n * 4, hipMemcpyDeviceToHost, s));
	if (update && h.pos_bytes)
		HIP_CHECK(hipMemcpyAsync(const_cast<float*>(positions), d + off_pos, h.pos_bytes, hipMemcpyDeviceToHost, s));
	if (update && h.attr_span)
		HIP_CHECK(hipMemcpyAsync(const_cast<float*>(attributes), d + off_attr, h.attr_span, hipMemcpyDeviceToHost, s));
	HIP_CHECK(hipStreamSynchronize(s));
	return n;
}

static size_t simplify_host(const unsigned int* indices, unsigned int* destination, size_t index_count, const float* positions, size_t vertex_count,
                            size_t position_stride, const float* attributes, size_t attribute_stride, const float* weights, size_t attribute_count,
                            const unsigned char* lock, size_t target_index_count, float target_error, unsigned int options, float* result_error)
{
	PooledContext pooled;
	Context& ctx = *pooled.ctx;
	std::lock_guard<std::mutex> guard(ctx.mutex);
	return simplify_host_on(ctx, indices, destination, index_count, positions, vertex_count, position_stride, attributes, attribute_stride, weights,
	                        attribute_count, lock, target_index_count, target_error, options, result_error);
}

// ------------------------------------------------------------------------------------------
// meshopt_simplifySloppy (src/simplifier.cpp:2647-2778) on device-resident inputs: the grid-size
// search runs on the host exactly as the reference's loop (one device count per pass), the
// cell and triangle steps on the device.

// interpolate, src/simplifier.cpp:2327-2334 (host arithmetic, no contraction)
static float sloppy_interpolate(float y, float x0, float y0, float x1, float y1, float x2, float y2)
{
	float num = (y1 - y) * (x1 - x2) * (x1 - x0) * (y2 - y0);
	float den = (y2 - y) * (x1 - x2) * (y0 - y1) + (y0 - y) * (x1 - x0) * (y1 - y2);
	return x1 + (den == 0.f ? 0.f : num / den);
}

static size_t sloppy_on_device(Context& ctx, unsigned int* destination, const unsigned int* indices, size_t index_count, const float* positions,
                               size_t vertex_count, size_t stride, const unsigned char* lock, size_t target_index_count, float target_error,
                               float* out_result_error, float* out_extent = nullptr)
{
	hipStream_t s = ctx.stream;
	const size_t V = vertex_count, I = index_count, T = I / 3;
	if (V >= 0xffffffffu || I >= 0xffffffffu)
		throw std::invalid_argument("meshopt_simplifySloppy: more than 2^32 - 1 vertices or indices");

	// carve the call's buffers from the context arena (256-byte aligned)
	const uint32_t vtable_size = (uint32_t)std::max<size_t>(pow2_at_least(2 * V + 2), 16), ttable_size = (uint32_t)std::max<size_t>(pow2_at_least(2 * T + 2), 16);
	const size_t sort_words = sort_temp_words(3 * T + 1);
	const size_t scan_words = std::max(scan_temp_words(V + 1), scan_temp_words(T + 1));
	size_t off = 0;
	auto take = [&](size_t bytes) {
		size_t at = off;
		off = (off + bytes + 255) & ~size_t(255);
		return at;
	};
	const size_t o_praw = take(16 * (V + 1)), o_pos = take(16 * (V + 1)), o_sc = take(sizeof(CallScalars));
	const size_t o_words = take(4 * 8), o_ids = take(4 * (V + 1)), o_vtable = take(4 * (size_t)vtable_size), o_rep = take(4 * (V + 1));
	const size_t o_cscan = take(4 * (V + 2)), o_cells = take(4 * (V + 1)), o_cfirst = take(4 * (V + 1));
	const size_t o_ek = take(4 * (3 * T + 1)), o_ev = take(4 * (3 * T + 1)), o_eka = take(4 * (3 * T + 1)), o_eva = take(4 * (3 * T + 1));
	const size_t o_coff = take(4 * (V + 2)), o_stemp = take(4 * sort_words), o_ctemp = take(4 * scan_words), o_cq = take(sizeof(Quad) * (V + 1));
	const size_t o_best = take(8 * (V + 1)), o_cremap = take(4 * (V + 1)), o_tri = take(16 * (T + 1)), o_ttable = take(4 * (size_t)ttable_size);
	const size_t o_keep = take(4 * (T + 2)), o_face = take(sizeof(Quad) * (T + 1));
	ctx.ensure_arena(off);
	char* base = ctx.arena;

	Mesh m{};
	m.vertex_count = (uint32_t)V;
	m.praw = reinterpret_cast<float4*>(base + o_praw);
	m.pos = reinterpret_cast<float4*>(base + o_pos);
	m.sc = reinterpret_cast<CallScalars*>(base + o_sc);
	uint32_t* words = reinterpret_cast<uint32_t*>(base + o_words); // count, cell count, kept, error bits, index check

	SloppyBuffers b{};
	b.indices = indices;
	b.pos = m.pos;
	b.lock = lock;
	b.vertex_count = (uint32_t)V;
	b.index_count = (uint32_t)I;
	b.count = words + 0;
	b.ids = reinterpret_cast<uint32_t*>(base + o_ids);
	b.vtable = reinterpret_cast<uint32_t*>(base + o_vtable);
	b.vtable_size = vtable_size;
	b.rep = reinterpret_cast<uint32_t*>(base + o_rep);
	b.cell_scan = reinterpret_cast<uint32_t*>(base + o_cscan);
	b.cells = reinterpret_cast<uint32_t*>(base + o_cells);
	b.cell_first = reinterpret_cast<uint32_t*>(base + o_cfirst);
	b.entry_keys = reinterpret_cast<uint32_t*>(base + o_ek);
	b.entry_vals = reinterpret_cast<uint32_t*>(base + o_ev);
	b.entry_keys_alt = reinterpret_cast<uint32_t*>(base + o_eka);
	b.entry_vals_alt = reinterpret_cast<uint32_t*>(base + o_eva);
	b.cell_off = reinterpret_cast<uint32_t*>(base + o_coff);
	b.sort_temp = reinterpret_cast<uint32_t*>(base + o_stemp);
	b.scan_temp = reinterpret_cast<uint32_t*>(base + o_ctemp);
	b.cq = reinterpret_cast<Quad*>(base + o_cq);
	b.best = reinterpret_cast<uint64_t*>(base + o_best);
	b.cell_remap = reinterpret_cast<uint32_t*>(base + o_cremap);
	b.error_bits = words + 3;
	b.tri = reinterpret_cast<uint4*>(base + o_tri);
	b.ttable = reinterpret_cast<uint32_t*>(base + o_ttable);
	b.ttable_size = ttable_size;
	b.keep = reinterpret_cast<uint32_t*>(base + o_keep);
	b.face = reinterpret_cast<Quad*>(base + o_face);

	uint32_t* h = ctx.sloppy_pinned();
	auto read_words = [&](int first, int count) {
		HIP_CHECK(hipMemcpyAsync(h + first, words + first, count * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
	};

	// every later gather trusts the indices: reject out-of-range ones before any kernel uses them
	if (I)
	{
		HIP_CHECK(hipMemsetAsync(words + 4, 0, sizeof(uint32_t), s));
		launch_check_indices(s, indices, I, V, words + 4);
		read_words(4, 1);
		if (h[4])
			throw std::invalid_argument("meshopt_simplifySloppy: index buffer references a vertex >= vertex_count");
	}

	// rescalePositions (src/simplifier.cpp:549-603)
	if (V)
		launch_gather_positions(s, positions, stride / sizeof(float), nullptr, m.praw, V);
	launch_rescale(s, m);
	// the extent meshopt_simplifyScale returns for the same vertices (rescalePositions, without the
	// rescale), asked for by the clusterlod fallback
	if (out_extent)
	{
		HIP_CHECK(hipMemcpyAsync(out_extent, &m.sc->extent, sizeof(float), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
	}

	auto count_at = [&](int grid_size) -> size_t {
		launch_sloppy_count(s, b, grid_size);
		read_words(0, 1);
		return h[0];
	};

	// the guided grid-size search (src/simplifier.cpp:2667-2723)
	const size_t target_cell_count = target_index_count / 6;
	const int interpolation_passes = 5;
	int min_grid = int(1.f / (target_error < 1e-3f ? 1e-3f : (target_error < 1.f ? target_error : 1.f)));
	int max_grid = 1025;
	size_t min_triangles = 0;
	size_t max_triangles = T;
	if (min_grid > 1 || lock)
		min_triangles = count_at(min_grid);

	int next_grid_size = int(sqrtf(float(target_cell_count)) + 0.5f);
	for (int pass = 0; pass < 10 + interpolation_passes; ++pass)
	{
		if (min_triangles >= target_index_count / 3 || max_grid - min_grid <= 1)
			break;
		int grid_size = next_grid_size;
		grid_size = (grid_size <= min_grid) ? min_grid + 1 : (grid_size >= max_grid ? max_grid - 1 : grid_size);
		const size_t triangles = count_at(grid_size);
		const float tip = sloppy_interpolate(float(size_t(target_index_count / 3)), float(min_grid), float(min_triangles), float(grid_size),
		                                     float(triangles), float(max_grid), float(max_triangles));
		if (triangles <= target_index_count / 3)
		{
			min_grid = grid_size;
			min_triangles = triangles;
		}
		else
		{
			max_grid = grid_size;
			max_triangles = triangles;
		}
		next_grid_size = (pass < interpolation_passes) ? int(tip + 0.5f) : (min_grid + max_grid) / 2;
	}

	if (min_triangles == 0)
	{
		if (out_result_error)
			*out_result_error = 1.f;
		return 0;
	}

	launch_sloppy_cells(s, b, min_grid, words + 1);
	read_words(1, 1);
	const uint32_t cell_count = h[1];
	launch_sloppy_quadrics(s, b, cell_count);
	launch_sloppy_filter(s, b, cell_count, destination, words + 2);
	read_words(2, 2);
	const uint32_t kept = h[2];
	float err_sq;
	memcpy(&err_sq, &h[3], sizeof(float));
	if (out_result_error)
		*out_result_error = sqrtf(err_sq);
	return (size_t)kept * 3;
}

static void check_args(size_t index_count, size_t stride, size_t target, float target_error, size_t attr_stride, size_t attr_count, const float* weights)
{
	// preconditions of src/simplifier.cpp:2348-2357 (asserts in the reference)
	if (index_count % 3 != 0 || stride < 12 || stride > 256 || stride % 4 != 0 || target > index_count || !(target_error >= 0) ||
	    attr_count > (size_t)kMaxAttributes || (attr_count && (attr_stride < attr_count * 4 || attr_stride > 256 || attr_stride % 4 != 0)))
		throw std::invalid_argument("meshopt_simplify: invalid arguments");
	for (size_t k = 0; k < attr_count; ++k)
		if (!(weights[k] >= 0))
			throw std::invalid_argument("meshopt_simplify: negative attribute weight");
}

} // namespace mhip

using namespace mhip;

[[noreturn]] static void die(const char* what)
{
	fprintf(stderr, "libmeshopt_hip: %s\n", what);
	fflush(stderr);
	abort();
}

extern "C" size_t meshopt_simplify(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions, size_t vertex_count,
                                   size_t vertex_positions_stride, size_t target_index_count, float target_error, unsigned int options, float* result_error)
{
	try
	{
		check_args(index_count, vertex_positions_stride, target_index_count, target_error, 0, 0, nullptr);
		if (options & kOptInternalSolve)
			throw std::invalid_argument("meshopt_simplify: the solve bit is not accepted here");
		return simplify_host(indices, destination, index_count, vertex_positions, vertex_count, vertex_positions_stride, nullptr, 0, nullptr, 0,
		                     nullptr, target_index_count, target_error, options, result_error);
	}
	catch (const std::exception& e)
	{
		die(e.what());
	}
}

extern "C" size_t meshopt_simplifyWithAttributes(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions,
                                                 size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes,
                                                 size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count,
                                                 const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options,
                                                 float* result_error)
{
	try
	{
		check_args(index_count, vertex_positions_stride, target_index_count, target_error, vertex_attributes_stride, attribute_count, attribute_weights);
		if (options & kOptInternalSolve)
			throw std::invalid_argument("meshopt_simplifyWithAttributes: the solve bit is not accepted here");
		return simplify_host(indices, destination, index_count, vertex_positions, vertex_count, vertex_positions_stride, vertex_attributes,
		                     vertex_attributes_stride, attribute_weights, attribute_count, vertex_lock, target_index_count, target_error, options,
		                     result_error);
	}
	catch (const std::exception& e)
	{
		die(e.what());
	}
}

extern "C" size_t meshopt_simplifyWithUpdate(unsigned int* indices, size_t index_count, float* vertex_positions, size_t vertex_count,
                                             size_t vertex_positions_stride, float* vertex_attributes, size_t vertex_attributes_stride,
                                             const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock,
                                             size_t target_index_count, float target_error, unsigned int options, float* result_error)
{
	try
	{
		check_args(index_count, vertex_positions_stride, target_index_count, target_error, vertex_attributes_stride, attribute_count, attribute_weights);
		return simplify_host(indices, indices, index_count, vertex_positions, vertex_count, vertex_positions_stride, vertex_attributes,
		                     vertex_attributes_stride, attribute_weights, attribute_count, vertex_lock, target_index_count, target_error,
		                     options | kOptInternalSolve, result_error);
	}
	catch (const std::exception& e)
	{
		die(e.what());
	}
}

extern "C" size_t meshopt_simplifySloppy(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions,
                                         size_t vertex_count, size_t vertex_positions_stride, const unsigned char* vertex_lock, size_t target_index_count,
                                         float target_error, float* result_error)
{
	try
	{
		check_args(index_count, vertex_positions_stride, target_index_count, target_error, 0, 0, nullptr);
		PooledContext pooled;
		Context& ctx = *pooled.ctx;
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.ensure_init();
		hipStream_t s = ctx.stream;
		const HostStaging h = plan_staging(index_count, vertex_positions, vertex_count, vertex_positions_stride, nullptr, 0, 0, vertex_lock);
		char* d = static_cast<char*>(ctx.staging.get(h.total));
		if (index_count)
			HIP_CHECK(hipMemcpyAsync(d + h.off_idx, indices, index_count * 4, hipMemcpyHostToDevice, s));
		if (h.pos_bytes)
			HIP_CHECK(hipMemcpyAsync(d + h.off_pos, vertex_positions, h.pos_bytes, hipMemcpyHostToDevice, s));
		if (h.lock_bytes)
			HIP_CHECK(hipMemcpyAsync(d + h.off_lock, vertex_lock, h.lock_bytes, hipMemcpyHostToDevice, s));
		unsigned int* d_dst = reinterpret_cast<unsigned int*>(d + h.off_dst);
		size_t n = sloppy_on_device(ctx, d_dst, reinterpret_cast<const unsigned int*>(d + h.off_idx), index_count,
		                            reinterpret_cast<const float*>(d + h.off_pos), vertex_count, vertex_positions_stride,
		                            h.lock_bytes ? reinterpret_cast<const unsigned char*>(d + h.off_lock) : nullptr, target_index_count, target_error,
		                            result_error);
		if (n)
			HIP_CHECK(hipMemcpyAsync(destination, d_dst, n * 4, hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		return n;
	}
	catch (const std::exception& e)
	{
		die(e.what());
	}
}

extern "C" size_t meshopt_hip_simplify_sloppy_device(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions,
                                                     size_t vertex_count, size_t vertex_positions_stride, const unsigned char* vertex_lock,
                                                     size_t target_index_count, float target_error, float* result_error)
{
	try
	{
		check_args(index_count, vertex_positions_stride, target_index_count, target_error, 0, 0, nullptr);
		PooledContext pooled;
		Context& ctx = *pooled.ctx;
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.ensure_init();
		size_t n = sloppy_on_device(ctx, destination, indices, index_count, vertex_positions, vertex_count, vertex_positions_stride, vertex_lock,
		                            target_index_count, target_error, result_error);
		HIP_CHECK(hipStreamSynchronize(ctx.stream));
		t_last_error.clear();
		return n;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return (size_t)-1;
	}
}

extern "C" size_t meshopt_hip_simplify_device(unsigned int* destination, const unsigned int* indices, size_t index_count, const float* vertex_positions,
                                              size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes,
                                              size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count,
                                              const unsigned char* vertex_lock, size_t target_index_count, float target_error, unsigned int options,
                                              float* result_error)
{
	try
	{
		check_args(index_count, vertex_positions_stride, target_index_count, target_error, vertex_attributes_stride, attribute_count, attribute_weights);
		PooledContext pooled;
		Context& ctx = *pooled.ctx;
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.ensure_init();
		CallArgs c{};
		c.destination = destination;
		c.indices = indices;
		c.index_count = index_count;
		c.positions = vertex_positions;
		c.vertex_count = vertex_count;
		c.position_stride = vertex_positions_stride;
		c.attributes = attribute_count ? vertex_attributes : nullptr;
		c.attribute_stride = vertex_attributes_stride;
		c.weights = attribute_weights;
		c.attribute_count = attribute_count;
		c.lock = vertex_lock;
		c.target_index_count = target_index_count;
		c.target_error = target_error;
		c.options = options & ~(unsigned int)kOptInternalSolve;
		size_t n = simplify_on_device(ctx, c, result_error);
		t_last_error.clear();
		return n;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return (size_t)-1;
	}
}

extern "C" size_t meshopt_hip_simplify_with_update_device(unsigned int* indices, size_t index_count, float* vertex_positions, size_t vertex_count,
                                                          size_t vertex_positions_stride, float* vertex_attributes, size_t vertex_attributes_stride,
                                                          const float* attribute_weights, size_t attribute_count, const unsigned char* vertex_lock,
                                                          size_t target_index_count, float target_error, unsigned int options, float* result_error)
{
	try
	{
		check_args(index_count, vertex_positions_stride, target_index_count, target_error, vertex_attributes_stride, attribute_count, attribute_weights);
		PooledContext pooled;
		Context& ctx = *pooled.ctx;
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.ensure_init();
		CallArgs c{};
		c.destination = indices;
		c.indices = indices;
		c.index_count = index_count;
		c.positions = vertex_positions;
		c.vertex_count = vertex_count;
		c.position_stride = vertex_positions_stride;
		c.attributes = attribute_count ? vertex_attributes : nullptr;
		c.attribute_stride = vertex_attributes_stride;
		c.weights = attribute_weights;
		c.attribute_count = attribute_count;
		c.lock = vertex_lock;
		c.target_index_count = target_index_count;
		c.target_error = target_error;
		c.options = options | kOptInternalSolve;
		c.out_positions = vertex_positions;
		c.out_attributes = attribute_count ? vertex_attributes : nullptr;
		size_t n = simplify_on_device(ctx, c, result_error);
		HIP_CHECK(hipStreamSynchronize(ctx.stream));
		t_last_error.clear();
		return n;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return (size_t)-1;
	}
}

// ------------------------------------------------------------------------------------------
// batch of independent meshes (clusterlod patches)

namespace mhip
{


// The tuning table of the batched calls' launch shape: the file named by MESHOPT_HIP_BATCH_SHAPE,
// else batch_shape.txt beside this library (written by meshoptimizer_amd/tune.py).  Lines
// "per_cu a b c d", "threads a b c d" and "cursor a b c d" (one number per size class); '#' starts
// a comment.  Returns false (and leaves *shape alone) when there is no valid table.
static bool read_batch_shape_table(BatchShape* shape)
{
	std::string path;
	if (const char* env = getenv("MESHOPT_HIP_BATCH_SHAPE"))
		path = env;
	else
	{
		Dl_info info{};
		if (!dladdr(reinterpret_cast<void*>(&read_batch_shape_table), &info) || !info.dli_fname)
			return false;
		path = info.dli_fname;
		size_t slash = path.rfind('/');
		path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/batch_shape.txt";
	}
	FILE* f = fopen(path.c_str(), "r");
	if (!f)
		return false;
	BatchShape t = batch_default_shape();
	int seen = 0;
	char line[256];
	while (fgets(line, sizeof(line), f))
	{
		if (char* hash = strchr(line, '#'))
			*hash = 0;
		char key[32];
		unsigned int v[kBatchClasses];
		static_assert(kBatchClasses == 4, "table rows hold four classes");
		if (sscanf(line, "%31s %u %u %u %u", key, &v[0], &v[1], &v[2], &v[3]) != 5)
			continue;
		uint32_t* dst = !strcmp(key, "per_cu") ? t.per_cu : !strcmp(key, "threads") ? t.threads : !strcmp(key, "cursor") ? t.cursor : nullptr;
		if (!dst)
			continue;
		for (int k = 0; k < kBatchClasses; ++k)
			dst[k] = v[k];
		seen |= dst == t.per_cu ? 1 : dst == t.threads ? 2 : 4;
	}
	fclose(f);
	if (seen != 7 || batch_shape_error(t))
		return false;
	*shape = t;
	return true;
}

// the size classes of the context's shape on its device: classes beyond 64 KB of LDS need the
// device's limit and the kernels' permission; without them the shape's classes are those of a
// 64 KB workgroup
static const BatchClasses& ensure_batch_classes(Context& ctx)
{
	if (ctx.batch_cls_ready)
		return ctx.batch_cls;
	if (!ctx.batch_shape_set)
	{
		BatchShape t;
		if (read_batch_shape_table(&t))
			ctx.batch_shape = t;
		ctx.batch_shape_set = true;
	}
	int max_block = 0, cus = 0;
	HIP_CHECK(hipDeviceGetAttribute(&max_block, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx.device));
	HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx.device));
	ctx.cu_count = (size_t)(cus > 0 ? cus : 1);
	BatchClasses c = batch_classes((uint32_t)max_block, ctx.batch_shape);
	bool allowed = true;
	for (int k = 0; k < kBatchClasses; ++k)
	{
		if (!c.icap[k])
			continue;
		const size_t bytes = batch_lds_bytes(c.shape.threads[k], c.vcap[k], c.icap[k], c.ccap[k]);
		if (bytes > 64 * 1024 && !batch_allow_lds(c.shape.threads[k], bytes))
			allowed = false;
	}
	if (!allowed)
	{
		(void)hipGetLastError();
		c = batch_classes(64 * 1024, ctx.batch_shape);
	}
	ctx.batch_cls = c;
	ctx.batch_cls_ready = true;
	return ctx.batch_cls;
}

// Memory of one batched call.  words: batch_list_words(P) device words (status, class counts, id
// lists, cursor); stats: 4 device words per mesh or null; quadrics: one per vertex of the batch;
// host: batch_host_words(P) words of host memory (status read-back, retry list).  Meshes that fit
// no LDS class take the plain path when plain_fallback is set, which allocates its own arena;
// otherwise the call reports them as an error.  timed: the kernel timing of the context applies.
struct BatchMemory
{
	uint32_t* words;
	uint32_t* stats;
	Quad* quadrics;
	uint32_t* host;
	bool plain_fallback;
	bool timed;
	// attribute state of batches with attributes (null otherwise): per vertex, attr_room
	// rescaled values and gradients, and one attribute quadric
	float* attr_values;
	Quad* attr_quadrics;
	Grad* attr_grads;
	size_t attr_room;
	// device address of `host` when it is pinned memory the kernels may write (then the status
	// words and class counts need no copy back); null for host memory the device cannot write
	uint32_t* host_dev = nullptr;
};

// the attributes of a batched call: device values packed per vertex like the positions, their
// stride in bytes, and the host array of weights
struct BatchAttributes
{
	const float* data;
	size_t stride;
	const float* weights;
	size_t count;
};

// the context's own batch buffers for P meshes and `vertices` vertices (grown on demand), with
// room for `attributes` attributes per vertex
static BatchMemory context_batch_memory(Context& ctx, size_t P, size_t vertices, bool need_quadrics, size_t attributes = 0)
{
	ctx.ensure_batch_buffers(P);
	if (need_quadrics)
		ctx.ensure_batch_quadrics(vertices ? vertices : 1);
	ctx.batch_last_meshes = P;
	BatchMemory m{ctx.batch_dev, ctx.batch_dev + batch_list_words(P), ctx.batch_quadrics, ctx.batch_pinned, true, timing_on(), nullptr, nullptr, nullptr, 0,
	              ctx.batch_pinned_dev};
	if (need_quadrics && attributes)
	{
		const size_t n = vertices ? vertices : 1;
		const size_t o_aq = 0, o_ag = o_aq + n * sizeof(Quad), o_va = o_ag + n * attributes * sizeof(Grad);
		char* base = static_cast<char*>(ctx.batch_attr.get(o_va + n * attributes * sizeof(float) + 256));
		m.attr_quadrics = reinterpret_cast<Quad*>(base + o_aq);
		m.attr_grads = reinterpret_cast<Grad*>(base + o_ag);
		m.attr_values = reinterpret_cast<float*>(base + o_va);
		m.attr_room = attributes;
	}
	return m;
}

// device-to-host copies a caller of the batched call wants along with the status read-back
struct BatchReadback
{
	void* host;
	const void* device;
	size_t bytes;
	bool valid; // set by the call: the copies hold the final results (no plain-path mesh ran after them)
};

// Batched call: sizes are validated and meshes sorted into LDS size classes on the device, the
// classes run side by side (the first on the call's stream, the others on auxiliary streams
// forked from it), and only the per-mesh status comes back to the host.  Meshes that fit no
// class (and Sparse batches) take the plain device path afterwards; their offsets and targets
// are read back only then (h_* may be null: device-resident offsets).  `readback` copies go with
// each status read-back, so a caller that needs the result counts saves a round trip.
static int simplify_batch_on_device(Context& ctx, const BatchMemory& mem, unsigned int* destination, unsigned int* result_counts,
                                    float* result_errors, const unsigned int* indices, const unsigned int* index_offsets, const float* positions,
                                    const unsigned int* vertex_offsets, const unsigned char* vertex_lock, size_t vertex_count, size_t stride,
                                    const unsigned int* target_index_counts,
                                    size_t patch_count, float target_error, unsigned int options, const unsigned int* h_index_offsets,
                                    const unsigned int* h_vertex_offsets, const unsigned int* h_targets,
                                    const unsigned int* h_size_offsets = nullptr, const BatchAttributes* attrs = nullptr,
                                    BatchReadback* readback = nullptr, int readback_count = 0)
{
	hipStream_t s = ctx.stream;
	auto enqueue_readback = [&]() {
		for (int r = 0; r < readback_count; ++r)
			HIP_CHECK(hipMemcpyAsync(readback[r].host, readback[r].device, readback[r].bytes, hipMemcpyDeviceToHost, s));
	};
	for (int r = 0; r < readback_count; ++r)
		readback[r].valid = false;
	if (stride < 12 || stride > 256 || stride % 4)
		throw std::invalid_argument("meshopt_hip_simplify_batch: invalid vertex stride");
	if (options & kOptInternalSolve)
		throw std::invalid_argument("meshopt_hip_simplify_batch: the solve bit is not accepted here");
	// the attributes with weight > 0 enter the quadrics (src/simplifier.cpp:2410-2421)
	const size_t attr_count = attrs ? attrs->count : 0;
	AttrMap attr_map{};
	uint32_t attr_used = 0;
	if (attr_count)
	{
		if (attr_count > (size_t)kMaxAttributes || attrs->stride < attr_count * 4 || attrs->stride > 256 || attrs->stride % 4 != 0 || !attrs->weights)
			throw std::invalid_argument("meshopt_hip_simplify_batch: invalid attribute count, stride or weights");
		for (size_t k = 0; k < attr_count; ++k)
		{
			if (!(attrs->weights[k] >= 0))
				throw std::invalid_argument("meshopt_hip_simplify_batch: negative attribute weight");
			if (attrs->weights[k] > 0)
			{
				attr_map.index[attr_used] = (uint32_t)k;
				attr_map.weight[attr_used] = attrs->weights[k];
				attr_used++;
			}
		}
		if (attr_used > mem.attr_room && !(options & (kOptSparse | kOptPrune)))
			throw std::logic_error("meshopt_hip_simplify_batch: attribute buffers too small");
	}
	if (!patch_count)
		return 0;

	const BatchClasses& classes = ensure_batch_classes(ctx);
	const size_t cu_count = ctx.cu_count;
	const BatchShape& shape = classes.shape;
	const size_t P = patch_count;
	// the status words go straight to pinned host memory when the device may write it (no copy
	// back); the class counts stay on the device (the class kernels read them), with a host copy
	uint32_t* status = mem.host_dev ? mem.host_dev : mem.words;
	uint32_t* counts = mem.words + P;
	uint32_t* host_counts = mem.host_dev ? mem.host_dev + P : nullptr;
	uint32_t* ids = mem.words + P + 4;
	uint32_t* h_status = mem.host;
	uint32_t* h_retry = mem.host + P + kBatchClasses;

	BatchArgs a;
	a.destination = destination;
	a.result_counts = result_counts;
	a.result_errors = result_errors;
	a.status = status;
	a.stats = mem.stats;
	a.indices = indices;
	a.index_offsets = index_offsets;
	a.positions = positions;
	a.vertex_offsets = vertex_offsets;
	a.vertex_lock = vertex_lock;
	a.vertex_count = vertex_count;
	a.quadrics = nullptr;
	a.position_stride_floats = stride / sizeof(float);
	a.target_index_counts = target_index_counts;
	a.patch_count = patch_count;
	a.target_error = target_error;
	a.options = options;
	a.attributes = attr_used ? attrs->data : nullptr;
	a.attribute_stride_floats = attr_used ? attrs->stride / sizeof(float) : 0;
	a.attribute_count = attr_used;
	a.attr_map = attr_map;
	a.attr_values = mem.attr_values;
	a.attr_quadrics = mem.attr_quadrics;
	a.attr_grads = mem.attr_grads;

	// Sparse and Prune meshes take the plain path, mesh by mesh (the batched kernel's LDS layouts
	// hold neither the sparse remap nor the component arrays)
	const bool batched = !(options & (kOptSparse | kOptPrune));
	if (batched)
	{
		a.quadrics = mem.quadrics;
		const bool timed = mem.timed;
		if (timed && !ctx.batch_ev[0])
		{
			HIP_CHECK(hipEventCreate(&ctx.batch_ev[0]));
			HIP_CHECK(hipEventCreate(&ctx.batch_ev[1]));
		}
		if (timed)
			HIP_CHECK(hipEventRecord(ctx.batch_ev[0], s));
		static_assert(kBatchClasses <= 4, "one count word per class before the id lists");
		if (!launch_batch_classify_order(s, a, classes, ids, counts, host_counts))
		{
			fill_u32(s, counts, 0u, kBatchClasses);
			launch_batch_classify(s, a, classes, ids, counts);
			launch_batch_order(s, a, ids, counts, ids + (size_t)kBatchClasses * P, host_counts);
		}
		ctx.ensure_aux_streams();
		HIP_CHECK(hipEventRecord(ctx.fork_ev, s));
		// when the host knows the index counts (host offsets, or the caller's host copy of them)
		// and one of them lies within class 3's index caps only, class 3 launches at once beside
		// the others instead; when every mesh is beyond class 2's index cap, classes 0-2 are empty
		// and not launched
		const int big = kBatchClasses - 1;
		const unsigned int* known = h_index_offsets ? h_index_offsets : h_size_offsets;
		bool small_empty = false, big_early = false;
		if (known && classes.icap[big])
		{
			small_empty = true;
			for (size_t p = 0; p < P; ++p)
			{
				const size_t n = known[p + 1] - known[p];
				small_empty &= n > classes.icap[big - 1];
				big_early |= n > classes.icap[big - 1] && n <= classes.icap[big];
			}
		}
		// a class's workgroups: one per mesh slot (those past the class's count leave at once), or
		// with the cursor form `per_cu` per CU looping over the class's list through cursor word
		// `slot` (one per class, and one more for the retry launch, which may run beside its
		// class's own)
		auto launch_class = [&](hipStream_t cs, int k, const uint32_t* list, const uint32_t* count_dev, size_t meshes_bound, int slot) {
			const bool from_cursor = shape.cursor[k] != 0;
			const size_t grid = from_cursor ? std::min(meshes_bound, (size_t)shape.per_cu[k] * cu_count) : meshes_bound;
			launch_batch(cs, shape.threads[k], a, classes.vcap[k], classes.icap[k], classes.ccap[k], list, count_dev, grid,
			             from_cursor ? ids + 2 * (size_t)kBatchClasses * P + slot : nullptr);
		};
		if (big_early)
		{
			HIP_CHECK(hipStreamWaitEvent(ctx.aux[big - 1], ctx.fork_ev, 0));
			launch_class(ctx.aux[big - 1], big, ids + (size_t)big * P, counts + big, P, big);
			HIP_CHECK(hipEventRecord(ctx.join_ev[big - 1], ctx.aux[big - 1]));
			HIP_CHECK(hipStreamWaitEvent(s, ctx.join_ev[big - 1], 0));
		}
		// classes 0, 1 and 2 launch at once (in the default shape with one workgroup per mesh slot),
		// the class of the largest meshes first: its longer workgroups start early and leave a
		// shorter tail (the 4096-patch batch split over classes 1 and 2 at 6 and 5 per CU measured
		// 0.557 ms with the largest class launched last, 0.512 ms with it first).
		// A workgroup of class 3 takes a whole CU's LDS, so even one that leaves at once waits for
		// a free CU: class 3 launches after the status read-back, only when it has meshes, with at
		// most one workgroup per CU looping over them (reading the counts back before launching
		// classes 2 and 3 instead measured 4.5% slower on the 4096-patch batch).
		static constexpr int kLaunchOrder[kBatchClasses - 1] = {2, 1, 0};
		for (int c = 0; c < kBatchClasses - 1; ++c)
		{
			const int k = kLaunchOrder[c];
			hipStream_t cs = c == 0 ? s : ctx.aux[c - 1];
			if (c)
				HIP_CHECK(hipStreamWaitEvent(cs, ctx.fork_ev, 0));
			if (classes.icap[k] && !small_empty)
				launch_class(cs, k, ids + (size_t)k * P, counts + k, P, k);
			if (c)
			{
				HIP_CHECK(hipEventRecord(ctx.join_ev[c - 1], cs));
				HIP_CHECK(hipStreamWaitEvent(s, ctx.join_ev[c - 1], 0));
			}
		}
		if (timed)
			HIP_CHECK(hipEventRecord(ctx.batch_ev[1], s));
		HIP_CHECK(hipGetLastError());
		if (!mem.host_dev)
			HIP_CHECK(hipMemcpyAsync(h_status, status, (P + kBatchClasses) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
		enqueue_readback();
		HIP_CHECK(hipStreamSynchronize(s));
		auto add_time = [&](size_t launches, size_t meshes) {
			float ms = 0;
			HIP_CHECK(hipEventElapsedTime(&ms, ctx.batch_ev[0], ctx.batch_ev[1]));
			TimingTotals& t = timing_totals();
			std::lock_guard<std::mutex> guard(t.m);
			t.batch_ms += ms;
			t.batch_launches += launches;
			t.batch_meshes += meshes;
		};
		if (timed)
			add_time(1, patch_count);
		// class 3's meshes (status 6 until they run), and the meshes with more candidates than
		// their class's arrays hold (status 5), which run again in the retry class: its arrays
		// hold the full bound and its caps every mesh of the classes before it
		const uint32_t large = big_early ? 0 : h_status[P + kBatchClasses - 1];
		uint32_t retries = 0;
		for (size_t p = 0; p < P; ++p)
			if (h_status[p] == 5)
			{
				// a retry class left without caps (LDS limits) sends the mesh to the plain path
				if (classes.icap[kBatchRetryClass])
					h_retry[retries++] = (uint32_t)p;
				else
					h_status[p] = 1;
			}
		if (retries || large)
		{
			if (timed)
				HIP_CHECK(hipEventRecord(ctx.batch_ev[0], s));
			hipStream_t ls = ctx.aux[kBatchClasses - 2];
			if (large)
			{
				const int c = kBatchClasses - 1;
				HIP_CHECK(hipEventRecord(ctx.fork_ev, s));
				HIP_CHECK(hipStreamWaitEvent(ls, ctx.fork_ev, 0));
				launch_class(ls, c, ids + (size_t)c * P, counts + c, large, c);
			}
			if (retries)
			{
				// the id list and count of class 0 hold the retry list (its meshes have run); the
				// copies read host memory, which stays untouched until the status read-back below
				const int c = kBatchRetryClass;
				h_retry[P] = retries;
				HIP_CHECK(hipMemcpyAsync(ids, h_retry, retries * sizeof(uint32_t), hipMemcpyHostToDevice, s));
				HIP_CHECK(hipMemcpyAsync(counts, h_retry + P, sizeof(uint32_t), hipMemcpyHostToDevice, s));
				launch_class(s, c, ids, counts, retries, kBatchClasses);
			}
			if (large)
			{
				HIP_CHECK(hipEventRecord(ctx.join_ev[kBatchClasses - 2], ls));
				HIP_CHECK(hipStreamWaitEvent(s, ctx.join_ev[kBatchClasses - 2], 0));
			}
			if (timed)
				HIP_CHECK(hipEventRecord(ctx.batch_ev[1], s));
			HIP_CHECK(hipGetLastError());
			if (!mem.host_dev)
				HIP_CHECK(hipMemcpyAsync(h_status, status, P * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
			enqueue_readback();
			HIP_CHECK(hipStreamSynchronize(s));
			if (timed)
				add_time(0, 0);
		}
	}
	else
		std::fill(h_status, h_status + P, 1u);

	bool fallback = false;
	for (size_t p = 0; p < P; ++p)
	{
		if (h_status[p] == 2)
			throw std::invalid_argument("meshopt_hip_simplify_batch: index buffer references a vertex >= vertex_count");
		if (h_status[p] == 4)
			throw std::invalid_argument("meshopt_hip_simplify_batch: invalid index count, target or offsets");
		if (h_status[p] == 3)
			throw std::runtime_error("meshopt_hip_simplify_batch: internal error (pass bound exceeded)");
		if (h_status[p] == 5)
			throw std::runtime_error("meshopt_hip_simplify_batch: internal error (candidate arrays of the last class exceeded)");
		if (h_status[p] == 6)
			throw std::runtime_error("meshopt_hip_simplify_batch: internal error (a mesh of the largest class did not run)");
		fallback |= h_status[p] == 1;
	}
	if (!fallback)
	{
		for (int r = 0; r < readback_count; ++r)
			readback[r].valid = true;
		return 0;
	}
	if (!mem.plain_fallback)
	{
		size_t first = 0;
		while (h_status[first] != 1)
			++first;
		char msg[256];
		snprintf(msg, sizeof(msg),
		         "meshopt_hip_simplify_batch: mesh %zu fits none of the batched kernel's LDS size classes (or the options need the plain path), "
		         "and this entry point does not run the plain path",
		         first);
		throw std::invalid_argument(msg);
	}

	// the plain device path for the meshes the batched kernel did not take
	std::vector<unsigned int> io, vo, tg;
	if (!h_index_offsets)
	{
		io.resize(P + 1);
		vo.resize(P + 1);
		tg.resize(P);
		HIP_CHECK(hipMemcpyAsync(io.data(), index_offsets, (P + 1) * sizeof(unsigned int), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipMemcpyAsync(vo.data(), vertex_offsets, (P + 1) * sizeof(unsigned int), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipMemcpyAsync(tg.data(), target_index_counts, P * sizeof(unsigned int), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		h_index_offsets = io.data();
		h_vertex_offsets = vo.data();
		h_targets = tg.data();
	}
	for (size_t p = 0; p < P; ++p)
	{
		if (h_status[p] != 1)
			continue;
		if (h_index_offsets[p + 1] < h_index_offsets[p] || h_vertex_offsets[p + 1] < h_vertex_offsets[p] ||
		    (h_index_offsets[p + 1] - h_index_offsets[p]) % 3 || h_targets[p] > h_index_offsets[p + 1] - h_index_offsets[p])
			throw std::invalid_argument("meshopt_hip_simplify_batch: invalid index count, target or offsets");
		CallArgs c{};
		c.destination = destination + h_index_offsets[p];
		c.indices = indices + h_index_offsets[p];
		c.index_count = h_index_offsets[p + 1] - h_index_offsets[p];
		c.positions = positions + (size_t)h_vertex_offsets[p] * (stride / sizeof(float));
		c.vertex_count = h_vertex_offsets[p + 1] - h_vertex_offsets[p];
		c.position_stride = stride;
		c.attributes = attr_count ? attrs->data + (size_t)h_vertex_offsets[p] * (attrs->stride / sizeof(float)) : nullptr;
		c.attribute_stride = attr_count ? attrs->stride : 0;
		c.weights = attr_count ? attrs->weights : nullptr;
		c.attribute_count = attr_count;
		c.lock = vertex_lock ? vertex_lock + h_vertex_offsets[p] : nullptr;
		c.target_index_count = h_targets[p];
		c.target_error = target_error;
		c.options = options;
		float err = 0;
		uint32_t n = (uint32_t)simplify_on_device(ctx, c, &err);
		HIP_CHECK(hipMemcpyAsync(result_counts + p, &n, sizeof(uint32_t), hipMemcpyHostToDevice, s));
		if (result_errors)
			HIP_CHECK(hipMemcpyAsync(result_errors + p, &err, sizeof(float), hipMemcpyHostToDevice, s));
		// the mesh's statistics as the batched kernel records them
		uint32_t st[4] = {(uint32_t)t_passes.size(), 0, 0, 0};
		for (const meshopt_hip_pass_stats& ps : t_passes)
		{
			st[1] += ps.live_triangles;
			st[2] += ps.candidates;
			st[3] += ps.committed;
		}
		if (mem.stats)
			HIP_CHECK(hipMemcpyAsync(mem.stats + 4 * p, st, sizeof(st), hipMemcpyHostToDevice, s));
		HIP_CHECK(hipStreamSynchronize(s));
	}
	return 0;
}

} // namespace mhip

static int batch_device_entry(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices,
                              const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets,
                              const BatchAttributes* attrs, const unsigned char* vertex_lock, size_t vertex_count, size_t vertex_positions_stride,
                              const unsigned int* target_index_counts, size_t mesh_count, float target_error, unsigned int options)
{
	try
	{
		Context& ctx = context();
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.ensure_init();
		const bool batched = !(options & (kOptSparse | kOptPrune));
		simplify_batch_on_device(ctx, context_batch_memory(ctx, mesh_count, vertex_count, batched, attrs ? attrs->count : 0), destination,
		                         result_counts, result_errors, indices, index_offsets, vertex_positions, vertex_offsets, vertex_lock, vertex_count,
		                         vertex_positions_stride, target_index_counts, mesh_count, target_error, options, nullptr, nullptr, nullptr, nullptr,
		                         attrs);
		t_last_error.clear();
		return 0;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return -1;
	}
}

extern "C" int meshopt_hip_simplify_batch_device(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices,
                                                 const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets,
                                                 const unsigned char* vertex_lock, size_t vertex_count, size_t vertex_positions_stride,
                                                 const unsigned int* target_index_counts, size_t mesh_count, float target_error, unsigned int options)
{
	return batch_device_entry(destination, result_counts, result_errors, indices, index_offsets, vertex_positions, vertex_offsets, nullptr, vertex_lock,
	                          vertex_count, vertex_positions_stride, target_index_counts, mesh_count, target_error, options);
}

extern "C" int meshopt_hip_simplify_batch_with_attributes_device(unsigned int* destination, unsigned int* result_counts, float* result_errors,
                                                                 const unsigned int* indices, const unsigned int* index_offsets,
                                                                 const float* vertex_positions, const unsigned int* vertex_offsets,
                                                                 const float* vertex_attributes, size_t vertex_attributes_stride,
                                                                 const float* attribute_weights, size_t attribute_count,
                                                                 const unsigned char* vertex_lock, size_t vertex_count, size_t vertex_positions_stride,
                                                                 const unsigned int* target_index_counts, size_t mesh_count, float target_error,
                                                                 unsigned int options)
{
	const BatchAttributes attrs{vertex_attributes, vertex_attributes_stride, attribute_weights, attribute_count};
	return batch_device_entry(destination, result_counts, result_errors, indices, index_offsets, vertex_positions, vertex_offsets,
	                          attribute_count ? &attrs : nullptr, vertex_lock, vertex_count, vertex_positions_stride, target_index_counts, mesh_count,
	                          target_error, options);
}

// ------------------------------------------------------------------------------------------
// allocation-free batched entry: the device scratch of its calls is one file-scope array

namespace mhip
{

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20; // per GPU, held until the process ends
constexpr size_t kNoAllocMaxMeshes = 16384;
constexpr int kNoAllocMaxDevices = 64;

// the batch's status words, class lists and vertex quadrics (the quadrics start 256-byte aligned
// after the lists); in host code this name is a stand-in, not a device address: host code uses
// the address looked up per device instead
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// the device address of `scratch` on each device (looked up once per device, under the context lock)
static void* g_scratch_device[kNoAllocMaxDevices];
// host words of the calls (status read-back, retry list), under the context lock
static uint32_t g_noalloc_host[2 * kNoAllocMaxMeshes + kBatchClasses + 1];

// set-up on the first call on a device: the context's streams and the device address of the scratch
static void* noalloc_setup(Context& ctx)
{
	int dev = 0;
	HIP_CHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= kNoAllocMaxDevices)
		throw std::invalid_argument("meshopt_hip_simplify_batch_device_noalloc: device index out of range");
	ctx.ensure_stream();
	if (dev != ctx.device)
		throw std::invalid_argument("meshopt_hip_simplify_batch_device_noalloc: the library's context lives on another device");
	ctx.ensure_aux_streams();
	if (!g_scratch_device[dev])
	{
		void* ptr = nullptr;
		HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
		if (!ptr)
			throw std::runtime_error("meshopt_hip_simplify_batch_device_noalloc: no device address for the scratch array");
		g_scratch_device[dev] = ptr;
	}
	return g_scratch_device[dev];
}

} // namespace mhip

extern "C" int meshopt_hip_simplify_batch_device_noalloc(unsigned int* destination, unsigned int* result_counts, float* result_errors,
                                                         const unsigned int* indices, const unsigned int* index_offsets, const float* vertex_positions,
                                                         const unsigned int* vertex_offsets, const unsigned char* vertex_lock, size_t vertex_count,
                                                         size_t vertex_positions_stride, const unsigned int* target_index_counts, size_t mesh_count,
                                                         float target_error, unsigned int options)
{
	try
	{
		if (mesh_count > kNoAllocMaxMeshes)
			throw std::invalid_argument("meshopt_hip_simplify_batch_device_noalloc: more than 16384 meshes");
		if (options & (kOptSparse | kOptPrune))
			throw std::invalid_argument("meshopt_hip_simplify_batch_device_noalloc: Sparse and Prune batches run on the plain path, which allocates");
		const size_t list_bytes = (batch_list_words(mesh_count) * sizeof(uint32_t) + 255) & ~size_t(255);
		const size_t need = list_bytes + (vertex_count ? vertex_count : 1) * sizeof(Quad);
		if (need > kNoAllocScratchBytes)
		{
			char msg[200];
			snprintf(msg, sizeof(msg), "meshopt_hip_simplify_batch_device_noalloc: the batch needs %zu bytes of scratch, more than the %zu it holds", need,
			         kNoAllocScratchBytes);
			throw std::invalid_argument(msg);
		}
		Context& ctx = context();
		std::lock_guard<std::mutex> guard(ctx.mutex);
		char* base = static_cast<char*>(noalloc_setup(ctx));
		const BatchMemory mem{reinterpret_cast<uint32_t*>(base), nullptr, reinterpret_cast<Quad*>(base + list_bytes), g_noalloc_host, false, false};
		simplify_batch_on_device(ctx, mem, destination, result_counts, result_errors, indices, index_offsets, vertex_positions, vertex_offsets,
		                         vertex_lock, vertex_count, vertex_positions_stride, target_index_counts, mesh_count, target_error, options, nullptr,
		                         nullptr, nullptr);
		t_last_error.clear();
		return 0;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return -1;
	}
}

extern "C" int meshopt_hip_batch_set_shape(const unsigned int* per_cu, const unsigned int* threads, const unsigned int* cursor)
{
	try
	{
		BatchShape t = batch_default_shape();
		if (per_cu && threads && cursor)
			for (int k = 0; k < kBatchClasses; ++k)
			{
				t.per_cu[k] = per_cu[k];
				t.threads[k] = threads[k];
				t.cursor[k] = cursor[k] != 0;
			}
		else if (per_cu || threads || cursor)
			throw std::invalid_argument("meshopt_hip_batch_set_shape: give all three arrays, or none for the default");
		if (const char* why = batch_shape_error(t))
			throw std::invalid_argument(std::string("meshopt_hip_batch_set_shape: ") + why);
		Context& ctx = context();
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.batch_shape = t;
		ctx.batch_shape_set = true;
		ctx.batch_cls_ready = false;
		t_last_error.clear();
		return 0;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return -1;
	}
}

extern "C" int meshopt_hip_batch_get_shape(unsigned int* per_cu, unsigned int* threads, unsigned int* cursor, unsigned int* vertex_caps,
                                           unsigned int* index_caps)
{
	try
	{
		Context& ctx = context();
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.ensure_stream();
		const BatchClasses& c = ensure_batch_classes(ctx);
		for (int k = 0; k < kBatchClasses; ++k)
		{
			if (per_cu)
				per_cu[k] = c.shape.per_cu[k];
			if (threads)
				threads[k] = c.shape.threads[k];
			if (cursor)
				cursor[k] = c.shape.cursor[k];
			if (vertex_caps)
				vertex_caps[k] = c.vcap[k];
			if (index_caps)
				index_caps[k] = c.icap[k];
		}
		t_last_error.clear();
		return 0;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return -1;
	}
}

extern "C" int meshopt_hip_last_batch_stats(unsigned int* stats, size_t max_meshes)
{
	try
	{
		Context& ctx = context();
		std::lock_guard<std::mutex> guard(ctx.mutex);
		const size_t n = std::min(max_meshes, ctx.batch_last_meshes);
		if (n)
		{
			HIP_CHECK(hipMemcpyAsync(stats, ctx.batch_dev + batch_list_words(ctx.batch_last_meshes), 4 * n * sizeof(uint32_t), hipMemcpyDeviceToHost,
			                         ctx.stream));
			HIP_CHECK(hipStreamSynchronize(ctx.stream));
		}
		t_last_error.clear();
		return (int)n;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return -1;
	}
}

// a batched call from host arrays on one context (the caller holds the context): stages the
// arrays, runs the batch and copies every mesh's result back; throws on invalid input
static void batch_host_on(Context& ctx, unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices,
                          const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets,
                          const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count,
                          const unsigned char* vertex_lock, size_t vertex_positions_stride, const unsigned int* target_index_counts,
                          size_t mesh_count, float target_error, unsigned int options)
{
	ctx.ensure_init();
	hipStream_t s = ctx.stream;
	size_t I = index_offsets[mesh_count], V = vertex_offsets[mesh_count];
	size_t pos_bytes = V ? (V - 1) * vertex_positions_stride + 12 : 0;
	size_t n1 = mesh_count + 1;
	size_t off_idx = 0, off_dst = (I * 4 + 255) & ~size_t(255);
	size_t off_io = off_dst + ((I * 4 + 255) & ~size_t(255));
	size_t off_vo = off_io + ((n1 * 4 + 255) & ~size_t(255));
	size_t off_tg = off_vo + ((n1 * 4 + 255) & ~size_t(255));
	size_t off_rc = off_tg + ((n1 * 4 + 255) & ~size_t(255));
	size_t off_re = off_rc + ((n1 * 4 + 255) & ~size_t(255));
	size_t off_pos = off_re + ((n1 * 4 + 255) & ~size_t(255));
	size_t off_lock = off_pos + ((pos_bytes + 255) & ~size_t(255));
	size_t lock_bytes = vertex_lock ? V : 0;
	if (attribute_count && (attribute_count > (size_t)kMaxAttributes || vertex_attributes_stride < attribute_count * 4 ||
	                        vertex_attributes_stride > 256 || vertex_attributes_stride % 4 != 0))
		throw std::invalid_argument("meshopt_hip_simplify_batch: invalid attribute count or stride");
	size_t attr_bytes = attribute_count && V ? (V - 1) * vertex_attributes_stride + attribute_count * 4 : 0;
	size_t off_attr = off_lock + ((lock_bytes + 255) & ~size_t(255));
	char* d = static_cast<char*>(ctx.staging.get(off_attr + attr_bytes + 256));
	if (I)
		HIP_CHECK(hipMemcpyAsync(d + off_idx, indices, I * 4, hipMemcpyHostToDevice, s));
	HIP_CHECK(hipMemcpyAsync(d + off_io, index_offsets, n1 * 4, hipMemcpyHostToDevice, s));
	HIP_CHECK(hipMemcpyAsync(d + off_vo, vertex_offsets, n1 * 4, hipMemcpyHostToDevice, s));
	if (mesh_count)
		HIP_CHECK(hipMemcpyAsync(d + off_tg, target_index_counts, mesh_count * 4, hipMemcpyHostToDevice, s));
	if (pos_bytes)
		HIP_CHECK(hipMemcpyAsync(d + off_pos, vertex_positions, pos_bytes, hipMemcpyHostToDevice, s));
	if (lock_bytes)
		HIP_CHECK(hipMemcpyAsync(d + off_lock, vertex_lock, lock_bytes, hipMemcpyHostToDevice, s));
	if (attr_bytes)
		HIP_CHECK(hipMemcpyAsync(d + off_attr, vertex_attributes, attr_bytes, hipMemcpyHostToDevice, s));
	const BatchAttributes attrs{reinterpret_cast<const float*>(d + off_attr), vertex_attributes_stride, attribute_weights, attribute_count};
	simplify_batch_on_device(ctx, context_batch_memory(ctx, mesh_count, V, !(options & (kOptSparse | kOptPrune)), attribute_count),
	                         reinterpret_cast<unsigned int*>(d + off_dst),
	                         reinterpret_cast<unsigned int*>(d + off_rc),
	                         reinterpret_cast<float*>(d + off_re), reinterpret_cast<const unsigned int*>(d + off_idx),
	                         reinterpret_cast<const unsigned int*>(d + off_io), reinterpret_cast<const float*>(d + off_pos),
	                         reinterpret_cast<const unsigned int*>(d + off_vo), lock_bytes ? reinterpret_cast<const unsigned char*>(d + off_lock) : nullptr,
	                         V, vertex_positions_stride,
	                         reinterpret_cast<const unsigned int*>(d + off_tg), mesh_count, target_error, options, index_offsets, vertex_offsets,
	                         target_index_counts, nullptr, attribute_count ? &attrs : nullptr);
	if (mesh_count)
	{
		HIP_CHECK(hipMemcpyAsync(result_counts, d + off_rc, mesh_count * 4, hipMemcpyDeviceToHost, s));
		if (result_errors)
			HIP_CHECK(hipMemcpyAsync(result_errors, d + off_re, mesh_count * 4, hipMemcpyDeviceToHost, s));
	}
	if (I)
		HIP_CHECK(hipMemcpyAsync(destination, d + off_dst, I * 4, hipMemcpyDeviceToHost, s));
	HIP_CHECK(hipStreamSynchronize(s));
}

static int batch_host_entry(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices,
                            const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets,
                            const float* vertex_attributes, size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count,
                            const unsigned char* vertex_lock, size_t vertex_positions_stride, const unsigned int* target_index_counts,
                            size_t mesh_count, float target_error, unsigned int options)
{
	try
	{
		// the offsets size every staging copy and bound every kernel's reads and writes: all of
		// them are checked here, before any copy or launch
		if (vertex_positions_stride < 12 || vertex_positions_stride > 256 || vertex_positions_stride % 4)
			throw std::invalid_argument("meshopt_hip_simplify_batch: invalid vertex stride");
		for (size_t p = 0; p < mesh_count; ++p)
			if (index_offsets[p + 1] < index_offsets[p] || vertex_offsets[p + 1] < vertex_offsets[p] ||
			    (index_offsets[p + 1] - index_offsets[p]) % 3 || target_index_counts[p] > index_offsets[p + 1] - index_offsets[p])
				throw std::invalid_argument("meshopt_hip_simplify_batch: invalid index count, target or offsets");
		Context& ctx = context();
		std::lock_guard<std::mutex> guard(ctx.mutex);
		batch_host_on(ctx, destination, result_counts, result_errors, indices, index_offsets, vertex_positions, vertex_offsets, vertex_attributes,
		              vertex_attributes_stride, attribute_weights, attribute_count, vertex_lock, vertex_positions_stride, target_index_counts,
		              mesh_count, target_error, options);
		t_last_error.clear();
		return 0;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return -1;
	}
}

extern "C" int meshopt_hip_simplify_batch(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices,
                                          const unsigned int* index_offsets, const float* vertex_positions, const unsigned int* vertex_offsets,
                                          const unsigned char* vertex_lock, size_t vertex_positions_stride, const unsigned int* target_index_counts,
                                          size_t mesh_count, float target_error, unsigned int options)
{
	return batch_host_entry(destination, result_counts, result_errors, indices, index_offsets, vertex_positions, vertex_offsets, nullptr, 0, nullptr, 0,
	                        vertex_lock, vertex_positions_stride, target_index_counts, mesh_count, target_error, options);
}

extern "C" int meshopt_hip_simplify_batch_with_attributes(unsigned int* destination, unsigned int* result_counts, float* result_errors,
                                                          const unsigned int* indices, const unsigned int* index_offsets, const float* vertex_positions,
                                                          const unsigned int* vertex_offsets, const float* vertex_attributes,
                                                          size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count,
                                                          const unsigned char* vertex_lock, size_t vertex_positions_stride,
                                                          const unsigned int* target_index_counts, size_t mesh_count, float target_error,
                                                          unsigned int options)
{
	return batch_host_entry(destination, result_counts, result_errors, indices, index_offsets, vertex_positions, vertex_offsets, vertex_attributes,
	                        vertex_attributes_stride, attribute_weights, attribute_count, vertex_lock, vertex_positions_stride, target_index_counts,
	                        mesh_count, target_error, options);
}

// The fallbacks of clodBuild's per-group simplify (demo/clusterlod.h:475-486) for the groups left
// above their target by the level's batch: with `permissive`, those groups run again as one
// sub-batch with Permissive added; then, with config.fallback_sloppy, every group still above its
// target takes meshopt_simplifySloppy over its deindexed triangles (simplifyFallback,
// demo/clusterlod.h:426-458: one vertex per index, the group's locks, FLT_MAX error), with the
// error made absolute by the subset's extent and scaled by config.error_factor_sloppy.  Groups
// whose target exceeds their size (tg[g] > size) have kept their indices and take no fallback.
static void clod_fallbacks(Context& ctx, const ClodBuffers& b, const meshopt_hip_clod_config& config, const std::vector<uint32_t>& io,
                           const std::vector<uint32_t>& tg, unsigned int* result_counts, float* result_errors, float target_error,
                           unsigned int batch_options, bool permissive, uint32_t* d_list, uint32_t* sub_io, uint32_t* sub_vo, uint32_t* sub_tg,
                           uint32_t* sub_idx, float* sub_pos, uint8_t* sub_locks, uint32_t* sub_result, uint32_t* sub_counts, float* sub_errors,
                           float* sub_attr, const float* attribute_weights, std::vector<uint32_t>& cnt, std::vector<uint32_t>& vo, bool have_counts)
{
	hipStream_t s = ctx.stream;
	const size_t G = tg.size();
	// the result counts and packed vertex offsets, unless they came back with the batch's status
	if (!have_counts)
	{
		HIP_CHECK(hipMemcpyAsync(cnt.data(), result_counts, 4 * G, hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipMemcpyAsync(vo.data(), b.unique, 4 * (G + 1), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
	}
	std::vector<uint32_t> fail;
	for (size_t g = 0; g < G; ++g)
		if (tg[g] <= io[g + 1] - io[g] && cnt[g] > tg[g])
			fail.push_back((uint32_t)g);
	if (fail.empty())
		return;

	if (permissive)
	{
		const size_t F = fail.size();
		std::vector<uint32_t> h_io(F + 1), h_vo(F + 1), h_tg(F), h_cnt(F);
		for (size_t f = 0; f < F; ++f)
		{
			const uint32_t g = fail[f];
			h_io[f + 1] = h_io[f] + (io[g + 1] - io[g]);
			h_vo[f + 1] = h_vo[f] + (vo[g + 1] - vo[g]);
			h_tg[f] = tg[g];
		}
		HIP_CHECK(hipMemcpyAsync(d_list, fail.data(), 4 * F, hipMemcpyHostToDevice, s));
		HIP_CHECK(hipMemcpyAsync(sub_io, h_io.data(), 4 * (F + 1), hipMemcpyHostToDevice, s));
		HIP_CHECK(hipMemcpyAsync(sub_vo, h_vo.data(), 4 * (F + 1), hipMemcpyHostToDevice, s));
		HIP_CHECK(hipMemcpyAsync(sub_tg, h_tg.data(), 4 * F, hipMemcpyHostToDevice, s));
		launch_clod_gather(s, b, d_list, (uint32_t)F, sub_io, sub_vo, sub_idx, sub_pos, sub_locks, sub_attr);
		const size_t A = b.attribute_count;
		const BatchAttributes attrs{sub_attr, A * sizeof(float), attribute_weights, A};
		simplify_batch_on_device(ctx, context_batch_memory(ctx, F, h_vo[F], !(batch_options & kOptPrune), A), sub_result, sub_counts,
		                         result_errors ? sub_errors : nullptr, sub_idx, sub_io, sub_pos, sub_vo, sub_locks, h_vo[F], 12, sub_tg, F, target_error,
		                         batch_options | kOptPermissive, nullptr, nullptr, nullptr, h_io.data(), A ? &attrs : nullptr);
		launch_clod_scatter(s, b, d_list, (uint32_t)F, sub_io, sub_result, sub_counts, result_errors ? sub_errors : nullptr, result_counts,
		                    result_errors);
		if (!config.fallback_sloppy)
			return;
		HIP_CHECK(hipMemcpyAsync(h_cnt.data(), sub_counts, 4 * F, hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		size_t still = 0;
		for (size_t f = 0; f < F; ++f)
			if (h_cnt[f] > h_tg[f])
				fail[still++] = fail[f];
		fail.resize(still);
	}

	// one group at a time: its subset lives in the sub-batch's buffers (a group has at most I
	// indices); the sloppy call's own buffers are the context's arena
	for (uint32_t g : fail)
	{
		const uint32_t n = io[g + 1] - io[g];
		launch_clod_deindex(s, b, g, n, sub_pos, sub_locks, sub_idx);
		float error = 0.f, extent = 0.f;
		const size_t count = sloppy_on_device(ctx, sub_result, sub_idx, n, sub_pos, n, 12, sub_locks, tg[g], FLT_MAX, &error, &extent);
		error *= extent;
		error *= config.error_factor_sloppy;
		launch_clod_sloppy_result(s, b, g, sub_result, (uint32_t)count, result_counts, result_errors, error);
	}
}

extern "C" int meshopt_hip_clod_level_device(unsigned int* destination, unsigned int* result_counts, float* result_errors, const unsigned int* indices,
                                             const unsigned int* index_offsets, size_t group_count, const float* vertex_positions,
                                             size_t vertex_count, size_t vertex_positions_stride, const float* vertex_attributes,
                                             size_t vertex_attributes_stride, const float* attribute_weights, size_t attribute_count,
                                             const unsigned int* remap,
                                             const unsigned char* base_locks, unsigned char* locks, const unsigned int* target_index_counts,
                                             float target_error, unsigned int options, const meshopt_hip_clod_config* config)
{
	try
	{
		if (vertex_positions_stride < 12 || vertex_positions_stride > 256 || vertex_positions_stride % 4)
			throw std::invalid_argument("meshopt_hip_clod_level_device: invalid vertex stride");
		if (!locks || !remap)
			throw std::invalid_argument("meshopt_hip_clod_level_device: remap and locks are required");
		if (vertex_count >= 0xffffffffu)
			throw std::invalid_argument("meshopt_hip_clod_level_device: too many vertices");
		if (attribute_count && (attribute_count > (size_t)kMaxAttributes || vertex_attributes_stride < attribute_count * 4 ||
		                        vertex_attributes_stride > 256 || vertex_attributes_stride % 4 != 0 || !vertex_attributes || !attribute_weights))
			throw std::invalid_argument("meshopt_hip_clod_level_device: invalid attribute count, stride or weights");
		Context& ctx = context();
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.ensure_init();
		hipStream_t s = ctx.stream;
		const size_t G = group_count;

		// the group offsets decide every buffer size and bound every probe loop, the targets which
		// groups keep their indices and which take a fallback: both read on the host
		std::vector<uint32_t> io(G + 1), tg(G);
		HIP_CHECK(hipMemcpyAsync(io.data(), index_offsets, (G + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
		if (G)
			HIP_CHECK(hipMemcpyAsync(tg.data(), target_index_counts, G * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		if (io[0] != 0)
			throw std::invalid_argument("meshopt_hip_clod_level_device: index_offsets[0] must be 0");
		for (size_t g = 0; g < G; ++g)
			if (io[g + 1] < io[g] || (io[g + 1] - io[g]) % 3)
				throw std::invalid_argument("meshopt_hip_clod_level_device: invalid group index counts");
		const size_t I = io[G];
		uint32_t max_group = 0;
		for (size_t g = 0; g < G; ++g)
			max_group = std::max(max_group, io[g + 1] - io[g]);

		// carve the level's buffers (256-byte aligned)
		size_t used = 0;
		auto take = [&](size_t bytes) {
			size_t off = (used + 255) & ~size_t(255);
			used = off + bytes;
			return off;
		};
		const size_t V = vertex_count;
		const size_t o_first = take(4 * (V + 1)), o_last = take(4 * (V + 1)), o_key = take(8 * I + 4), o_sfirst = take(8 * I + 4);
		const size_t o_local = take(8 * I + 4), o_staged = take(4 * I + 4), o_unique = take(4 * (G + 2)), o_order = take(4 * I + 4), o_pidx = take(4 * I + 4);
		const size_t A = attribute_count;
		const size_t o_ppos = take(12 * I + 4), o_plock = take(I + 4), o_pres = take(4 * I + 4), o_pattr = take(A ? 4 * A * I + 4 : 0);
		const size_t o_temp = take(4 * (scan_temp_words(G + 1) + 64)), o_flag = take(16);
		// a group whose target exceeds its size keeps its indices (demo/clusterlod.h:463-464): the
		// batch runs it at its size and its result is replaced afterwards
		std::vector<uint32_t> keep, clamped;
		for (size_t g = 0; g < G; ++g)
			if (tg[g] > io[g + 1] - io[g])
				keep.push_back((uint32_t)g);
		if (!keep.empty())
		{
			clamped = tg;
			for (uint32_t g : keep)
				clamped[g] = io[g + 1] - io[g];
		}
		const bool want_permissive = config && config->fallback_permissive && !(options & kOptPermissive);
		const bool want_sloppy = config && config->fallback_sloppy;
		const bool fallbacks = want_permissive || want_sloppy;
		// the fallbacks' sub-batch (at most every group: I indices, I packed vertices), reused by the
		// sloppy fallback's deindexed subsets (one group at a time)
		const size_t o_tg = take(keep.empty() ? 0 : 4 * G), o_list = take(4 * (G + 1));
		const size_t o_sio = take(fallbacks ? 4 * (G + 1) : 0), o_svo = take(fallbacks ? 4 * (G + 1) : 0), o_stg = take(fallbacks ? 4 * G : 0);
		const size_t o_sidx = take(fallbacks ? 4 * I + 4 : 0), o_spos = take(fallbacks ? 12 * I + 4 : 0), o_slock = take(fallbacks ? I + 4 : 0);
		const size_t o_sres = take(fallbacks ? 4 * I + 4 : 0), o_scnt = take(fallbacks ? 4 * G : 0), o_serr = take(fallbacks ? 4 * G : 0);
		const size_t o_sattr = take(fallbacks && A ? 4 * A * I + 4 : 0);
		char* base = static_cast<char*>(ctx.ensure_clod(used + 256));

		ClodBuffers b;
		b.indices = indices;
		b.index_offsets = index_offsets;
		b.remap = remap;
		b.group_count = (uint32_t)G;
		b.vertex_count = (uint32_t)V;
		b.lds_slots = clod_lds_slots(max_group);
		const bool large_groups = 2 * (size_t)max_group > b.lds_slots;
		b.first_group = reinterpret_cast<uint32_t*>(base + o_first);
		b.last_group = reinterpret_cast<uint32_t*>(base + o_last);
		b.slot_key = reinterpret_cast<uint32_t*>(base + o_key);
		b.slot_first = reinterpret_cast<uint32_t*>(base + o_sfirst);
		b.slot_local = reinterpret_cast<uint32_t*>(base + o_local);
		b.staged = reinterpret_cast<uint32_t*>(base + o_staged);
		b.unique = reinterpret_cast<uint32_t*>(base + o_unique);
		b.order = reinterpret_cast<uint32_t*>(base + o_order);
		b.packed_indices = reinterpret_cast<uint32_t*>(base + o_pidx);
		b.packed_positions = reinterpret_cast<float*>(base + o_ppos);
		b.packed_locks = reinterpret_cast<uint8_t*>(base + o_plock);
		b.packed_result = reinterpret_cast<uint32_t*>(base + o_pres);
		b.packed_attributes = A ? reinterpret_cast<float*>(base + o_pattr) : nullptr;
		b.attribute_count = (uint32_t)A;
		uint32_t* temp = reinterpret_cast<uint32_t*>(base + o_temp);
		uint32_t* flag = reinterpret_cast<uint32_t*>(base + o_flag);

		// every index below vertex_count before anything reads through it
		HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(uint32_t), s));
		if (I)
		{
			launch_check_indices(s, indices, I, V, flag);
			uint32_t bad = 0;
			HIP_CHECK(hipMemcpyAsync(&bad, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
			HIP_CHECK(hipStreamSynchronize(s));
			if (bad)
				throw std::invalid_argument("meshopt_hip_clod_level_device: index buffer references a vertex >= vertex_count");
		}

		if (V)
		{
			HIP_CHECK(hipMemsetAsync(b.first_group, 0xff, 4 * V, s));
			HIP_CHECK(hipMemsetAsync(b.last_group, 0, 4 * V, s));
		}
		if (large_groups)
		{
			HIP_CHECK(hipMemsetAsync(b.slot_key, 0, 8 * I, s));
			HIP_CHECK(hipMemsetAsync(b.slot_first, 0xff, 8 * I, s));
		}
		HIP_CHECK(hipMemsetAsync(b.unique, 0, 4 * (G + 1), s));
		launch_clod_group(s, b, large_groups);
		launch_clod_locks(s, b, base_locks, locks);
		scan_exclusive(s, b.unique, b.unique, G, nullptr, b.unique + G, temp);
		launch_clod_place(s, b, vertex_positions, vertex_positions_stride / sizeof(float), locks, vertex_attributes, vertex_attributes_stride / sizeof(float));

		// the packed groups are the meshes of one batched call (Sparse is what the packing did)
		const unsigned int* batch_targets = target_index_counts;
		if (!keep.empty())
		{
			uint32_t* d_tg = reinterpret_cast<uint32_t*>(base + o_tg);
			HIP_CHECK(hipMemcpyAsync(d_tg, clamped.data(), 4 * G, hipMemcpyHostToDevice, s));
			batch_targets = d_tg;
		}
		const unsigned int batch_options = options & ~(unsigned int)kOptSparse;
		const BatchAttributes attrs{b.packed_attributes, A * sizeof(float), attribute_weights, A};
		// the fallbacks need the result counts and the packed vertex offsets: they come back with
		// the batch's status read-back (the kept groups' counts, rewritten below, are not used)
		std::vector<uint32_t> h_cnt(fallbacks ? G : 0), h_vo(fallbacks ? G + 1 : 0);
		BatchReadback rb[2] = {{h_cnt.data(), result_counts, 4 * G, false}, {h_vo.data(), b.unique, 4 * (G + 1), false}};
		simplify_batch_on_device(ctx, context_batch_memory(ctx, G, I, !(options & kOptPrune), A), b.packed_result, result_counts, result_errors,
		                         b.packed_indices, index_offsets, b.packed_positions, b.unique,
		                         b.packed_locks, I, 12, batch_targets, G, target_error, batch_options, nullptr, nullptr,
		                         nullptr, io.data(), A ? &attrs : nullptr, fallbacks ? rb : nullptr, fallbacks ? 2 : 0);
		uint32_t* d_list = reinterpret_cast<uint32_t*>(base + o_list);
		if (!keep.empty())
		{
			HIP_CHECK(hipMemcpyAsync(d_list, keep.data(), 4 * keep.size(), hipMemcpyHostToDevice, s));
			launch_clod_keep(s, b, d_list, (uint32_t)keep.size(), result_counts, result_errors);
		}
		if (fallbacks)
			clod_fallbacks(ctx, b, *config, io, tg, result_counts, result_errors, target_error, batch_options, want_permissive, d_list,
			               reinterpret_cast<uint32_t*>(base + o_sio), reinterpret_cast<uint32_t*>(base + o_svo),
			               reinterpret_cast<uint32_t*>(base + o_stg), reinterpret_cast<uint32_t*>(base + o_sidx),
			               reinterpret_cast<float*>(base + o_spos), reinterpret_cast<uint8_t*>(base + o_slock),
			               reinterpret_cast<uint32_t*>(base + o_sres), reinterpret_cast<uint32_t*>(base + o_scnt),
			               reinterpret_cast<float*>(base + o_serr), A ? reinterpret_cast<float*>(base + o_sattr) : nullptr, attribute_weights, h_cnt, h_vo,
			               rb[0].valid && rb[1].valid);
		if (config && config->error_edge_limit > 0 && result_errors)
			launch_clod_edge_limit(s, b, target_index_counts, config->error_edge_limit, result_errors);
		launch_clod_unpack(s, b, result_counts, destination);
		HIP_CHECK(hipStreamSynchronize(s));
		t_last_error.clear();
		return 0;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return -1;
	}
}

extern "C" float meshopt_simplifyScale(const float* vertex_positions, size_t vertex_count, size_t vertex_positions_stride)
{
	try
	{
		PooledContext pooled;
		Context& ctx = *pooled.ctx;
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.ensure_init();
		hipStream_t s = ctx.stream;
		size_t span = vertex_count ? (vertex_count - 1) * vertex_positions_stride + 12 : 0;
		size_t off_p = (span + 255) & ~size_t(255);
		size_t off_sc = (off_p + (vertex_count + 1) * sizeof(float4) + 255) & ~size_t(255);
		char* d = static_cast<char*>(ctx.staging.get(off_sc + sizeof(CallScalars) + 256));
		if (span)
			HIP_CHECK(hipMemcpyAsync(d, vertex_positions, span, hipMemcpyHostToDevice, s));
		Mesh m{};
		m.praw = reinterpret_cast<float4*>(d + off_p);
		m.sc = reinterpret_cast<CallScalars*>(d + off_sc);
		m.vertex_count = (uint32_t)vertex_count;
		launch_gather_positions(s, reinterpret_cast<const float*>(d), vertex_positions_stride / 4, nullptr, m.praw, vertex_count);
		launch_extent_only(s, m);
		HIP_CHECK(hipMemcpyAsync(ctx.sc_host, m.sc, sizeof(CallScalars), hipMemcpyDeviceToHost, s));
		HIP_CHECK(hipStreamSynchronize(s));
		return ctx.sc_host->extent;
	}
	catch (const std::exception& e)
	{
		die(e.what());
	}
}

// the callbacks meshopt_setAllocator sets, exported for the header's index adapters
// (src/allocator.cpp:6-12 under MESHOPTIMIZER_ALLOC_EXPORT)
meshopt_Allocator::Storage& meshopt_Allocator::storage()
{
	static Storage s = {::operator new, ::operator delete};
	return s;
}

extern "C" void meshopt_setAllocator(void* (*allocate)(size_t), void (*deallocate)(void*))
{
	meshopt_Allocator::Storage& s = meshopt_Allocator::storage();
	s.allocate = allocate ? allocate : static_cast<void* (*)(size_t)>(::operator new);
	s.deallocate = deallocate ? deallocate : static_cast<void (*)(void*)>(::operator delete);
}

extern "C" int meshopt_hip_set_round_policy(int deep_heads, int deep_from_head)
{
	if (deep_from_head < 0 || deep_from_head > kMaxHeadRounds)
	{
		t_last_error = "meshopt_hip_set_round_policy: deep_from_head out of range";
		return -1;
	}
	g_deep_heads.store(deep_heads < -1 ? kDeepHeads : deep_heads);
	g_deep_from_head.store(deep_from_head);
	t_last_error.clear();
	return 0;
}

extern "C" int meshopt_hip_reset_round_counts(void)
{
	for (int b = 0; b < 33; ++b)
	{
		g_head_rounds[b].store(0, std::memory_order_relaxed);
		g_ext_head_rounds[b].store(0, std::memory_order_relaxed);
	}
	t_last_error.clear();
	return 0;
}

extern "C" int meshopt_hip_last_call_stats(meshopt_hip_call_stats* call, meshopt_hip_pass_stats* passes, int max_passes)
{
	if (call)
		*call = t_call;
	int n = (int)t_passes.size();
	for (int i = 0; i < n && i < max_passes && passes; ++i)
		passes[i] = t_passes[i];
	return n;
}

extern "C" void meshopt_hip_batch_kernel_timing(double* total_ms, size_t* launches, size_t* meshes)
{
	TimingTotals& t = timing_totals();
	std::lock_guard<std::mutex> guard(t.m);
	if (total_ms)
		*total_ms = t.batch_ms;
	if (launches)
		*launches = t.batch_launches;
	if (meshes)
		*meshes = t.batch_meshes;
	t.batch_ms = 0;
	t.batch_launches = 0;
	t.batch_meshes = 0;
}

extern "C" void meshopt_hip_kernel_timing(int enable, double* total_ms, size_t* launches, size_t* ranks)
{
	TimingTotals& t = timing_totals();
	std::lock_guard<std::mutex> guard(t.m);
	if (total_ms)
		*total_ms = t.x_ms;
	if (launches)
		*launches = t.x_launches;
	if (ranks)
		*ranks = t.x_visits;
	t.on = enable != 0;
	t.x_ms = 0;
	t.x_launches = 0;
	t.x_visits = 0;
}

extern "C" const char* meshopt_hip_last_error(void)
{
	return t_last_error.c_str();
}

extern "C" void meshopt_hip_release(void)
{
	{
		Context& ctx = context();
		std::lock_guard<std::mutex> guard(ctx.mutex);
		ctx.release();
	}
	pool().release_idle();
}

// ------------------------------------------------------------------------------------------
// serving: a request loop over several contexts (one HIP stream, arena and staging buffer each)

namespace mhip
{

// device bytes a context holds for one host call of this shape (arena with its growth slack,
// staging buffer with its slack)
static size_t request_device_bytes(const meshopt_hip_request& r)
{
	size_t A = 0;
	for (size_t k = 0; k < r.attribute_count && k < (size_t)kMaxAttributes; ++k)
		A += r.attribute_weights[k] > 0;
	const bool sparse = (r.options & kOptSparse) != 0;
	Arena dry;
	Plan plan;
	carve(dry, plan, r.vertex_count, sparse ? std::min(r.vertex_count, r.index_count) : r.vertex_count, r.index_count, A, sparse,
	      (r.options & kOptPrune) != 0, false);
	size_t arena = dry.used + 4096;
	arena += arena / 4;
	const HostStaging h = plan_staging(r.index_count, r.vertex_positions, r.vertex_count, r.vertex_positions_stride, r.vertex_attributes,
	                                   r.vertex_attributes_stride, r.attribute_count, r.vertex_lock);
	return arena + h.total + h.total / 4 + 256;
}

struct ServerTicket
{
	meshopt_hip_request req{};
	std::vector<float> weights; // the request's weights, copied at submit
	size_t need = 0;
	double submit_ms = 0, start_ms = 0, finish_ms = 0;
	size_t result = 0;
	float error = 0;
	int slot = -1;
	int batched = 1; // requests in the call that ran this one
	bool done = false;
	bool claimed = false; // a thread waits on it (only one may: that thread frees it)
	std::string failure;
};

// requests the batched kernel takes whole (the largest LDS size class: 1984 vertices, 5952
// indices; a larger mesh would still be simplified exactly, by the plain path inside the call),
// without Sparse or Prune
constexpr size_t kServeBatchVertices = 1984, kServeBatchIndices = 5952;
// requests one batched call of the server takes at most, and queue entries it looks through
constexpr size_t kServeBatchMax = 4096, kServeScan = 8192;

static bool serve_batchable(const ServerTicket& t)
{
	const meshopt_hip_request& r = t.req;
	return !(r.options & (kOptSparse | kOptPrune)) && r.index_count <= kServeBatchIndices && r.vertex_count <= kServeBatchVertices &&
	       r.attribute_count <= (size_t)kMaxAttributes && t.weights.size() == r.attribute_count;
}

// one batched call takes one target error, one option set, one attribute count with one set of
// weights, and locks for all meshes or none
static bool serve_compatible(const ServerTicket& a, const ServerTicket& b)
{
	return a.req.target_error == b.req.target_error && a.req.options == b.req.options &&
	       (a.req.vertex_lock == nullptr) == (b.req.vertex_lock == nullptr) && a.req.attribute_count == b.req.attribute_count &&
	       a.weights == b.weights;
}

// small requests taken together from the queue: packed (positions as xyz, attributes as
// attribute_count floats per vertex, indices local to each request's vertices) and simplified by
// one batched call on the slot's context; each gets what the plain path returns for it alone.  Throws when the call fails (a bad index in any request); the
// worker then runs each request on its own.
static void serve_batch(Context& ctx, const std::vector<ServerTicket*>& g)
{
	const size_t P = g.size();
	std::vector<uint32_t> io(P + 1, 0), vo(P + 1, 0), tg(P), counts(P);
	std::vector<float> errs(P);
	for (size_t p = 0; p < P; ++p)
	{
		io[p + 1] = io[p] + (uint32_t)g[p]->req.index_count;
		vo[p + 1] = vo[p] + (uint32_t)g[p]->req.vertex_count;
		tg[p] = (uint32_t)g[p]->req.target_index_count;
	}
	const bool locks = g[0]->req.vertex_lock != nullptr;
	const size_t A = g[0]->req.attribute_count;
	std::vector<uint32_t> idx(std::max<size_t>(io[P], 1)), dst(std::max<size_t>(io[P], 1));
	std::vector<float> pos(3 * std::max<size_t>(vo[P], 1)), attr(A * std::max<size_t>(vo[P], 1));
	std::vector<uint8_t> lock(locks ? std::max<size_t>(vo[P], 1) : 0);
	for (size_t p = 0; p < P; ++p)
	{
		const meshopt_hip_request& r = g[p]->req;
		if (r.index_count)
			memcpy(idx.data() + io[p], r.indices, r.index_count * sizeof(uint32_t));
		const char* src = reinterpret_cast<const char*>(r.vertex_positions);
		for (size_t v = 0; v < r.vertex_count; ++v)
			memcpy(pos.data() + 3 * (vo[p] + v), src + v * r.vertex_positions_stride, 3 * sizeof(float));
		if (locks && r.vertex_count)
			memcpy(lock.data() + vo[p], r.vertex_lock, r.vertex_count);
		const char* asrc = reinterpret_cast<const char*>(r.vertex_attributes);
		for (size_t v = 0; A && v < r.vertex_count; ++v)
			memcpy(attr.data() + A * (vo[p] + v), asrc + v * r.vertex_attributes_stride, A * sizeof(float));
	}
	batch_host_on(ctx, dst.data(), counts.data(), errs.data(), idx.data(), io.data(), pos.data(), vo.data(), A ? attr.data() : nullptr,
	              A * sizeof(float), A ? g[0]->weights.data() : nullptr, A, locks ? lock.data() : nullptr, 3 * sizeof(float), tg.data(), P,
	              g[0]->req.target_error, g[0]->req.options);
	for (size_t p = 0; p < P; ++p)
	{
		if (counts[p])
			memcpy(g[p]->req.destination, dst.data() + io[p], counts[p] * sizeof(uint32_t));
		g[p]->result = counts[p];
		g[p]->error = errs[p];
		g[p]->batched = (int)P;
	}
}

// Requests are admitted in arrival order (FIFO, no overtaking) to the first idle slot whose
// memory, next to what the other slots hold, fits the budget.  A slot keeps its arena between
// requests unless it holds more than budget / slots; idle slots give theirs up when the head of
// the queue needs the room.  Slots run concurrently, each on its own HIP stream, so the device
// overlaps their work.  A small request at the head takes the compatible small requests queued
// behind it (while their memory fits) into one batched call on its slot.
struct Server
{
	size_t budget = 0;
	int slots = 0;
	int device = 0;
	std::vector<std::unique_ptr<Context>> ctxs;
	std::vector<size_t> held;
	std::vector<char> busy;
	std::vector<std::thread> workers;
	std::mutex m;
	std::condition_variable cv_work, cv_done;
	std::deque<uint64_t> queue;
	std::unordered_map<uint64_t, std::unique_ptr<ServerTicket>> tickets;
	uint64_t next_id = 1;
	bool stopping = false;
	std::chrono::steady_clock::time_point epoch = std::chrono::steady_clock::now();

	double now_ms() const
	{
		return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - epoch).count();
	}

	// the queue head fits on slot s (caller holds m)
	bool head_fits(int s) const
	{
		const ServerTicket& t = *tickets.at(queue.front());
		size_t others = 0;
		for (int k = 0; k < slots; ++k)
			if (k != s)
				others += held[k];
		return others + std::max(held[s], t.need) <= budget;
	}

	void worker(int s)
	{
		(void)hipSetDevice(device);
		Context& ctx = *ctxs[s];
		std::unique_lock<std::mutex> lk(m);
		for (;;)
		{
			cv_work.wait(lk, [&] { return stopping || !queue.empty(); });
			if (queue.empty())
			{
				if (stopping)
					break;
				continue;
			}
			if (!head_fits(s))
			{
				// make room from idle slots (this one included), then wait for running ones
				for (int k = 0; k < slots; ++k)
					if (!busy[k] && held[k])
					{
						ctxs[k]->release_memory();
						held[k] = 0;
					}
				if (!head_fits(s))
				{
					cv_work.wait(lk);
					continue;
				}
			}
			const uint64_t id = queue.front();
			queue.pop_front();
			ServerTicket& t = *tickets[id];
			busy[s] = 1;
			t.start_ms = now_ms();
			t.slot = s;
			std::vector<ServerTicket*> group{&t};
			size_t need = t.need;
			if (serve_batchable(t))
			{
				size_t others = 0;
				for (int k = 0; k < slots; ++k)
					if (k != s)
						others += held[k];
				// compatible requests from anywhere in the first kServeScan entries, in queue order;
				// the others keep their order
				for (size_t i = 0, seen = 0; i < queue.size() && seen < kServeScan && group.size() < kServeBatchMax; ++seen)
				{
					ServerTicket& u = *tickets[queue[i]];
					if (!serve_batchable(u) || !serve_compatible(t, u))
					{
						++i;
						continue;
					}
					if (others + std::max(held[s], need + u.need) > budget)
						break;
					queue.erase(queue.begin() + (std::ptrdiff_t)i);
					need += u.need;
					u.start_ms = t.start_ms;
					u.slot = s;
					group.push_back(&u);
				}
			}
			held[s] = std::max(held[s], need);
			lk.unlock();

			auto run_alone = [&](ServerTicket& r1) {
				const meshopt_hip_request& r = r1.req;
				try
				{
					r1.result = simplify_host_on(ctx, r.indices, r.destination, r.index_count, r.vertex_positions, r.vertex_count,
					                             r.vertex_positions_stride, r.attribute_count ? r.vertex_attributes : nullptr,
					                             r.vertex_attributes_stride, r1.weights.empty() ? nullptr : r1.weights.data(), r.attribute_count,
					                             r.vertex_lock, r.target_index_count, r.target_error, r.options, &r1.error);
				}
				catch (const std::exception& e)
				{
					r1.result = 0;
					r1.error = 0;
					r1.failure = e.what();
				}
			};
			bool ran = false;
			if (group.size() > 1)
			{
				try
				{
					serve_batch(ctx, group);
					ran = true;
				}
				catch (const std::exception&)
				{
					// some request is invalid for the device (a bad index): each runs on its own,
					// so that only the invalid one fails
				}
			}
			if (!ran)
				for (ServerTicket* u : group)
					run_alone(*u);
			size_t now_held = ctx.held_bytes();
			if (now_held > budget / (size_t)slots)
			{
				ctx.release_memory();
				now_held = 0;
			}

			lk.lock();
			held[s] = now_held;
			busy[s] = 0;
			const double finish = now_ms();
			for (ServerTicket* u : group)
			{
				u->finish_ms = finish;
				u->done = true;
			}
			cv_done.notify_all();
			cv_work.notify_all();
		}
	}
};

} // namespace mhip

extern "C" meshopt_hip_server* meshopt_hip_server_create(size_t memory_budget_bytes, int slots)
{
	try
	{
		if (slots < 1 || slots > 64 || memory_budget_bytes == 0)
			throw std::invalid_argument("meshopt_hip_server_create: slots must be in [1, 64] and the budget positive");
		std::unique_ptr<Server> srv(new Server());
		srv->budget = memory_budget_bytes;
		srv->slots = slots;
		HIP_CHECK(hipGetDevice(&srv->device));
		for (int s = 0; s < slots; ++s)
		{
			srv->ctxs.emplace_back(new Context());
			srv->ctxs.back()->ensure_init();
		}
		srv->held.assign(slots, 0);
		srv->busy.assign(slots, 0);
		for (int s = 0; s < slots; ++s)
			srv->workers.emplace_back(&Server::worker, srv.get(), s);
		t_last_error.clear();
		return reinterpret_cast<meshopt_hip_server*>(srv.release());
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return nullptr;
	}
}

extern "C" size_t meshopt_hip_server_submit(meshopt_hip_server* server, const meshopt_hip_request* request)
{
	Server* srv = reinterpret_cast<Server*>(server);
	try
	{
		if (!srv || !request)
			throw std::invalid_argument("meshopt_hip_server_submit: null server or request");
		const meshopt_hip_request& r = *request;
		check_args(r.index_count, r.vertex_positions_stride, r.target_index_count, r.target_error, r.vertex_attributes_stride, r.attribute_count,
		           r.attribute_weights);
		if (r.options & kOptInternalSolve)
			throw std::invalid_argument("meshopt_hip_server_submit: the solve bit is not accepted here");
		std::unique_ptr<ServerTicket> t(new ServerTicket());
		t->req = r;
		if (r.attribute_count)
			t->weights.assign(r.attribute_weights, r.attribute_weights + r.attribute_count);
		t->req.attribute_weights = t->weights.empty() ? nullptr : t->weights.data();
		t->need = request_device_bytes(t->req);
		if (t->need > srv->budget)
			throw std::invalid_argument("meshopt_hip_server_submit: the request needs more device memory than the server's budget");
		std::lock_guard<std::mutex> lk(srv->m);
		if (srv->stopping)
			throw std::runtime_error("meshopt_hip_server_submit: the server is shutting down");
		const uint64_t id = srv->next_id++;
		t->submit_ms = srv->now_ms();
		srv->tickets[id] = std::move(t);
		srv->queue.push_back(id);
		srv->cv_work.notify_all();
		t_last_error.clear();
		return id;
	}
	catch (const std::exception& e)
	{
		t_last_error = e.what();
		return 0;
	}
}

extern "C" int meshopt_hip_server_wait(meshopt_hip_server* server, size_t ticket, meshopt_hip_response* response)
{
	Server* srv = reinterpret_cast<Server*>(server);
	if (!srv)
	{
		t_last_error = "meshopt_hip_server_wait: null server";
		return -1;
	}
	std::unique_lock<std::mutex> lk(srv->m);
	auto it = srv->tickets.find(ticket);
	if (it == srv->tickets.end())
	{
		t_last_error = "meshopt_hip_server_wait: unknown ticket";
		return -1;
	}
	// the ticket object keeps its address; the map may rehash while this thread waits.  Only one
	// thread may wait on a ticket, since the waiter frees it
	ServerTicket* tp = it->second.get();
	if (tp->claimed)
	{
		t_last_error = "meshopt_hip_server_wait: another thread is already waiting on this ticket";
		return -1;
	}
	tp->claimed = true;
	srv->cv_done.wait(lk, [&] { return tp->done; });
	it = srv->tickets.find(ticket);
	std::unique_ptr<ServerTicket> t = std::move(it->second);
	srv->tickets.erase(it);
	lk.unlock();
	if (response)
	{
		response->index_count = t->failure.empty() ? t->result : (size_t)-1;
		response->result_error = t->error;
		response->slot = t->slot;
		response->submit_ms = t->submit_ms;
		response->start_ms = t->start_ms;
		response->finish_ms = t->finish_ms;
		response->device_bytes = t->need;
		response->batched = t->batched;
	}
	if (!t->failure.empty())
	{
		t_last_error = t->failure;
		return -1;
	}
	t_last_error.clear();
	return 0;
}

extern "C" double meshopt_hip_server_clock_ms(const meshopt_hip_server* server)
{
	const Server* srv = reinterpret_cast<const Server*>(server);
	return srv ? srv->now_ms() : 0.0;
}

extern "C" void meshopt_hip_server_destroy(meshopt_hip_server* server)
{
	Server* srv = reinterpret_cast<Server*>(server);
	if (!srv)
		return;
	{
		std::lock_guard<std::mutex> lk(srv->m);
		srv->stopping = true;
		srv->cv_work.notify_all();
	}
	for (auto& w : srv->workers)
		w.join();
	for (auto& c : srv->ctxs)
		c->release();
	delete srv;
}
