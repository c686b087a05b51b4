// meshopt_setAllocator through the C++ overloads of include/meshoptimizer.h, the pattern of the
// reference's customAllocator test (demo/tests.cpp:1712-1762): a counting allocator sees the
// blocks of the 16-bit index adapters, every block is freed before the call returns, 32-bit calls
// take no block, and after the default allocator is restored the counting one is never called.
// The 16-bit results must equal the 32-bit ones.  Prints "alloc ok" and exits 0 on success.
#include "meshoptimizer.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

static size_t alloc_count, free_count;

static void* counting_alloc(size_t size)
{
	alloc_count++;
	return malloc(size);
}

static void counting_free(void* ptr)
{
	free_count++;
	free(ptr);
}

#define CHECK(cond)                                                                  \
	do                                                                               \
	{                                                                                \
		if (!(cond))                                                                 \
		{                                                                            \
			fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond);       \
			return 1;                                                                \
		}                                                                            \
	} while (0)

int main()
{
	// a 16 x 16 grid of quads with a gentle bump: enough to simplify
	const int n = 16;
	std::vector<float> vb;
	for (int y = 0; y <= n; ++y)
		for (int x = 0; x <= n; ++x)
		{
			vb.push_back(float(x) / n);
			vb.push_back(float(y) / n);
			vb.push_back(0.05f * float((x * 7 + y * 3) % 5) / 5.f);
		}
	std::vector<unsigned int> ib;
	for (int y = 0; y < n; ++y)
		for (int x = 0; x < n; ++x)
		{
			unsigned int a = y * (n + 1) + x, b = a + 1, c = a + (n + 1), d = c + 1;
			unsigned int t[6] = {a, b, c, c, b, d};
			ib.insert(ib.end(), t, t + 6);
		}
	std::vector<unsigned short> ibs(ib.begin(), ib.end());
	const size_t vc = vb.size() / 3, ic = ib.size(), target = ic / 4 / 3 * 3;

	meshopt_setAllocator(counting_alloc, counting_free);
	CHECK(alloc_count == 0 && free_count == 0);

	std::vector<unsigned int> d32(ic);
	float e32 = 0;
	size_t n32 = meshopt_simplify(&d32[0], &ib[0], ic, &vb[0], vc, 12, target, 1e-1f, 0, &e32);
	CHECK(alloc_count == 0 && free_count == 0); // 32-bit indices: no adapter block

	std::vector<unsigned short> d16(ic);
	float e16 = 0;
	size_t n16 = meshopt_simplify(&d16[0], &ibs[0], ic, &vb[0], vc, 12, target, 1e-1f, 0, &e16);
	CHECK(alloc_count == 2 && free_count == 2); // the input and output adapters
	CHECK(n16 == n32 && n32 > 0 && n32 < ic && memcmp(&e16, &e32, sizeof(float)) == 0);
	for (size_t i = 0; i < n32; ++i)
		CHECK(d16[i] == d32[i]);

	std::vector<unsigned short> s16(ic);
	size_t ns = meshopt_simplifySloppy(&s16[0], &ibs[0], ic, &vb[0], vc, 12, target, 1e-1f);
	CHECK(alloc_count == 4 && free_count == 4 && ns <= ic);

	std::vector<unsigned short> u16(ibs);
	std::vector<float> vu(vb);
	size_t nu = meshopt_simplifyWithUpdate(&u16[0], ic, &vu[0], vc, 12, NULL, 0, NULL, 0, NULL, target, 1e-1f);
	CHECK(alloc_count == 5 && free_count == 5 && nu <= ic); // one adapter updated in place

	meshopt_setAllocator(::operator new, ::operator delete);
	meshopt_simplify(&d16[0], &ibs[0], ic, &vb[0], vc, 12, target, 1e-1f);
	CHECK(alloc_count == 5 && free_count == 5); // the counting allocator is no longer called

	printf("alloc ok: %zu -> %zu indices\n", ic, n32);
	return 0;
}
