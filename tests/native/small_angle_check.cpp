// The tetrahedron polar decomposition narrows sin / cos of a promoted float
// half-angle (quat.h:83-88: (float)sin((double)theta)). For every float
// theta in [0, 2e-4) this checks, with the host's libm (the one the
// reference's build uses), that (float)sin((double)theta) == theta and
// (float)cos((double)theta) == 1.0f, which lets the device skip the double
// sincos there. Exits non-zero on the first counterexample.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

int main() {
	const float limit = 2.0e-4f;
	uint32_t end;
	std::memcpy(&end, &limit, 4);   // positive floats are ordered like their bit patterns
	const unsigned nthreads = std::max(1u, std::thread::hardware_concurrency());
	std::vector<uint64_t> bad(nthreads, 0), first(nthreads, ~0ull);
	std::vector<std::thread> pool;
	for (unsigned w = 0; w < nthreads; ++w)
		pool.emplace_back([&, w] {
			for (uint64_t b = w; b < end; b += nthreads) {
				const uint32_t bits = (uint32_t)b;
				float t;
				std::memcpy(&t, &bits, 4);
				const float s = (float)std::sin((double)t), c = (float)std::cos((double)t);
				if (std::memcmp(&s, &t, 4) != 0 || c != 1.0f) {
					if (!bad[w]) first[w] = b;
					++bad[w];
				}
			}
		});
	for (auto& th : pool) th.join();
	uint64_t total = 0, f = ~0ull;
	for (unsigned w = 0; w < nthreads; ++w) {
		total += bad[w];
		if (first[w] < f) f = first[w];
	}
	std::printf("floats checked %u, counterexamples %llu", end, (unsigned long long)total);
	if (total) std::printf(" (first bit pattern 0x%08llx)", (unsigned long long)f);
	std::printf("\n");
	return total ? 1 : 0;
}
