// Writes xorwow_vectors.json: draws of rocRAND's XORWOW (the generator behind
// hipRAND's default device state on ROCm) through its host implementation,
// for seeds and subsequences of the reference's initial load. The fixture
// pins the oracle's restatement of the generator (tests/test_oracle.py).
// Build and run with tests/golden/make_xorwow_vectors.py.
#include <rocrand/rocrand_xorwow.h>
#include <rocrand/rocrand_uniform.h>
#include <rocrand/rocrand_normal.h>

#include <cstdio>

int main() {
  const unsigned long long seeds[] = {12345ull, 12445ull, 0ull, (1ull << 40) + 7ull};
  const unsigned long long subsequences[] = {0ull, 1ull, 2ull, 5ull, 1000ull, 123456ull, (1ull << 31) + 3ull};
  std::printf("{\n  \"raw\": [\n");
  bool first = true;
  for (unsigned long long seed : seeds)
    for (unsigned long long sub : subsequences) {
      rocrand_state_xorwow st;
      rocrand_init(seed, sub, 0, &st);
      std::printf("%s    {\"seed\": %llu, \"subsequence\": %llu, \"draws\": [", first ? "" : ",\n", seed, sub);
      for (int i = 0; i < 16; ++i) std::printf("%s%u", i ? ", " : "", rocrand(&st));
      std::printf("]}");
      first = false;
    }
  std::printf("\n  ],\n  \"load_draws\": [\n");
  first = true;
  for (unsigned long long seed : {12345ull, 12445ull})
    for (unsigned long long sub : {0ull, 1ull, 63ull, 4096ull, 99999ull}) {
      rocrand_state_xorwow st;
      rocrand_init(seed, sub, 0, &st);
      const double u0 = rocrand_uniform_double(&st), u1 = rocrand_uniform_double(&st),
                   u2 = rocrand_uniform_double(&st);
      const double2 pair = rocrand_normal_double2(&st);
      const double n = rocrand_normal_double(&st);
      std::printf("%s    {\"seed\": %llu, \"subsequence\": %llu, \"values\": [\"%a\", \"%a\", \"%a\", \"%a\", \"%a\", \"%a\"]}",
                  first ? "" : ",\n", seed, sub, u0, u1, u2, pair.x, pair.y, n);
      first = false;
    }
  std::printf("\n  ]\n}\n");
  return 0;
}
