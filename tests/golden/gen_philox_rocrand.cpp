// Writes Philox4x32-10 vectors from rocRAND's header implementation (an
// independent implementation of Salmon et al., SC'11) for pinning the oracle.
// Build and run: tests/golden/gen_philox_rocrand.sh
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct Engine : rocrand_device::philox4x32_10_engine {
  __host__ uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
  Engine e;
  unsigned s = 12345u;
  auto next = [&]() { s = s * 1664525u + 1013904223u; return s; };
  std::printf("[\n");
  for (int i = 0; i < 64; ++i) {
    uint4 c;
    uint2 k;
    if (i < 32) {  // counters laid out as (voxel/4, field, step, purpose)
      c = uint4{(unsigned)(i * 977), (unsigned)(i % 4), (unsigned)(i / 4), (unsigned)(i % 2)};
      k = uint2{(unsigned)(i % 3), 0u};
    } else {
      c = uint4{next(), next(), next(), next()};
      k = uint2{next(), next()};
    }
    uint4 r = e.block(c, k);
    std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n", c.x, c.y, c.z,
                c.w, k.x, k.y, r.x, r.y, r.z, r.w, i == 63 ? "" : ",");
  }
  std::printf("]\n");
  return 0;
}
