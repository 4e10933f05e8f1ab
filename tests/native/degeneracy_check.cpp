// CPU check of the host degeneracy order (kclistpp_amd/csrc/host_order.cpp):
// reads little-endian uint32 pairs from argv[1], writes the ranks of nodes
// 0..max id (uint32) to argv[2].  Built and run by tests/test_host.py, which
// compares the ranks with the oracle's restatement of Graph.c:57-91 on
// graphs that take the threaded and the sequential adjacency build.
#include <cstdint>
#include <cstdio>
#include <vector>

namespace kcl {
void degeneracy_rank(uint32_t n, uint64_t m, const uint32_t *pairs, uint32_t *rank);
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 3;
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<uint32_t> p((size_t)bytes / 4);
  if (std::fread(p.data(), 4, p.size(), f) != p.size()) return 4;
  std::fclose(f);
  uint32_t n = 0;
  for (uint32_t x : p) n = x + 1 > n ? x + 1 : n;
  std::vector<uint32_t> rank(n);
  kcl::degeneracy_rank(n, p.size() / 2, p.data(), rank.data());
  FILE *o = std::fopen(argv[2], "wb");
  if (!o || std::fwrite(rank.data(), 4, rank.size(), o) != rank.size()) return 5;
  std::fclose(o);
  return 0;
}
